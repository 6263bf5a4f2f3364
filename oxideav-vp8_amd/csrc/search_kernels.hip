// Exhaustive +-R whole-pel block search on gfx950 (a project-defined judge
// and labeller beside the reference's 4-step ARNR descent).
//
// For every 16x16 block of the centre frame and every other frame of the
// group, the SADs of all (2R+1)^2 candidates seed + (dx, dy) are computed
// with the reference's block_sad semantics (clamped fetches at the visible
// frame edges, src/arnr.rs:149-153, 200-225; partial edge blocks cover
// bw x bh pixels) and the minimum is taken; ties go to the smallest
// (|mvx| + |mvy|, mvy, mvx).
//
// Work decomposition: a lane group per block.  The block's reference
// window (rows by+sy-R .. by+sy+R+15, dword-aligned columns) is staged into
// LDS once; lane (j, db) owns the candidates dx in quad j (4 consecutive
// positions) and K consecutive dy rows starting at db*K.  It slides down
// the window rows: every row's five aligned dwords feed the K row
// accumulators that need them through v_qsad_pk_u16_u8, which adds the
// SADs of one 4-byte source dword against the 4 byte-shifted positions of
// an 8-byte reference pair into four 16-bit accumulators (no byte-alignment
// shifts; a full 16x16 SAD is at most 65280, so 16 bits never overflow).
// The lane's 4K candidates are folded into one 64-bit key
// (sad, |mv|_1, mvy, mvx) and the group reduces the key.
#include "vp8g_device.h"
#include "vp8g_internal.h"

namespace vp8g {

template <int R>
struct SearchGeom {
  static constexpr int D = 2 * R + 1;               // dy (and dx) candidates
  static constexpr int J = R / 2 + 1;               // dx quads: covers shift0 + 2R for any shift0 <= 3
  static constexpr int K = R <= 16 ? 5 : 9;         // dy rows per lane
  static constexpr int DB = (D + K - 1) / K;        // dy blocks
  static constexpr int LPB = J * DB;                // working lanes per block
  static constexpr int G = LPB <= 2 ? 2 : (LPB <= 64 ? 64 : (LPB + 63) / 64 * 64);  // lane group
  static constexpr int WROWS = DB * K + 15;         // window rows staged
  static constexpr int PD = J + 4;                  // window pitch (dwords)
  static constexpr int WIN_DW = WROWS * PD;
  static constexpr int WG = G >= 256 ? G : (G <= 2 ? 64 : 256);  // threads per workgroup
  static constexpr int BPW = WG / G;                // blocks per workgroup
};

__device__ __forceinline__ uint32_t gather4_clamped(const uint8_t* row, int x0, int w) {
  uint32_t v = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) v |= (uint32_t)row[clampi(x0 + b, 0, w - 1)] << (8 * b);
  return v;
}

__device__ __forceinline__ uint64_t pack_key(uint32_t sad, int mvx, int mvy) {
  const uint32_t l1 = (uint32_t)(abs(mvx) + abs(mvy));
  return ((uint64_t)sad << 48) | ((uint64_t)l1 << 32) | ((uint64_t)(uint32_t)(mvy + 32768) << 16) |
         (uint64_t)(uint32_t)(mvx + 32768);
}

__device__ __forceinline__ uint64_t min_u64(uint64_t a, uint64_t b) { return a < b ? a : b; }

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  const int lo = __shfl_xor((int)(uint32_t)v, m, 64);
  const int hi = __shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Source access: registers (S) for groups of at most one wave, or an LDS
// copy (s_src, read with wave-uniform addresses) for multi-wave groups,
// where a 64-entry register array is not promoted out of scratch.
template <int R>
struct SrcRows {
  const uint32_t (&S)[16][4];
  const uint32_t* lds;
  __device__ __forceinline__ uint4 row(int r) const {
    if (SearchGeom<R>::G > 64) return *reinterpret_cast<const uint4*>(lds + 4 * r);
    return make_uint4(S[r][0], S[r][1], S[r][2], S[r][3]);
  }
};

template <int R, bool FULL>
__device__ __forceinline__ void accumulate(const uint32_t* win, int j, int db, const SrcRows<R>& src,
                                           int bh, int bw, uint64_t (&acc)[SearchGeom<R>::K]) {
  using SG = SearchGeom<R>;
  constexpr int K = SG::K;
  const int bwq = bw >> 2, rem = bw & 3;
  const uint32_t remmask = rem ? (0xFFFFFFFFu >> (32 - 8 * rem)) : 0u;
#pragma unroll
  for (int yy = 0; yy < K + 15; yy++) {
    const uint32_t* wr = win + (db * K + yy) * SG::PD + j;
    const uint32_t w0 = wr[0], w1 = wr[1], w2 = wr[2], w3 = wr[3], w4 = wr[4];
    const uint64_t p[4] = {(uint64_t)w0 | ((uint64_t)w1 << 32), (uint64_t)w1 | ((uint64_t)w2 << 32),
                           (uint64_t)w2 | ((uint64_t)w3 << 32), (uint64_t)w3 | ((uint64_t)w4 << 32)};
    const uint32_t wv[5] = {w0, w1, w2, w3, w4};
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int r = yy - k;
      if (r < 0 || r > 15) continue;
      if (!FULL && r >= bh) continue;
      const uint4 s4 = src.row(r);
      const uint32_t sv[4] = {s4.x, s4.y, s4.z, s4.w};
      if (FULL) {
#pragma unroll
        for (int q = 0; q < 4; q++) acc[k] = __builtin_amdgcn_qsad_pk_u16_u8(p[q], sv[q], acc[k]);
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          if (q < bwq) {
            acc[k] = __builtin_amdgcn_qsad_pk_u16_u8(p[q], sv[q], acc[k]);
          } else if (q == bwq && rem) {
            // Last, partial source dword: mask the reference bytes too.
            uint64_t add = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
              const uint32_t ref = funnel(wv[q], wv[q + 1], i) & remmask;
              add |= (uint64_t)sad4(sv[q], ref, 0) << (16 * i);
            }
            acc[k] += add;
          }
        }
      }
    }
  }
}

// Shared memory of one block-search workgroup.
template <int R>
struct SearchSmem {
  uint32_t win[SearchGeom<R>::BPW][SearchGeom<R>::WIN_DW];
  uint64_t red[SearchGeom<R>::G > 64 ? SearchGeom<R>::G / 64 : 1];
  alignas(16) uint32_t src[SearchGeom<R>::G > 64 ? SearchGeom<R>::BPW * 64 : 4];
};

// The exhaustive +-R search of one (frame, block) item `b` (frame-major
// over the non-centre frames) by the lane group of workgroup slot `slot`.
// Every thread of the workgroup calls it (it holds workgroup barriers);
// the returned key (sad, |mv|_1, mvy, mvx) is valid in every lane of the
// group.  `*tab` receives the item's table offset.
template <int R>
__device__ __forceinline__ uint64_t search_item(const SearchArgs& a, SearchSmem<R>& sm, int b, int slot, int l,
                                                size_t* tab_out) {
  using SG = SearchGeom<R>;
  constexpr int G = SG::G, K = SG::K, J = SG::J;
  const int tid = threadIdx.x;
  const int per_frame = a.brows * a.bcols;
  const int kf = b / per_frame;
  const int fi = kf < a.center ? kf : kf + 1;
  const int rest = b - kf * per_frame;
  const int brow = rest / a.bcols, bcol = rest - brow * a.bcols;
  const int bx = bcol * 16, by = brow * 16;
  const int bw = min(16, a.w - bx), bh = min(16, a.h - by);
  int sx = 0, sy = 0;
  const size_t tab = 3 * ((size_t)fi * per_frame + rest);
  *tab_out = tab;
  if (a.seeds) {
    sx = clampi(a.seeds[tab], -kMaxSeed, kMaxSeed);
    sy = clampi(a.seeds[tab + 1], -kMaxSeed, kMaxSeed);
  }
  const FrameDesc F = a.frames[fi];
  const FrameDesc C = a.frames[a.center];
  const int x0 = bx + sx - R;                 // leftmost candidate column
  const int x0a = x0 >= 0 ? (x0 & ~3) : -(((-x0) + 3) & ~3);
  const int shift0 = x0 - x0a;                // 0..3
  const int y0 = by + sy - R;

  // Stage the clamped window.
  uint32_t* win = sm.win[slot];
  const bool interior = a.aligned4 && x0a >= 0 && x0a + 4 * SG::PD <= a.w;
  for (int s = l; s < SG::WIN_DW; s += G) {
    const int wy = s / SG::PD, wd = s - wy * SG::PD;
    const uint8_t* row = F.y + (size_t)clampi(y0 + wy, 0, a.h - 1) * F.y_stride;
    win[s] = interior ? *reinterpret_cast<const uint32_t*>(row + x0a + 4 * wd)
                      : gather4_clamped(row, x0a + 4 * wd, a.w);
  }
  // Source block (zero outside bw x bh).
  auto src_dword = [&](int r, int q) -> uint32_t {
    uint32_t v = 0;
    if (r < bh) {
      const uint8_t* row = C.y + (size_t)(by + r) * C.y_stride + bx;
      if (bw == 16 && a.aligned4) {
        v = *reinterpret_cast<const uint32_t*>(row + 4 * q);
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++)
          if (4 * q + e < bw) v |= (uint32_t)row[4 * q + e] << (8 * e);
      }
    }
    return v;
  };
  uint32_t S[16][4];
  if (G > 64) {
    if (l < 64) sm.src[slot * 64 + l] = src_dword(l >> 2, l & 3);
  } else {
#pragma unroll
    for (int r = 0; r < 16; r++) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t v = src_dword(r, q);
        // One block per wave: the source is wave-uniform and lives in SGPRs.
        S[r][q] = G == 64 ? (uint32_t)__builtin_amdgcn_readfirstlane((int)v) : v;
      }
    }
  }
  __syncthreads();
  const SrcRows<R> src{S, sm.src + (G > 64 ? slot * 64 : 0)};

  const int j = l % J, db = l / J;
  uint64_t best = ~0ull;
  if (l < SG::LPB) {
    uint64_t acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0;
    if (bw == 16 && bh == 16) {
      accumulate<R, true>(win, j, db, src, bh, bw, acc);
    } else {
      accumulate<R, false>(win, j, db, src, bh, bw, acc);
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int dyi = db * K + k;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int dxi = 4 * j + i - shift0;
        if (dyi < SG::D && dxi >= 0 && dxi < SG::D) {
          const uint32_t sad = (uint32_t)(acc[k] >> (16 * i)) & 0xFFFFu;
          best = min_u64(best, pack_key(sad, sx - R + dxi, sy - R + dyi));
        }
      }
    }
  }
  // Reduce over the block's lane group.
#pragma unroll
  for (int m = 1; m < (G < 64 ? G : 64); m <<= 1) best = min_u64(best, shfl_xor_u64(best, m));
  if (G > 64) {
    if ((tid & 63) == 0) sm.red[tid >> 6] = best;
    __syncthreads();
    best = sm.red[0];
#pragma unroll
    for (int wv = 1; wv < (G > 64 ? G / 64 : 1); wv++) best = min_u64(best, sm.red[wv]);
  }
  return best;
}

__device__ __forceinline__ void store_key(int32_t* o, uint64_t key) {
  o[0] = (int32_t)(key & 0xFFFF) - 32768;
  o[1] = (int32_t)((key >> 16) & 0xFFFF) - 32768;
  o[2] = (int32_t)(key >> 48);
}

template <int R>
__global__ void __launch_bounds__(SearchGeom<R>::WG) block_search_kernel(const SearchArgs a) {
  using SG = SearchGeom<R>;
  __shared__ SearchSmem<R> sm;
  const int tid = threadIdx.x;
  const int slot = tid / SG::G;         // block slot in the workgroup
  const int l = tid - slot * SG::G;     // lane within the block's group
  const long long bi = (long long)blockIdx.x * SG::BPW + slot;
  const bool live = bi < a.total_blocks;
  // Out-of-range slots still run the staging/barriers with a valid block.
  size_t tab;
  const uint64_t best = search_item<R>(a, sm, live ? (int)bi : 0, slot, l, &tab);
  if (live && l == 0) store_key(a.out + tab, best);
}

// ---------------------------------------------------------------- fallback
// Exhaustive fallback of a seeded search.  fallback_mark_kernel copies the
// first-stage table to the output and queues every (frame, block) item whose
// first-stage SAD fails the occlusion cut-off (sad >= 13 * bw * bh, the
// reference's sad / (bw * bh) > 12, src/arnr.rs:603-608) and, with the
// neighbour gate, has a 4-neighbour in the same frame that passes it.  One
// atomic per wave appends the wave's queued items.  fallback_search_kernel
// then runs the exhaustive +-R search around (0, 0) of every queued item
// (persistent workgroups walking the queue) and keeps the better of that and
// the first-stage entry under (sad, |mvx| + |mvy|, mvy, mvx).
__device__ __forceinline__ bool entry_fails(const int32_t* e, int area) {
  return (uint32_t)e[2] >= (kSadPerPixelCutoff + 1) * (uint32_t)area;
}

__global__ void __launch_bounds__(256) fallback_mark_kernel(const FallbackArgs f) {
  const int item = blockIdx.x * 256 + threadIdx.x;
  const int per_frame = f.brows * f.bcols;
  bool queue = false;
  if (item < f.total_blocks) {
    const int k = item / per_frame, rest = item - k * per_frame;
    const int fi = k < f.center ? k : k + 1;
    const int brow = rest / f.bcols, bcol = rest - brow * f.bcols;
    const size_t base = (size_t)fi * per_frame;
    const int32_t* e = f.first + 3 * (base + rest);
    int32_t* o = f.out + 3 * (base + rest);
    o[0] = e[0];
    o[1] = e[1];
    o[2] = e[2];
    auto area = [&](int r, int c) { return min(16, f.w - 16 * c) * min(16, f.h - 16 * r); };
    queue = entry_fails(e, area(brow, bcol));
    if (queue && f.gate) {
      bool any_pass = false;
      const int nr[4] = {brow, brow, brow - 1, brow + 1}, nc[4] = {bcol - 1, bcol + 1, bcol, bcol};
#pragma unroll
      for (int t = 0; t < 4; t++) {
        if (nr[t] < 0 || nr[t] >= f.brows || nc[t] < 0 || nc[t] >= f.bcols) continue;
        if (!entry_fails(f.first + 3 * (base + (size_t)nr[t] * f.bcols + nc[t]), area(nr[t], nc[t])))
          any_pass = true;
      }
      queue = any_pass;
    }
  }
  const uint64_t mask = __ballot(queue);
  if (mask == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(f.count, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, leader, 64);
  if (queue) f.queue[base + __popcll(mask & ((1ull << lane) - 1))] = (uint32_t)item;
}

template <int R>
__global__ void __launch_bounds__(SearchGeom<R>::WG) fallback_search_kernel(const SearchArgs a, const FallbackArgs f) {
  using SG = SearchGeom<R>;
  __shared__ SearchSmem<R> sm;
  const int tid = threadIdx.x;
  const int slot = tid / SG::G;
  const int l = tid - slot * SG::G;
  const uint32_t cnt = uniform_u32(f.count);
  if (f.queued && blockIdx.x == 0 && tid == 0) *f.queued = cnt;
  for (uint32_t base = blockIdx.x * SG::BPW; base < cnt; base += gridDim.x * SG::BPW) {
    const uint32_t qi = base + slot;
    const bool live = qi < cnt;
    // A dead slot repeats the workgroup's first item (base < cnt).
    const int item = (int)f.queue[live ? qi : base];
    size_t tab;
    const uint64_t best = search_item<R>(a, sm, item, slot, l, &tab);
    if (live && l == 0) {
      const int32_t* e = f.first + tab;
      const uint64_t prev = pack_key((uint32_t)e[2], e[0], e[1]);
      store_key(f.out + tab, min_u64(best, prev));
    }
    __syncthreads();  // the next item's staging overwrites the windows
  }
}

hipError_t launch_search_fallback(const SearchArgs& a, const FallbackArgs& f, int radius, int cus,
                                  hipStream_t stream) {
  if (a.total_blocks <= 0) return f.queued ? launch_zero_u32(f.queued, 1, stream) : hipSuccess;
  hipError_t e = launch_zero_u32(f.count, 1, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fallback_mark_kernel, dim3((a.total_blocks + 255) / 256), dim3(256), 0, stream, f);
  switch (radius) {
#define VP8G_FALLBACK_CASE(RR)                                                                     \
  case RR: {                                                                                       \
    using SG = SearchGeom<RR>;                                                                     \
    const int need = (a.total_blocks + SG::BPW - 1) / SG::BPW;                                     \
    const int grid = min(need, max(1, cus) * (SG::WG >= 512 ? 2 : 8));                             \
    hipLaunchKernelGGL(fallback_search_kernel<RR>, dim3(grid), dim3(SG::WG), 0, stream, a, f);     \
    break;                                                                                         \
  }
    VP8G_FALLBACK_CASE(16)
    VP8G_FALLBACK_CASE(64)
#undef VP8G_FALLBACK_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}


// ---------------------------------------------------------------- pyramid
// One level of a resolution pyramid for coarse-to-fine search: each output
// pixel is the rounded mean of a 2x2 input block, input coordinates clamped
// to the plane (odd sizes repeat the last row or column).  blockIdx.z =
// 3 * frame + plane (Y, U, V); one thread per 4 consecutive output pixels of
// a row.
__global__ void __launch_bounds__(256) downsample2_kernel(const Downsample2Args a) {
  const int fi = blockIdx.z / 3, p = blockIdx.z - 3 * fi;
  const Downsample2Frame& F = a.f[fi];
  const uint8_t* src = p == 0 ? F.src_y : (p == 1 ? F.src_u : F.src_v);
  uint8_t* dst = p == 0 ? F.dst_y : (p == 1 ? F.dst_u : F.dst_v);
  const int w = p == 0 ? a.w : (a.w + 1) >> 1, h = p == 0 ? a.h : (a.h + 1) >> 1;
  const uint32_t ss = p == 0 ? F.src_y_stride : F.src_uv_stride;
  const uint32_t ds = p == 0 ? F.dst_y_stride : F.dst_uv_stride;
  const int ow = (w + 1) >> 1, oh = (h + 1) >> 1;
  const int x4 = 4 * (blockIdx.x * 64 + (threadIdx.x & 63)), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x4 >= ow || y >= oh) return;
  const uint8_t* r0 = src + (size_t)(2 * y) * ss;
  const uint8_t* r1 = src + (size_t)min(2 * y + 1, h - 1) * ss;
  uint8_t* o = dst + (size_t)y * ds;
  // Interior: the 8 source bytes of each row are one aligned uint2 (8-byte
  // aligned planes and strides), the 4 outputs one aligned dword.
  if (((uintptr_t)r0 & 7) == 0 && ((uintptr_t)r1 & 7) == 0 && ((uintptr_t)(o + x4) & 3) == 0 && 2 * x4 + 8 <= w &&
      x4 + 4 <= ow) {
    const uint2 a = *reinterpret_cast<const uint2*>(r0 + 2 * x4), b = *reinterpret_cast<const uint2*>(r1 + 2 * x4);
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t wa = k < 2 ? a.x : a.y, wb = k < 2 ? b.x : b.y;
      const int sh = 16 * (k & 1);
      const uint32_t sum = ((wa >> sh) & 255) + ((wa >> (sh + 8)) & 255) + ((wb >> sh) & 255) + ((wb >> (sh + 8)) & 255);
      out |= ((sum + 2) >> 2) << (8 * k);
    }
    *reinterpret_cast<uint32_t*>(o + x4) = out;
    return;
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int x = x4 + k;
    if (x < ow) {
      const int x0 = 2 * x, x1 = min(2 * x + 1, w - 1);
      o[x] = (uint8_t)((r0[x0] + r0[x1] + r1[x0] + r1[x1] + 2) >> 2);
    }
  }
}

// Seeds for the next finer level: fine block (r, c) of frame f takes the
// vector of coarse block (min(r / 2, cbh - 1), min(c / 2, cbw - 1)), both
// components doubled; the third entry is 0.
__global__ void __launch_bounds__(256) upsample_table_kernel(const int32_t* coarse, int cbw, int cbh, int32_t* fine,
                                                             int fbw, int fbh, int total) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const int per = fbw * fbh;
  const int f = i / per, rem = i - f * per;
  const int r = rem / fbw, c = rem - r * fbw;
  const int32_t* src = coarse + 3 * (((size_t)f * cbh + min(r >> 1, cbh - 1)) * cbw + min(c >> 1, cbw - 1));
  int32_t* dst = fine + 3 * (size_t)i;
  dst[0] = 2 * src[0];
  dst[1] = 2 * src[1];
  dst[2] = 0;
}

hipError_t launch_downsample2(const Downsample2Args& a, int frames, hipStream_t stream) {
  if (frames <= 0 || frames > kDownsampleFrames) return hipErrorInvalidValue;
  const int ow = (a.w + 1) >> 1, oh = (a.h + 1) >> 1;
  hipLaunchKernelGGL(downsample2_kernel, dim3((ow + 255) / 256, (oh + 3) / 4, 3 * frames), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_upsample_table(const int32_t* coarse, int cbw, int cbh, int32_t* fine, int fbw, int fbh, int n,
                                 hipStream_t stream) {
  if (n <= 0 || fbw <= 0 || fbh <= 0 || cbw <= 0 || cbh <= 0) return hipErrorInvalidValue;
  const int64_t total = (int64_t)n * fbw * fbh;
  if (total > INT32_MAX - 255) return hipErrorInvalidValue;
  hipLaunchKernelGGL(upsample_table_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, coarse, cbw,
                     cbh, fine, fbw, fbh, (int)total);
  return hipGetLastError();
}

// ---------------------------------------------------------------- candidates
// Exact +-2 search of one block against one neighbour frame around centre
// (sx, sy), for refine_candidates_kernel: the window is loaded and scored
// at once (no prefetch).  Lane layout, SADs and tie keys as in
// refine2_kernel.  Returns {mvx, mvy, sad} of the best candidate.
struct R2Best {
  int mvx, mvy;
  uint32_t sad;
};

__device__ __forceinline__ R2Best r2_search_at(const SearchArgs& a, const FrameDesc& F, uint32_t* win,
                                               const uint4 (&src)[4], int bx, int by, int bw, int bh, bool full,
                                               int sx, int sy, int l) {
  constexpr int R = 2, PD = 6, WROWS = 20, WDW = WROWS * PD;
  const int dy = l >> 3, j = (l >> 2) & 1, h = l & 3;
  auto floor4 = [](int x) { return x >= 0 ? (x & ~3) : -(((-x) + 3) & ~3); };
  const int wy0 = l / PD, wd0 = l - PD * wy0;
  const int s1 = min(l + 64, WDW - 1), wy1 = s1 / PD, wd1 = s1 - PD * wy1;
  const int x0a = floor4(bx + sx - R), y0 = by + sy - R;
  uint32_t v0, v1;
  if (a.aligned4 && x0a >= 0 && x0a + 4 * PD <= a.w && y0 >= 0 && y0 + WROWS <= a.h) {
    const uint8_t* base = F.y + (size_t)y0 * F.y_stride + x0a;
    v0 = *reinterpret_cast<const uint32_t*>(base + (uint32_t)wy0 * F.y_stride + 4 * wd0);
    v1 = *reinterpret_cast<const uint32_t*>(base + (uint32_t)wy1 * F.y_stride + 4 * wd1);
  } else {
    v0 = gather4_clamped(F.y + (size_t)clampi(y0 + wy0, 0, a.h - 1) * F.y_stride, x0a + 4 * wd0, a.w);
    v1 = gather4_clamped(F.y + (size_t)clampi(y0 + wy1, 0, a.h - 1) * F.y_stride, x0a + 4 * wd1, a.w);
  }
  wave_lds_sync();  // the previous search's window reads are done
  win[l] = v0;
  if (l + 64 < WDW) win[l + 64] = v1;
  wave_lds_sync();
  const int shift0 = (bx + sx - R) - floor4(bx + sx - R);
  uint64_t acc = 0;
  if (full) {
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const uint32_t* wr = win + (dy + 4 * h + rr) * PD + j;
      const uint32_t w0 = wr[0], w1 = wr[1], w2 = wr[2], w3 = wr[3], w4 = wr[4];
      acc = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w0 | ((uint64_t)w1 << 32), src[rr].x, acc);
      acc = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w1 | ((uint64_t)w2 << 32), src[rr].y, acc);
      acc = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w2 | ((uint64_t)w3 << 32), src[rr].z, acc);
      acc = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w3 | ((uint64_t)w4 << 32), src[rr].w, acc);
    }
  } else {
    const int bwq = bw >> 2, rem = bw & 3;
    const uint32_t remmask = rem ? (0xFFFFFFFFu >> (32 - 8 * rem)) : 0u;
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      if (4 * h + rr < bh) {
        const uint32_t* wr = win + (dy + 4 * h + rr) * PD + j;
        const uint32_t wv[5] = {wr[0], wr[1], wr[2], wr[3], wr[4]};
        const uint32_t sv[4] = {src[rr].x, src[rr].y, src[rr].z, src[rr].w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint64_t pair = (uint64_t)wv[q] | ((uint64_t)wv[q + 1] << 32);
          if (q < bwq) {
            acc = __builtin_amdgcn_qsad_pk_u16_u8(pair, sv[q], acc);
          } else if (q == bwq && rem) {
            uint64_t add = 0;
#pragma unroll
            for (int i = 0; i < 4; i++)
              add |= (uint64_t)sad4(sv[q], funnel(wv[q], wv[q + 1], i) & remmask, 0) << (16 * i);
            acc += add;
          }
        }
      }
    }
  }
  uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
  lo += dpp<kDppQuadXor1>(lo);
  hi += dpp<kDppQuadXor1>(hi);
  lo += dpp<kDppQuadXor2>(lo);
  hi += dpp<kDppQuadXor2>(hi);
  const int mvy = sy - R + dy, mvx0 = sx - R + 4 * j - shift0;
  const int l1_base = abs(sx) + abs(sy) - 2 * R;
  const uint32_t key_lo = (uint32_t)((dy << 3) + 4 * j - shift0);
  uint32_t lane_key = 0xFFFFFFFFu;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int dxi = 4 * j + i - shift0;
    const uint32_t sad = ((i < 2 ? lo : hi) >> (16 * (i & 1))) & 0xFFFFu;
    const uint32_t rel = (uint32_t)(abs(mvx0 + i) + abs(mvy) - l1_base);
    if (l < 40 && dxi >= 0 && dxi <= 2 * R) lane_key = min(lane_key, (sad << 10) | (rel << 6) | (key_lo + (uint32_t)i));
  }
  const uint32_t best = wave_min_any(lane_key);
  return R2Best{sx - R + (int)(best & 7), sy - R + (int)((best >> 3) & 7), best >> 10};
}

// Exact +-2 search around each block's seed: one wave per (neighbour frame,
// block position), 4 per workgroup, frame-major.  The centre block is staged
// through LDS into the lane's source registers, then r2_search_at loads the
// frame's 20x24 window and scores the 25 vectors.
__global__ void __launch_bounds__(256) refine2_kernel(const SearchArgs a) {
  __shared__ __align__(16) uint32_t s_win[4][120];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int per_frame = a.brows * a.bcols;
  const int item = xcd_band(blockIdx.x, gridDim.x) * 4 + wave;
  if (item >= a.total_blocks) return;  // whole wave; no workgroup barrier follows
  const int k = item / per_frame, rest = item - k * per_frame;
  const int fi = k < a.center ? k : k + 1;
  const int brow = rest / a.bcols, bcol = rest - brow * a.bcols;
  const int bx = bcol * 16, by = brow * 16;
  const int bw = min(16, a.w - bx), bh = min(16, a.h - by);
  const bool full = bw == 16 && bh == 16;
  uint32_t* win = s_win[wave];
  int sx = 0, sy = 0;
  if (a.seeds) {
    const int32_t* sd = a.seeds + 3 * ((size_t)fi * per_frame + rest);
    sx = clampi((int)uniform_u32(sd), -kMaxSeed, kMaxSeed);
    sy = clampi((int)uniform_u32(sd + 1), -kMaxSeed, kMaxSeed);
  }
  uint4 src[4];
  {
    const FrameDesc C = a.frames[a.center];
    const int r = l >> 2, q = l & 3;
    uint32_t v = 0;
    if (r < bh) {
      const uint8_t* row = C.y + (size_t)(by + r) * C.y_stride + bx;
      if (full && a.aligned4) {
        v = *reinterpret_cast<const uint32_t*>(row + 4 * q);
      } else {
        for (int e = 0; e < 4; e++)
          if (4 * q + e < bw) v |= (uint32_t)row[4 * q + e] << (8 * e);
      }
    }
    win[l] = v;
    wave_lds_sync();
#pragma unroll
    for (int rr = 0; rr < 4; rr++) src[rr] = reinterpret_cast<const uint4*>(win)[4 * (l & 3) + rr];
  }
  const R2Best r = r2_search_at(a, a.frames[fi], win, src, bx, by, bw, bh, full, sx, sy, l);
  if (l == 0) {
    int32_t* o = a.out + 3 * ((size_t)fi * per_frame + rest);
    o[0] = r.mvx;
    o[1] = r.mvy;
    o[2] = (int32_t)r.sad;
  }
}

// (sad, |mvx| + |mvy|, mvy, mvx) as one ordered 64-bit key (components of at
// most 16386 in magnitude).
__device__ __forceinline__ uint64_t abs_key(int mvx, int mvy, uint32_t sad) {
  return ((uint64_t)sad << 48) | ((uint64_t)(abs(mvx) + abs(mvy)) << 32) | ((uint64_t)(mvy + 32768) << 16) |
         (uint64_t)(mvx + 32768);
}

// Second stage of the learned-seeded search: for every block and neighbour
// frame, exact +-2 searches around extra centres, and the best of those and
// the first stage's result.  The centres are the block's own first-stage
// vector when it lies on the border of the first square around the seed
// (the minimum may continue past it), then the first-stage vectors of the
// left, right, upper and lower blocks, each taken once and only when it
// differs from the seed and the own vector.  One wave per (neighbour frame,
// block position), 4 per workgroup, frame-major: the searches of one block
// are a chain of dependent LDS round trips, so the frames run in parallel
// waves rather than one after another.
__global__ void __launch_bounds__(256) refine_candidates_kernel(const SearchArgs a, const int32_t* first) {
  constexpr int kMaxCand = 5;
  __shared__ __align__(16) uint32_t s_win[4][120];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int per_frame = a.brows * a.bcols;
  const int item = xcd_band(blockIdx.x, gridDim.x) * 4 + wave;
  if (item >= a.total_blocks) return;  // whole wave; no workgroup barrier follows
  const int k = item / per_frame, rest = item - k * per_frame;
  const int fi = k < a.center ? k : k + 1;
  const int brow = rest / a.bcols, bcol = rest - brow * a.bcols;
  const int bx = bcol * 16, by = brow * 16;
  const int bw = min(16, a.w - bx), bh = min(16, a.h - by);
  const bool full = bw == 16 && bh == 16;
  uint32_t* win = s_win[wave];
  // The frame's seed, own first-stage entry and neighbours' vectors (an
  // absent neighbour repeats the own vector), requested before the centre
  // block so that the loads overlap.
  const int32_t* own = first + 3 * ((size_t)fi * per_frame + rest);
  const int ox = (int)uniform_u32(own), oy = (int)uniform_u32(own + 1);
  const uint32_t osad = uniform_u32(own + 2);
  int sx = 0, sy = 0;
  if (a.seeds) {
    const int32_t* sd = a.seeds + 3 * ((size_t)fi * per_frame + rest);
    sx = clampi((int)uniform_u32(sd), -kMaxSeed, kMaxSeed);
    sy = clampi((int)uniform_u32(sd + 1), -kMaxSeed, kMaxSeed);
  }
  int cx[kMaxCand], cy[kMaxCand];
  {
    const int nr[4] = {brow, brow, brow - 1, brow + 1}, ncl[4] = {bcol - 1, bcol + 1, bcol, bcol};
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const bool in = nr[t] >= 0 && nr[t] < a.brows && ncl[t] >= 0 && ncl[t] < a.bcols;
      const int32_t* e = first + 3 * ((size_t)fi * per_frame + (size_t)(in ? nr[t] : brow) * a.bcols +
                                      (in ? ncl[t] : bcol));
      cx[t + 1] = (int)uniform_u32(e);
      cy[t + 1] = (int)uniform_u32(e + 1);
    }
  }
  uint4 src[4];
  {
    const FrameDesc C = a.frames[a.center];
    const int r = l >> 2, q = l & 3;
    uint32_t v = 0;
    if (r < bh) {
      const uint8_t* row = C.y + (size_t)(by + r) * C.y_stride + bx;
      if (full && a.aligned4) {
        v = *reinterpret_cast<const uint32_t*>(row + 4 * q);
      } else {
        for (int e = 0; e < 4; e++)
          if (4 * q + e < bw) v |= (uint32_t)row[4 * q + e] << (8 * e);
      }
    }
    win[l] = v;
    wave_lds_sync();
#pragma unroll
    for (int rr = 0; rr < 4; rr++) src[rr] = reinterpret_cast<const uint4*>(win)[4 * (l & 3) + rr];
  }
  const FrameDesc F = a.frames[fi];
  // slot 0: the own vector, searched again only when it lies on the first
  // square's border (the minimum may continue past it); slots 1..4: left,
  // right, upper, lower
  bool use[kMaxCand];
  cx[0] = ox;
  cy[0] = oy;
  use[0] = abs(ox - sx) == 2 || abs(oy - sy) == 2;
#pragma unroll
  for (int t = 1; t < kMaxCand; t++) use[t] = true;
#pragma unroll
  for (int t = 0; t < kMaxCand; t++) {
    cx[t] = clampi(cx[t], -kMaxSeed, kMaxSeed);
    cy[t] = clampi(cy[t], -kMaxSeed, kMaxSeed);
    use[t] = use[t] && !(cx[t] == sx && cy[t] == sy);  // that square was searched
#pragma unroll
    for (int u = 0; u < t; u++)
      if (cx[u] == cx[t] && cy[u] == cy[t]) use[t] = false;  // the own vector or an earlier slot
  }
  int bxv = ox, byv = oy;
  uint32_t bsad = osad;
  uint64_t bkey = abs_key(ox, oy, osad);
#pragma unroll
  for (int t = 0; t < kMaxCand; t++) {
    if (!use[t]) continue;
    const R2Best r = r2_search_at(a, F, win, src, bx, by, bw, bh, full, cx[t], cy[t], l);
    const uint64_t key = abs_key(r.mvx, r.mvy, r.sad);
    if (key < bkey) {
      bkey = key;
      bxv = r.mvx;
      byv = r.mvy;
      bsad = r.sad;
    }
  }
  if (l == 0) {
    int32_t* o = a.out + 3 * ((size_t)fi * per_frame + rest);
    o[0] = bxv;
    o[1] = byv;
    o[2] = (int32_t)bsad;
  }
}

hipError_t launch_refine_candidates(const SearchArgs& a, const int32_t* first, hipStream_t stream) {
  if (a.total_blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(refine_candidates_kernel, dim3((a.total_blocks + 3) / 4), dim3(256), 0, stream, a, first);
  return hipGetLastError();
}

bool block_search_supported(int radius) { return radius == 2 || radius == 16 || radius == 64; }

hipError_t launch_block_search(const SearchArgs& a, int radius, hipStream_t stream) {
  if (a.total_blocks <= 0) return hipSuccess;
  switch (radius) {
#define VP8G_SEARCH_CASE(RR)                                                                       \
  case RR: {                                                                                       \
    using SG = SearchGeom<RR>;                                                                     \
    const int grid = (a.total_blocks + SG::BPW - 1) / SG::BPW;                                     \
    hipLaunchKernelGGL(block_search_kernel<RR>, dim3(grid), dim3(SG::WG), 0, stream, a);           \
    break;                                                                                         \
  }
    case 2:
      hipLaunchKernelGGL(refine2_kernel, dim3((a.total_blocks + 3) / 4), dim3(256), 0, stream, a);
      break;
    VP8G_SEARCH_CASE(16)
    VP8G_SEARCH_CASE(64)
#undef VP8G_SEARCH_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace vp8g
