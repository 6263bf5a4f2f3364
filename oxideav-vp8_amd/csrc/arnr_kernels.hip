// ARNR altref synthesis on gfx950, search + occlusion guard + blend + resolve
// fused into one launch.
//
// Work decomposition: one workgroup per strip of M horizontally adjacent
// anchor macroblocks of one MB row; one wave per neighbour frame.  Each wave
// stages its neighbour's clamped luma and chroma windows covering the whole
// strip (+-15 px luma, +-7 px chroma) into LDS once, then runs the M block
// searches and blends out of LDS.  The centre strip is staged once per
// workgroup.  Per-pixel sums from all waves meet in LDS accumulators and the
// workgroup resolves the strip.
//
// Semantics follow src/arnr.rs:476-675 of the reference crate:
//   * search: refine_search (src/arnr.rs:264-300) — SAD at (0,0), then for
//     step 8,4,2,1 the 8 probes around the step's frozen centre in the order
//     (-s,-s),(0,-s),(s,-s),(-s,0),(s,0),(-s,s),(0,s),(s,s); a probe wins on
//     strictly smaller SAD, so the first minimum in that order wins.  The
//     reference's row early exit (src/arnr.rs:217-219) never changes the
//     winner, so full SADs are computed.
//   * edge clamping uses the visible dims (pel, src/arnr.rs:149-153): the
//     staged windows hold the clamped samples, so every later read is plain.
//   * occlusion guard: skip the block when sad / (bw*bh) > 12
//     (src/arnr.rs:603-608).
//   * blend: acc += wt[|ref - ctr|] * ref, wsum += wt[..] with the 256-entry
//     table of src/arnr.rs:307-323; chroma uses (mv/2 truncated toward zero)
//     and the chroma-plane clamp (src/arnr.rs:402-459, 620-639).  Both sums
//     are packed into one u32 (acc < 2^20, wsum < 2^12 for <= 64 frames)
//     and added with LDS atomics: integer addition is exact and commutative,
//     so the result does not depend on wave scheduling.
//   * resolve: (acc + ws/2) / ws with the centre's weight-16 seed
//     (src/arnr.rs:662-675).
// Chroma output clamping (src/arnr.rs:442-444) never fires from the block
// grid: cbx + ceil(bw/2) = ceil(min(bx+16, w)/2) <= ceil(w/2).
#include "vp8g_device.h"
#include "vp8g_internal.h"

namespace vp8g {


constexpr int kWinRows = 46;   // by-15 .. by+30
constexpr int kCWinRows = 22;  // cby-7 .. cby+14
constexpr int kAccShift = 20;  // packed accumulator: wsum << 20 | acc

// Strip geometry for M MBs per workgroup; compile-time so that every
// index division in the staging and resolve loops is a multiply-shift.
template <int M_>
struct StripGeom {
  static constexpr int m = M_;                 // MBs per strip
  static constexpr int lw = 16 * M_ + 32;      // luma window width (cols bx0-16 .. bx0+16M+15)
  // Row pitch = width + 8 bytes: two rows apart (the row pair of one search
  // lane) then land 4 banks apart per lane, so the 8 lanes of a probe group
  // read disjoint LDS banks.
  static constexpr int lp = lw + 8;
  static constexpr int cp = 8 * M_ + 16;       // chroma window pitch (cols cbx0-8 .. cbx0+8M+7)
  static constexpr int wave_bytes = (kWinRows * lp + 2 * kCWinRows * cp + 15) & ~15;
};

__device__ __forceinline__ uint32_t load_clamped4(const uint8_t* row, int x0, int w, bool aligned4) {
  if (aligned4 && x0 >= 0 && x0 + 4 <= w) return *reinterpret_cast<const uint32_t*>(row + x0);
  uint32_t v = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) v |= (uint32_t)row[clampi(x0 + j, 0, w - 1)] << (8 * j);
  return v;
}

__device__ __forceinline__ void lds_add(uint32_t* p, uint32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stage frame f's clamped windows for the strip starting at (bx0, by):
// luma rows by-15..by+30, cols bx0-16..bx0+16M+15; chroma rows
// cby-7..cby+14, cols cbx0-8..cbx0+8M+7 (U then V).
template <class G>
__device__ void stage_windows(uint8_t* win, const FrameDesc f, const ArnrGroup& a, const G g,
                              int bx0, int by, int lane) {
  const int cbx0 = bx0 >> 1, cby = by >> 1;
  const ptrdiff_t v_minus_u = f.v - f.u;  // plane choice as integer arithmetic
  uint8_t* cwin = win + kWinRows * g.lp;
  constexpr int lpieces = G::lw >> 4;  // 16-byte pieces per luma row
  constexpr int cdw = G::cp >> 2;      // dwords per chroma row
  const bool interior = a.aligned16 && bx0 >= 16 && bx0 + 16 * G::m + 16 <= a.w && by >= 15 &&
                        by + 31 <= a.h;
  if (interior) {
    // Luma: whole aligned 16-byte pieces; chroma (then interior as well:
    // cbx0 >= 8, cbx0 + 8M + 8 <= cw, cby >= 8, cby + 15 <= ch): dwords.
    // Loads go out in batches of 4 per lane before their LDS stores.
    const uint8_t* ybase = f.y + (size_t)(by - 15) * f.y_stride + bx0 - 16;
    constexpr int total = kWinRows * lpieces;
    auto yld = [&](int s) -> uint4 {
      s = min(s, total - 1);
      const int wy = s / lpieces, pc = s - wy * lpieces;
      return *reinterpret_cast<const uint4*>(ybase + (size_t)wy * f.y_stride + 16 * pc);
    };
    auto yst = [&](int s, const uint4& v) {
      if (s < total) {
        const int wy = s / lpieces, pc = s - wy * lpieces;
        uint2* d = reinterpret_cast<uint2*>(win + wy * G::lp + 16 * pc);  // 8-byte aligned
        d[0] = make_uint2(v.x, v.y);
        d[1] = make_uint2(v.z, v.w);
      }
    };
    for (int s0 = lane; s0 < total; s0 += 256) {
      const uint4 v0 = yld(s0), v1 = yld(s0 + 64), v2 = yld(s0 + 128), v3 = yld(s0 + 192);
      yst(s0, v0);
      yst(s0 + 64, v1);
      yst(s0 + 128, v2);
      yst(s0 + 192, v3);
    }
    const size_t coff = (size_t)(cby - 7) * f.uv_stride + cbx0 - 8;
    constexpr int ctotal = 2 * kCWinRows * cdw;
    const bool al4 = a.aligned4c != 0;
    auto cld = [&](int s) -> uint32_t {
      s = min(s, ctotal - 1);
      const int pl = s >= kCWinRows * cdw;
      const int t = s - pl * kCWinRows * cdw;
      const int wy = t / cdw, dq = t - wy * cdw;
      const uint8_t* p = f.u + (pl ? v_minus_u : 0) + coff + (size_t)wy * f.uv_stride + 4 * dq;
      return al4 ? *reinterpret_cast<const uint32_t*>(p)
                 : (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    };
    uint32_t* cw32 = reinterpret_cast<uint32_t*>(cwin);
    for (int s0 = lane; s0 < ctotal; s0 += 256) {
      const uint32_t c0 = cld(s0), c1 = cld(s0 + 64), c2 = cld(s0 + 128), c3 = cld(s0 + 192);
      cw32[s0] = c0;
      if (s0 + 64 < ctotal) cw32[s0 + 64] = c1;
      if (s0 + 128 < ctotal) cw32[s0 + 128] = c2;
      if (s0 + 192 < ctotal) cw32[s0 + 192] = c3;
    }
    return;
  }
  // Frame-edge strips: clamped dword gathers.
  constexpr int ldw = G::lw >> 2;
  for (int s = lane; s < kWinRows * ldw; s += 64) {
    const int wy = s / ldw, dq = s - wy * ldw;
    const int y = clampi(by - 15 + wy, 0, a.h - 1);
    reinterpret_cast<uint32_t*>(win)[wy * (G::lp >> 2) + dq] =
        load_clamped4(f.y + (size_t)y * f.y_stride, bx0 - 16 + 4 * dq, a.w, a.aligned4 != 0);
  }
  for (int s = lane; s < 2 * kCWinRows * cdw; s += 64) {
    const int pl = s >= kCWinRows * cdw;
    const int t = s - pl * kCWinRows * cdw;
    const int wy = t / cdw, dq = t - wy * cdw;
    const int y = clampi(cby - 7 + wy, 0, a.ch - 1);
    reinterpret_cast<uint32_t*>(cwin)[s] = load_clamped4(f.u + (pl ? v_minus_u : 0) + (size_t)y * f.uv_stride,
                                                         cbx0 - 8 + 4 * dq, a.cw, a.aligned4c != 0);
  }
}

// Probe g of a step sits at centre + s*(dx_g, dy_g) with the reference's
// order (-1,-1),(0,-1),(1,-1),(-1,0),(1,0),(-1,1),(0,1),(1,1)
// (src/arnr.rs:274-283); (d + 1) is packed 2 bits per probe.
constexpr uint32_t kProbeDx1 = (0u << 0) | (1u << 2) | (2u << 4) | (0u << 6) | (2u << 8) | (0u << 10) |
                               (1u << 12) | (2u << 14);
constexpr uint32_t kProbeDy1 = (0u << 0) | (0u << 2) | (0u << 4) | (1u << 6) | (1u << 8) | (2u << 10) |
                               (2u << 12) | (2u << 14);

// refine_search (src/arnr.rs:264-300) for one block against the staged
// window.  Lanes are 8 probe groups x 8 row pairs; `row0` points at this
// lane's first source row's (0,0)-displaced reference row.  The probes of
// a step never leave +-15 (8+4+2+1), so the reference's range check never
// fires and is not evaluated.  FULL: 16x16 block; otherwise columns >= bw
// and rows >= bh are masked out of every SAD (the centre strip holds zeros
// there, the reference row is masked to zero).
template <bool FULL, class G>
__device__ __forceinline__ void descend(const uint8_t* row0, const uint4 src0, const uint4 src1,
                                        int gd_off, int gq, int kk, int bw, int bh, int& bestx,
                                        int& besty, uint32_t& best_sad, const uint8_t* zrow, uint32_t zsrc) {
  static_assert((G::lp & 3) == 0, "window rows are dword aligned");
  uint32_t colmask[4] = {~0u, ~0u, ~0u, ~0u};
  uint32_t rowmask0 = ~0u, rowmask1 = ~0u;
  if (!FULL) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int left = bw - 4 * q;
      colmask[q] = left >= 4 ? 0xFFFFFFFFu : (left <= 0 ? 0u : (0xFFFFFFFFu >> (32 - 8 * left)));
    }
    rowmask0 = 2 * kk < bh ? ~0u : 0u;
    rowmask1 = 2 * kk + 1 < bh ? ~0u : 0u;
  }
  // The lane's two rows at byte offset `off`: one alignment for both (the
  // row pitch is a dword multiple), ten naturally aligned dword reads and
  // byte funnel shifts (unaligned LDS reads stall the LDS pipe).
  auto sad_at = [&](int off) -> uint32_t {
    const uint8_t* p = row0 + off;
    const uint32_t sh = (uint32_t)(uintptr_t)p & 3;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p - sh);
    constexpr int P = G::lp >> 2;
    const uint32_t a0 = w[0], a1 = w[1], a2 = w[2], a3 = w[3], a4 = w[4];
    const uint32_t b0 = w[P], b1 = w[P + 1], b2 = w[P + 2], b3 = w[P + 3], b4 = w[P + 4];
    uint4 q0 = make_uint4(funnel(a0, a1, sh), funnel(a1, a2, sh), funnel(a2, a3, sh), funnel(a3, a4, sh));
    uint4 q1 = make_uint4(funnel(b0, b1, sh), funnel(b1, b2, sh), funnel(b2, b3, sh), funnel(b3, b4, sh));
    if (!FULL) {
      q0.x &= colmask[0] & rowmask0; q0.y &= colmask[1] & rowmask0;
      q0.z &= colmask[2] & rowmask0; q0.w &= colmask[3] & rowmask0;
      q1.x &= colmask[0] & rowmask1; q1.y &= colmask[1] & rowmask1;
      q1.z &= colmask[2] & rowmask1; q1.w &= colmask[3] & rowmask1;
    }
    uint32_t acc = sad4(src0.x, q0.x, 0);
    acc = sad4(src0.y, q0.y, acc);
    acc = sad4(src0.z, q0.z, acc);
    acc = sad4(src0.w, q0.w, acc);
    acc = sad4(src1.x, q1.x, acc);
    acc = sad4(src1.y, q1.y, acc);
    acc = sad4(src1.z, q1.z, acc);
    acc = sad4(src1.w, q1.w, acc);
    return sum8(acc);
  };
  int bx = 0, by = 0, boff = 0;
  uint32_t bsad;
  if (FULL) {
    // SAD at (0, 0) over the whole wave: lane = (row l / 4, dword l % 4),
    // both operands dword aligned.
    uint32_t z = sad4(zsrc, *reinterpret_cast<const uint32_t*>(zrow), 0);
    z = sum16(z);
    bsad = (uint32_t)__builtin_amdgcn_readlane((int)z, 0) + (uint32_t)__builtin_amdgcn_readlane((int)z, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)z, 32) + (uint32_t)__builtin_amdgcn_readlane((int)z, 48);
  } else {
    bsad = (uint32_t)__builtin_amdgcn_readfirstlane((int)sad_at(0));
  }
#pragma unroll
  for (int sh = 3; sh >= 0; sh--) {
    const uint32_t s = sad_at(boff + gd_off * (1 << sh));
    // Probe order is the group index and the incumbent keeps ties, so the
    // key (sad << 4 | group + 1) against (best << 4) yields exactly the
    // reference's winner as one unsigned minimum over the wave.
    uint32_t key = (s << 4) | (uint32_t)(gq + 1);
    key = min(key, dpp<0x128>(key));  // row_ror:8: the row's other group
    const uint32_t k0 = (uint32_t)__builtin_amdgcn_readlane((int)key, 0);
    const uint32_t k1 = (uint32_t)__builtin_amdgcn_readlane((int)key, 16);
    const uint32_t k2 = (uint32_t)__builtin_amdgcn_readlane((int)key, 32);
    const uint32_t k3 = (uint32_t)__builtin_amdgcn_readlane((int)key, 48);
    const uint32_t kmin = min(min(k0, k1), min(k2, k3));
    if (kmin < (bsad << 4)) {
      const int w = (int)(kmin & 15) - 1;
      const int dx = (int)((kProbeDx1 >> (2 * w)) & 3) - 1;
      const int dy = (int)((kProbeDy1 >> (2 * w)) & 3) - 1;
      bsad = kmin >> 4;
      bx += dx * (1 << sh);
      by += dy * (1 << sh);
      boff += (dy * G::lp + dx) * (1 << sh);
    }
  }
  bestx = bx;
  besty = by;
  best_sad = bsad;
}

// Maps the launch index onto strips so that each of the 8 XCDs (which
// receive workgroups round-robin) works on one contiguous band and keeps
// that band's frame rows in its own L2.
__device__ __forceinline__ int xcd_band(int i, int total) {
  const int full = (total >> 3) << 3;
  if (i >= full) return i;
  return (i & 7) * (total >> 3) + (i >> 3);
}

// One strip of group `a`: M anchor MBs of MB row `brow`, strip column
// `scol`.  `frame_at(i)` returns frame i of the group, `wt4` the weight LUT.
// TABLE: the per-block vectors and SADs come from `table` (layout
// n x brows x bcols x {mvx, mvy, sad}) instead of the descent, and the
// blend reads the neighbour straight from memory with clamped coordinates
// (the vectors are not bounded by the staged +-15 window).
// Ends with the strip resolved into the output planes; the caller
// synchronises the workgroup before reusing LDS.
template <int M, bool TABLE, class FrameAt>
__device__ __forceinline__ void process_strip(const ArnrGroup& a, const uint32_t* wt4,
                                              const FrameAt& frame_at, int brow, int scol,
                                              uint8_t* smem, const int32_t* table = nullptr) {
  using G = StripGeom<M>;
  const G g;
  // LDS: packed accumulators M*384 u32 | weight LUT 256 B | centre luma
  // 16 x 16M | centre U, V 8 x 8M each | per-wave windows.
  uint32_t* s_acc = reinterpret_cast<uint32_t*>(smem);
  uint8_t* s_lut = smem + M * 384 * 4;
  uint8_t* s_cy = s_lut + 256;
  uint8_t* s_cu = s_cy + 256 * M;
  uint8_t* s_cv = s_cu + 64 * M;
  uint8_t* s_win = s_cv + 64 * M;

  const int nw = blockDim.x >> 6;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;

  const int bx0 = scol * 16 * M, by = brow * 16;
  const int cbx0 = bx0 >> 1, cby = by >> 1;
  const int bh = min(16, a.h - by), cbh = (bh + 1) >> 1;
  const FrameDesc C = frame_at(a.center);
  const int nnb = a.filter_on ? a.n - 1 : 0;
  uint8_t* win = s_win + wave * g.wave_bytes;
  uint8_t* cwin = win + kWinRows * g.lp;

  // The first neighbour's windows are staged before anything else.
  int k = wave;
  if (!TABLE && k < nnb) stage_windows(win, frame_at(k < a.center ? k : k + 1), a, g, bx0, by, lane);

  // Centre strip (zero outside the frame), accumulators, LUT.
  for (int i = tid; i < 64 * M; i += blockDim.x) {  // luma: 16 rows x 4M dwords
    const int r = i / (4 * M), q = i - r * 4 * M;
    const int x = bx0 + 4 * q;
    uint32_t v = 0;
    if (r < bh && x < a.w) {
      const uint8_t* row = C.y + (size_t)(by + r) * C.y_stride;
      if (a.aligned4 && x + 4 <= a.w) {
        v = *reinterpret_cast<const uint32_t*>(row + x);
      } else {
        for (int j = 0; j < 4; j++)
          if (x + j < a.w) v |= (uint32_t)row[x + j] << (8 * j);
      }
    }
    reinterpret_cast<uint32_t*>(s_cy)[i] = v;
  }
  for (int i = tid; i < 32 * M; i += blockDim.x) {  // chroma: 2 planes x 8 rows x 2M dwords
    const int pl = i >= 16 * M;
    const int t = i - pl * 16 * M;
    const int r = t / (2 * M), q = t - r * 2 * M;
    const int x = cbx0 + 4 * q;
    uint32_t v = 0;
    if (r < cbh && x < a.cw) {
      const uint8_t* row = (pl ? C.v : C.u) + (size_t)(cby + r) * C.uv_stride;
      for (int j = 0; j < 4; j++)
        if (x + j < a.cw) v |= (uint32_t)row[x + j] << (8 * j);
    }
    reinterpret_cast<uint32_t*>(pl ? s_cv : s_cu)[t] = v;
  }
  for (int i = tid; i < 384 * M; i += blockDim.x) s_acc[i] = 0;
  for (int i = tid; i < 64; i += blockDim.x) reinterpret_cast<uint32_t*>(s_lut)[i] = wt4[i];
  __syncthreads();

  const int gq = lane >> 3, kk = lane & 7;   // search: probe group, row pair
  // This lane's probe direction as a window offset per unit step.
  const int gd_off = (((kProbeDy1 >> (2 * gq)) & 3) - 1) * G::lp + (((kProbeDx1 >> (2 * gq)) & 3) - 1);
  const int lr = lane >> 2, lq = lane & 3;   // blend: luma row, dword
  const int cr = lane >> 3, cc = lane & 7;   // blend: chroma pixel

  for (; k < nnb; k += nw) {
    const int fi = k < a.center ? k : k + 1;
    const FrameDesc F = frame_at(fi);
    if (!TABLE && k != wave) stage_windows(win, F, a, g, bx0, by, lane);
    wave_lds_sync();

#pragma unroll 1
    for (int m = 0; m < M; m++) {
      const int bx = bx0 + 16 * m;
      if (bx >= a.w) break;
      const int bw = min(16, a.w - bx);
      const int cbw = (bw + 1) >> 1;
      int bestx, besty;
      uint32_t best_sad;
      if (TABLE) {
        const int32_t* t = table + 3 * (((size_t)fi * a.brows + brow) * a.bcols + (bx >> 4));
        bestx = __builtin_amdgcn_readfirstlane(t[0]);
        besty = __builtin_amdgcn_readfirstlane(t[1]);
        best_sad = (uint32_t)__builtin_amdgcn_readfirstlane(t[2]);
      } else {
        const uint8_t* s_src = s_cy + 2 * kk * 16 * M + 16 * m;
        const uint4 src0 = *reinterpret_cast<const uint4*>(s_src);
        const uint4 src1 = *reinterpret_cast<const uint4*>(s_src + 16 * M);
        const uint8_t* row0 = win + (2 * kk + 15) * G::lp + 16 * m + 16;
        if (bw == 16 && bh == 16) {
          // (0, 0) operands of the whole-wave SAD: row lane / 4, dword lane % 4
          const uint8_t* zrow = win + ((lane >> 2) + 15) * G::lp + 16 * m + 16 + 4 * (lane & 3);
          const uint32_t zsrc = *reinterpret_cast<const uint32_t*>(s_cy + (lane >> 2) * 16 * M + 16 * m + 4 * (lane & 3));
          descend<true, G>(row0, src0, src1, gd_off, gq, kk, 16, 16, bestx, besty, best_sad, zrow, zsrc);
        } else {
          descend<false, G>(row0, src0, src1, gd_off, gq, kk, bw, bh, bestx, besty, best_sad, nullptr, 0);
        }
      }
      if (!TABLE && a.mv_out && lane == 0) {
        int32_t* o = a.mv_out + 3 * (((size_t)fi * a.brows + brow) * a.bcols + (bx >> 4));
        o[0] = bestx; o[1] = besty; o[2] = (int32_t)best_sad;
      }

      // sad / (bw*bh) <= 12  <=>  sad < 13 * bw * bh (integer division).
      if (best_sad < (kSadPerPixelCutoff + 1) * (uint32_t)(bw * bh)) {
        uint32_t* acc = s_acc + 384 * m;
        if (lr < bh) {
          uint32_t ref4;
          if (TABLE) {
            const uint8_t* rrow = F.y + (size_t)clampi(by + lr + besty, 0, a.h - 1) * F.y_stride;
            const int xr = bx + 4 * lq + bestx;
            ref4 = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) ref4 |= (uint32_t)rrow[clampi(xr + j, 0, a.w - 1)] << (8 * j);
          } else {
            const uint8_t* rp4 = win + (lr + besty + 15) * g.lp + 16 * m + 16 + 4 * lq + bestx;
            const uint32_t rsh = (uint32_t)(uintptr_t)rp4 & 3;
            const uint32_t* rw = reinterpret_cast<const uint32_t*>(rp4 - rsh);
            ref4 = funnel(rw[0], rw[1], rsh);
          }
          const uint32_t ctr4 = *reinterpret_cast<const uint32_t*>(s_cy + lr * 16 * M + 16 * m + 4 * lq);
#pragma unroll
          for (int j = 0; j < 4; j++) {
            if (4 * lq + j < bw) {
              const int rp = (ref4 >> (8 * j)) & 255;
              const int cp = (ctr4 >> (8 * j)) & 255;
              const uint32_t wgt = s_lut[rp > cp ? rp - cp : cp - rp];
              lds_add(&acc[lr * 16 + 4 * lq + j], (wgt << kAccShift) | (wgt * (uint32_t)rp));
            }
          }
        }
        if (cr < cbh && cc < cbw) {
          const int cmvx = bestx / 2, cmvy = besty / 2;  // truncation toward zero
          int ru, rv;
          if (TABLE) {
            const size_t o = (size_t)clampi(cby + cr + cmvy, 0, a.ch - 1) * F.uv_stride +
                             clampi((bx >> 1) + cc + cmvx, 0, a.cw - 1);
            ru = F.u[o];
            rv = F.v[o];
          } else {
            const int o = (cr + cmvy + 7) * g.cp + 8 * m + 8 + cc + cmvx;
            ru = cwin[o];
            rv = cwin[kCWinRows * g.cp + o];
          }
          const int cu = s_cu[cr * 8 * M + 8 * m + cc], cv = s_cv[cr * 8 * M + 8 * m + cc];
          const uint32_t wu = s_lut[ru > cu ? ru - cu : cu - ru];
          const uint32_t wv = s_lut[rv > cv ? rv - cv : cv - rv];
          lds_add(&acc[256 + cr * 8 + cc], (wu << kAccShift) | (wu * (uint32_t)ru));
          lds_add(&acc[320 + cr * 8 + cc], (wv << kAccShift) | (wv * (uint32_t)rv));
        }
      }
    }
    wave_lds_sync();  // window reads done before the next neighbour restages it
  }
  __syncthreads();

  // Resolve the strip: each thread handles 4 horizontally adjacent pixels.
  for (int i = tid; i < 96 * M; i += blockDim.x) {
    const int m = i / 96, j = i - m * 96;
    const int bx = bx0 + 16 * m;
    if (bx >= a.w) continue;
    const uint32_t* acc = s_acc + 384 * m;
    int x, y, plane_w, base;
    const uint8_t* ctrp;
    uint8_t* out;
    if (j < 64) {  // luma: 16 rows x 4 quads
      const int r = j >> 2, q = j & 3;
      x = bx + 4 * q; y = by + r; plane_w = a.w; base = r * 16 + 4 * q;
      ctrp = s_cy + r * 16 * M + 16 * m + 4 * q;
      out = a.out_y;
      if (r >= bh) continue;
    } else {  // chroma: 2 planes x 8 rows x 2 quads
      const int t = j - 64, pl = t >= 16, u = t & 15, r = u >> 1, q = u & 1;
      x = (bx >> 1) + 4 * q; y = cby + r; plane_w = a.cw; base = 256 + 64 * pl + r * 8 + 4 * q;
      ctrp = (pl ? s_cv : s_cu) + r * 8 * M + 8 * m + 4 * q;
      out = pl ? a.out_v : a.out_u;
      if (r >= cbh) continue;
    }
    uint32_t px4 = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const uint32_t packed = acc[base + e];
      const uint32_t sum = kWMax * ctrp[e] + (packed & ((1u << kAccShift) - 1));
      const uint32_t ws = kWMax + (packed >> kAccShift);
      px4 |= ((sum + (ws >> 1)) / ws) << (8 * e);
    }
    uint8_t* dst = out + (size_t)y * plane_w + x;
    if (x + 4 <= plane_w && ((((uintptr_t)dst) & 3) == 0)) {
      *reinterpret_cast<uint32_t*>(dst) = px4;
    } else {
      for (int e = 0; e < 4; e++)
        if (x + e < plane_w) dst[e] = (uint8_t)(px4 >> (8 * e));
    }
  }
}

template <int M>
__global__ void __launch_bounds__(1024) arnr_fused_kernel(const ArnrArgs a) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int total = a.g.strips_per_row * a.g.brows;
  const int strip = a.xcd_remap ? xcd_band(blockIdx.x, total) : (int)blockIdx.x;
  const int brow = strip / a.g.strips_per_row;
  const int scol = strip - brow * a.g.strips_per_row;
  process_strip<M, false>(a.g, a.wt4, [&](int i) { return a.frames[i]; }, brow, scol, smem);
}

// The altref from a given alignment table (guard, blend, resolve).
template <int M>
__global__ void __launch_bounds__(1024) arnr_table_kernel(const ArnrArgs a, const int32_t* table) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int total = a.g.strips_per_row * a.g.brows;
  const int strip = a.xcd_remap ? xcd_band(blockIdx.x, total) : (int)blockIdx.x;
  const int brow = strip / a.g.strips_per_row;
  const int scol = strip - brow * a.g.strips_per_row;
  process_strip<M, true>(a.g, a.wt4, [&](int i) { return a.frames[i]; }, brow, scol, smem, table);
}

// Descriptor fields read from global memory and made wave-uniform, so that
// they live in scalar registers and drive uniform branches.
__device__ __forceinline__ int uload(const int* p) { return __builtin_amdgcn_readfirstlane(*p); }
__device__ __forceinline__ uint32_t uload(const uint32_t* p) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)*p);
}
template <class T>
__device__ __forceinline__ T* uload(T* const* p) {
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)v.x);
  const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)v.y);
  return reinterpret_cast<T*>(lo | (hi << 32));
}

__device__ __forceinline__ FrameDesc load_frame(const FrameDesc* p) {
  FrameDesc f;
  f.y = uload(&p->y);
  f.u = uload(&p->u);
  f.v = uload(&p->v);
  f.y_stride = uload(&p->y_stride);
  f.uv_stride = uload(&p->uv_stride);
  return f;
}

__device__ __forceinline__ ArnrGroup load_group(const ArnrGroup* p) {
  ArnrGroup g;
  g.out_y = uload(&p->out_y);
  g.out_u = uload(&p->out_u);
  g.out_v = uload(&p->out_v);
  g.mv_out = uload(&p->mv_out);
  g.w = uload(&p->w);
  g.h = uload(&p->h);
  g.cw = uload(&p->cw);
  g.ch = uload(&p->ch);
  g.bcols = uload(&p->bcols);
  g.brows = uload(&p->brows);
  g.n = uload(&p->n);
  g.center = uload(&p->center);
  g.filter_on = uload(&p->filter_on);
  g.aligned4 = uload(&p->aligned4);
  g.aligned16 = uload(&p->aligned16);
  g.aligned4c = uload(&p->aligned4c);
  g.strips_per_row = uload(&p->strips_per_row);
  return g;
}

// Persistent batch kernel: every workgroup loops taking units (one strip of
// one group) until all chunks are drained; a unit is processed exactly as
// the single-group kernel processes that strip.
template <int M>
__global__ void __launch_bounds__(1024) arnr_batch_kernel(const ArnrBatchArgs b) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ int s_unit;
  const int first = blockIdx.x % kBatchChunks;
  int c = 0;
  while (c < kBatchChunks) {
    const int chunk = (first + c) % kBatchChunks;
    const int lo = (int)((long long)b.total_units * chunk / kBatchChunks);
    const int hi = (int)((long long)b.total_units * (chunk + 1) / kBatchChunks);
    if (threadIdx.x == 0) {
      const int t = (int)atomicAdd(&b.ticket[chunk], 1u);
      s_unit = t < hi - lo ? lo + t : -1;
    }
    __syncthreads();
    const int unit = __builtin_amdgcn_readfirstlane(s_unit);
    if (unit < 0) {
      c++;  // chunk drained: move on (every workgroup ends after kBatchChunks misses)
    } else {
      const int j = __builtin_amdgcn_readfirstlane((int)b.unit_job[unit]);
      const ArnrBatchJob* job = b.jobs + j;
      const ArnrGroup g = load_group(&job->g);
      const int fb = uload(&job->frame_base);
      const int local = unit - uload(&job->unit_base);
      const int brow = local / g.strips_per_row;
      const int scol = local - brow * g.strips_per_row;
      process_strip<M, false>(g, job->wt4, [&](int i) { return load_frame(b.frames + fb + i); }, brow, scol,
                       smem);
    }
    __syncthreads();  // LDS and s_unit are free for the next unit
  }
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

size_t arnr_fused_lds_bytes(int nw, int m) {
  const int lp = 16 * m + 40, cp = 8 * m + 16;  // StripGeom<m>::lp, ::cp
  const int wave_bytes = (kWinRows * lp + 2 * kCWinRows * cp + 15) & ~15;
  return (size_t)m * 384 * 4 + 256 + 256 * m + 128 * m + (size_t)nw * wave_bytes;
}

int arnr_supported_mbs(int m) { return m >= 8 ? 8 : (m >= 4 ? 4 : (m >= 2 ? 2 : 1)); }

hipError_t launch_arnr_fused(const ArnrArgs& a, int nw, hipStream_t stream) {
  const int blocks = a.g.strips_per_row * a.g.brows;
  const size_t lds = arnr_fused_lds_bytes(nw, a.mbs_per_wg);
  switch (a.mbs_per_wg) {
    case 1: hipLaunchKernelGGL(arnr_fused_kernel<1>, dim3(blocks), dim3(64 * nw), lds, stream, a); break;
    case 2: hipLaunchKernelGGL(arnr_fused_kernel<2>, dim3(blocks), dim3(64 * nw), lds, stream, a); break;
    case 4: hipLaunchKernelGGL(arnr_fused_kernel<4>, dim3(blocks), dim3(64 * nw), lds, stream, a); break;
    case 8: hipLaunchKernelGGL(arnr_fused_kernel<8>, dim3(blocks), dim3(64 * nw), lds, stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int M>
static int batch_occupancy(int nw) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, arnr_batch_kernel<M>, 64 * nw,
                                                   arnr_fused_lds_bytes(nw, M)) != hipSuccess)
    return 0;
  return per_cu;
}

int arnr_batch_workgroups_per_cu(int nw, int m) {
  switch (m) {
    case 1: return batch_occupancy<1>(nw);
    case 2: return batch_occupancy<2>(nw);
    case 4: return batch_occupancy<4>(nw);
    case 8: return batch_occupancy<8>(nw);
    default: return 0;
  }
}

hipError_t launch_arnr_batch(const ArnrBatchArgs& b, int nw, int workgroups, hipStream_t stream) {
  const size_t lds = arnr_fused_lds_bytes(nw, b.mbs_per_wg);
  switch (b.mbs_per_wg) {
    case 1: hipLaunchKernelGGL(arnr_batch_kernel<1>, dim3(workgroups), dim3(64 * nw), lds, stream, b); break;
    case 2: hipLaunchKernelGGL(arnr_batch_kernel<2>, dim3(workgroups), dim3(64 * nw), lds, stream, b); break;
    case 4: hipLaunchKernelGGL(arnr_batch_kernel<4>, dim3(workgroups), dim3(64 * nw), lds, stream, b); break;
    case 8: hipLaunchKernelGGL(arnr_batch_kernel<8>, dim3(workgroups), dim3(64 * nw), lds, stream, b); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_arnr_table(const ArnrArgs& a, const int32_t* table, int nw, hipStream_t stream) {
  const int blocks = a.g.strips_per_row * a.g.brows;
  const size_t lds = arnr_fused_lds_bytes(0, a.mbs_per_wg);  // no staged windows
  switch (a.mbs_per_wg) {
    case 1: hipLaunchKernelGGL(arnr_table_kernel<1>, dim3(blocks), dim3(64 * nw), lds, stream, a, table); break;
    case 2: hipLaunchKernelGGL(arnr_table_kernel<2>, dim3(blocks), dim3(64 * nw), lds, stream, a, table); break;
    case 4: hipLaunchKernelGGL(arnr_table_kernel<4>, dim3(blocks), dim3(64 * nw), lds, stream, a, table); break;
    case 8: hipLaunchKernelGGL(arnr_table_kernel<8>, dim3(blocks), dim3(64 * nw), lds, stream, a, table); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace vp8g
