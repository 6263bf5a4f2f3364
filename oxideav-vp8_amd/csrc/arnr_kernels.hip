// ARNR altref synthesis on gfx950: one workgroup per anchor macroblock,
// one wave per neighbour frame, search + occlusion guard + blend + resolve
// fused into a single launch.
//
// Semantics follow src/arnr.rs:476-675 of the reference crate:
//   * search: refine_search (src/arnr.rs:264-300) — SAD at (0,0), then for
//     step 8,4,2,1 the 8 probes around the step's frozen centre in the order
//     (-s,-s),(0,-s),(s,-s),(-s,0),(s,0),(-s,s),(0,s),(s,s); a probe wins on
//     strictly smaller SAD, so the first minimum in that order wins.  The
//     reference's row early exit (src/arnr.rs:217-219) never changes the
//     winner, so full SADs are computed.
//   * edge clamping uses the visible dims (pel, src/arnr.rs:149-153); every
//     sample a search or blend can touch (|mv| <= 15, chroma |mv/2| <= 7) is
//     staged once per (block, neighbour) into clamped LDS windows (luma
//     46x48, chroma 22x24 per plane), after which all reads are LDS reads.
//   * occlusion guard: skip the block when sad / (bw*bh) > 12
//     (src/arnr.rs:603-608).
//   * blend: acc += wt[|ref - ctr|] * ref, wsum += wt[..] with the 256-entry
//     table of src/arnr.rs:307-323; chroma uses (mv/2 truncated toward zero)
//     and the chroma-plane clamp (src/arnr.rs:402-459, 620-639).  The two
//     sums are packed into one u32 (acc < 2^20, wsum < 2^12 for <= 64
//     frames) and added with LDS atomics: integer addition is exact and
//     commutative, so the result does not depend on wave scheduling.
//   * resolve: (acc + ws/2) / ws with the centre's weight-16 seed
//     (src/arnr.rs:662-675).
// Chroma output clamping (src/arnr.rs:442-444) never fires from the block
// grid: cbx + ceil(bw/2) = ceil(min(bx+16, w)/2) <= ceil(w/2).
#include "vp8g_device.h"
#include "vp8g_internal.h"

namespace vp8g {

constexpr int kWinRows = 46;                         // by-15 .. by+30
constexpr int kWinCols = 48;                         // bx-16 .. bx+31
constexpr int kWinDwords = kWinRows * kWinCols / 4;  // 552
constexpr int kCWinRows = 22;                        // cby-7 .. cby+14
constexpr int kCWinCols = 24;                        // cbx-8 .. cbx+15
constexpr int kCWinDwords = kCWinRows * kCWinCols / 4;  // 132 per plane
constexpr int kWaveLdsDwords = kWinDwords + 2 * kCWinDwords;  // 816
constexpr int kAccShift = 20;                        // packed: wsum << 20 | acc

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

// Stage the neighbour's clamped luma and chroma windows into this wave's
// LDS region.  All global loads are issued before the first LDS store.
__device__ __forceinline__ void stage_windows(uint32_t* win, const FrameDesc& f, const ArnrArgs& a,
                                              int bx, int by, bool interior16, int lane) {
  const int cbx = bx >> 1, cby = by >> 1;
  if (interior16) {
    // 46 rows x 3 aligned 16-byte pieces, no clamping needed.
    uint4 v[3];
#pragma unroll
    for (int it = 0; it < 3; it++) {
      const int s = lane + 64 * it;
      if (s < kWinRows * 3) {
        const int wy = s / 3, part = s - wy * 3;
        v[it] = *reinterpret_cast<const uint4*>(f.y + (size_t)(by - 15 + wy) * f.y_stride + bx - 16 +
                                                16 * part);
      }
    }
    // The luma window being interior implies the chroma windows are too
    // (cbx >= 8, cbx + 16 <= cw, cby >= 8, cby + 15 <= ch).
    uint32_t c[5];
#pragma unroll
    for (int it = 0; it < 5; it++) {
      const int s = lane + 64 * it;
      if (s < 2 * kCWinDwords) {
        const int pl = s >= kCWinDwords;
        const int t = s - pl * kCWinDwords;
        const int wy = t / 6, dq = t - wy * 6;
        const uint8_t* row = (pl ? f.v : f.u) + (size_t)(cby - 7 + wy) * f.uv_stride + cbx - 8 + 4 * dq;
        c[it] = a.aligned4c ? *reinterpret_cast<const uint32_t*>(row)
                            : (uint32_t)row[0] | ((uint32_t)row[1] << 8) | ((uint32_t)row[2] << 16) |
                                  ((uint32_t)row[3] << 24);
      }
    }
#pragma unroll
    for (int it = 0; it < 3; it++) {
      const int s = lane + 64 * it;
      if (s < kWinRows * 3) reinterpret_cast<uint4*>(win)[s] = v[it];
    }
#pragma unroll
    for (int it = 0; it < 5; it++) {
      const int s = lane + 64 * it;
      if (s < 2 * kCWinDwords) win[kWinDwords + s] = c[it];
    }
    return;
  }
  // Frame-edge blocks (a few percent of a frame): clamped dword gathers.
#pragma unroll 1
  for (int s = lane; s < kWinDwords; s += 64) {
    const int wy = s / 12, dq = s - wy * 12;
    const int y = clampi(by - 15 + wy, 0, a.h - 1);
    win[s] = load_clamped4(f.y + (size_t)y * f.y_stride, bx - 16 + 4 * dq, a.w, a.aligned4 != 0);
  }
#pragma unroll 1
  for (int s = lane; s < 2 * kCWinDwords; s += 64) {
    const int pl = s >= kCWinDwords;
    const int t = s - pl * kCWinDwords;
    const int wy = t / 6, dq = t - wy * 6;
    const int y = clampi(cby - 7 + wy, 0, a.ch - 1);
    win[kWinDwords + s] = load_clamped4((pl ? f.v : f.u) + (size_t)y * f.uv_stride, cbx - 8 + 4 * dq,
                                        a.cw, a.aligned4c != 0);
  }
}

// Maps the launch index onto the MB grid so that each of the 8 XCDs (which
// receive workgroups round-robin) works on one contiguous band of MBs and
// keeps that band's frame rows in its own L2.
__device__ __forceinline__ int xcd_band(int i, int total) {
  const int full = (total >> 3) << 3;
  if (i >= full) return i;
  return (i & 7) * (total >> 3) + (i >> 3);
}

__global__ void __launch_bounds__(1024) arnr_fused_kernel(const ArnrArgs a) {
  extern __shared__ __align__(16) uint32_t smem[];
  uint32_t* s_acc = smem;                                   // 256 luma + 64 U + 64 V, packed
  uint8_t* s_lut = reinterpret_cast<uint8_t*>(smem + 384);  // 256 bytes
  uint32_t* s_win = smem + 384 + 64;                        // per wave: luma | U | V windows

  const int nw = blockDim.x >> 6;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;

  const int mb = a.xcd_remap ? xcd_band(blockIdx.x, a.bcols * a.brows) : (int)blockIdx.x;
  const int bcol = mb % a.bcols;
  const int brow = mb / a.bcols;
  const int bx = bcol * 16, by = brow * 16;
  const int bw = min(16, a.w - bx), bh = min(16, a.h - by);
  const int cbx = bx >> 1, cby = by >> 1;
  const int cbw = (bw + 1) >> 1, cbh = (bh + 1) >> 1;
  const FrameDesc C = a.frames[a.center];
  const bool full_block = bw == 16 && bh == 16;
  const bool interior16 = a.aligned16 && bx >= 16 && bx + 32 <= a.w && by >= 15 && by + 31 <= a.h;

  const int nnb = a.filter_on ? a.n - 1 : 0;
  uint32_t* win = s_win + wave * kWaveLdsDwords;
  const int g = lane >> 3, kk = lane & 7;   // search: probe group, row pair
  const int lr = lane >> 2, lq = lane & 3;  // blend: luma row, dword
  const int cr = lane >> 3, cc = lane & 7;  // blend: chroma pixel

  // First neighbour's windows are issued before anything else.
  int k = wave;
  if (k < nnb) stage_windows(win, a.frames[k < a.center ? k : k + 1], a, bx, by, interior16, lane);

  // Centre samples this lane needs: 2 source rows for the search, one
  // dword for the luma blend, one U and one V pixel.  Zero outside bw x bh.
  uint32_t src[8];
  uint32_t ctr_l = 0;
  int cu8 = 0, cv8 = 0;
  const bool chroma_in = cr < cbh && cc < cbw;
  if (full_block && a.aligned16) {
#pragma unroll
    for (int rr = 0; rr < 2; rr++) {
      const uint4 q = *reinterpret_cast<const uint4*>(C.y + (size_t)(by + 2 * kk + rr) * C.y_stride + bx);
      src[rr * 4 + 0] = q.x; src[rr * 4 + 1] = q.y; src[rr * 4 + 2] = q.z; src[rr * 4 + 3] = q.w;
    }
    ctr_l = *reinterpret_cast<const uint32_t*>(C.y + (size_t)(by + lr) * C.y_stride + bx + 4 * lq);
  } else {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int r = 2 * kk + (i >> 2), q = i & 3;
      uint32_t v = 0;
      if (r < bh) {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (4 * q + j < bw) v |= (uint32_t)C.y[(size_t)(by + r) * C.y_stride + bx + 4 * q + j] << (8 * j);
      }
      src[i] = v;
    }
    if (lr < bh) {
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (4 * lq + j < bw) ctr_l |= (uint32_t)C.y[(size_t)(by + lr) * C.y_stride + bx + 4 * lq + j] << (8 * j);
    }
  }
  if (chroma_in) {
    cu8 = C.u[(size_t)(cby + cr) * C.uv_stride + cbx + cc];
    cv8 = C.v[(size_t)(cby + cr) * C.uv_stride + cbx + cc];
  }
  for (int i = tid; i < 384; i += blockDim.x) s_acc[i] = 0;
  for (int i = tid; i < 64; i += blockDim.x) reinterpret_cast<uint32_t*>(s_lut)[i] = a.wt4[i];
  __syncthreads();

  uint32_t colmask[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (4 * q + j < bw) m |= 0xFFu << (8 * j);
    colmask[q] = m;
  }
  const int gdx = (g == 0 || g == 3 || g == 5) ? -1 : ((g == 1 || g == 6) ? 0 : 1);
  const int gdy = g < 3 ? -1 : (g < 5 ? 0 : 1);
  const uint32_t area = (uint32_t)(bw * bh);

  for (; k < nnb; k += nw) {
    const int fi = k < a.center ? k : k + 1;
    if (k != wave) stage_windows(win, a.frames[fi], a, bx, by, interior16, lane);
    wave_lds_sync();

    // SAD of probe (cx, cy) summed over this lane's 8-lane group.
    auto probe = [&](int cx, int cy) -> uint32_t {
      uint32_t acc = 0;
#pragma unroll
      for (int rr = 0; rr < 2; rr++) {
        const int r = 2 * kk + rr;
        if (r < bh) {
          const int off = (r + cy + 15) * kWinCols + cx + 16;
          const uint32_t* p = win + (off >> 2);
          const uint32_t sh = off & 3;
          const uint32_t w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
          uint32_t a0 = funnel(w0, w1, sh), a1 = funnel(w1, w2, sh);
          uint32_t a2 = funnel(w2, w3, sh), a3 = funnel(w3, w4, sh);
          if (bw < 16) {
            a0 &= colmask[0]; a1 &= colmask[1]; a2 &= colmask[2]; a3 &= colmask[3];
          }
          acc = sad4(src[rr * 4 + 0], a0, acc);
          acc = sad4(src[rr * 4 + 1], a1, acc);
          acc = sad4(src[rr * 4 + 2], a2, acc);
          acc = sad4(src[rr * 4 + 3], a3, acc);
        }
      }
      return sum8(acc);
    };

    int bestx = 0, besty = 0;
    uint32_t best_sad = (uint32_t)__builtin_amdgcn_readfirstlane((int)probe(0, 0));
#pragma unroll 1
    for (int step = 8; step >= 1; step >>= 1) {
      const int px = bestx + gdx * step, py = besty + gdy * step;
      const bool valid = px >= -15 && px <= 15 && py >= -15 && py <= 15;
      uint32_t s = probe(valid ? px : 0, valid ? py : 0);
      if (!valid) s = 0xFFFFFFFFu;
      const int cx0 = bestx, cy0 = besty;
#pragma unroll
      for (int gg = 0; gg < 8; gg++) {
        const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)s, gg * 8);
        if (v < best_sad) {
          const int ddx = (gg == 0 || gg == 3 || gg == 5) ? -1 : ((gg == 1 || gg == 6) ? 0 : 1);
          const int ddy = gg < 3 ? -1 : (gg < 5 ? 0 : 1);
          best_sad = v;
          bestx = cx0 + ddx * step;
          besty = cy0 + ddy * step;
        }
      }
    }
    if (a.mv_out && lane == 0) {
      int32_t* o = a.mv_out + 3 * ((size_t)fi * a.brows * a.bcols + mb);
      o[0] = bestx; o[1] = besty; o[2] = (int32_t)best_sad;
    }

    if (best_sad / area <= kSadPerPixelCutoff) {
      if (lr < bh) {
        const int off = (lr + besty + 15) * kWinCols + 4 * lq + bestx + 16;
        const uint32_t ref4 = funnel(win[off >> 2], win[(off >> 2) + 1], off & 3);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (4 * lq + j < bw) {
            const int rp = (ref4 >> (8 * j)) & 255;
            const int cp = (ctr_l >> (8 * j)) & 255;
            const uint32_t wgt = s_lut[rp > cp ? rp - cp : cp - rp];
            lds_add(&s_acc[lr * 16 + 4 * lq + j], (wgt << kAccShift) | (wgt * (uint32_t)rp));
          }
        }
      }
      if (chroma_in) {
        const int cmvx = bestx / 2, cmvy = besty / 2;  // truncation toward zero
        const uint8_t* cw8 = reinterpret_cast<const uint8_t*>(win + kWinDwords);
        const int o = (cr + cmvy + 7) * kCWinCols + cc + cmvx + 8;
        const int ru = cw8[o], rv = cw8[kCWinDwords * 4 + o];
        const uint32_t wu = s_lut[ru > cu8 ? ru - cu8 : cu8 - ru];
        const uint32_t wv = s_lut[rv > cv8 ? rv - cv8 : cv8 - rv];
        lds_add(&s_acc[256 + cr * 8 + cc], (wu << kAccShift) | (wu * (uint32_t)ru));
        lds_add(&s_acc[320 + cr * 8 + cc], (wv << kAccShift) | (wv * (uint32_t)rv));
      }
    }
    wave_lds_sync();  // window reads done before the next neighbour restages it
  }
  __syncthreads();

  // Resolve: lane-parallel over the block's pixels, dword stores where the
  // row is whole and aligned.
  for (int i = tid; i < 384; i += blockDim.x) {
    uint32_t ctr;
    uint8_t* dst;
    bool in;
    if (i < 256) {
      const int r = i >> 4, c = i & 15;
      in = r < bh && c < bw;
      ctr = in ? C.y[(size_t)(by + r) * C.y_stride + bx + c] : 0;
      dst = a.out_y + (size_t)(by + r) * a.w + bx + c;
    } else {
      const int j = (i - 256) & 63, r = j >> 3, c = j & 7;
      in = r < cbh && c < cbw;
      const uint8_t* pl = i < 320 ? C.u : C.v;
      ctr = in ? pl[(size_t)(cby + r) * C.uv_stride + cbx + c] : 0;
      dst = (i < 320 ? a.out_u : a.out_v) + (size_t)(cby + r) * a.cw + cbx + c;
    }
    const uint32_t packed = s_acc[i];
    const uint32_t acc = kWMax * ctr + (packed & ((1u << kAccShift) - 1));
    const uint32_t ws = kWMax + (packed >> kAccShift);
    if (in) *dst = (uint8_t)((acc + (ws >> 1)) / ws);
  }
}

size_t arnr_fused_lds_bytes(int nw) { return (size_t)(384 + 64 + nw * kWaveLdsDwords) * 4; }

hipError_t launch_arnr_fused(const ArnrArgs& a, int nw, hipStream_t stream) {
  const int blocks = a.bcols * a.brows;
  hipLaunchKernelGGL(arnr_fused_kernel, dim3(blocks), dim3(64 * nw), arnr_fused_lds_bytes(nw),
                     stream, a);
  return hipGetLastError();
}

}  // namespace vp8g
