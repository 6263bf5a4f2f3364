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
//   * edge clamping uses the visible luma dims (pel, src/arnr.rs:149-153);
//     every sample a search or blend can touch (|mv| <= 15) is staged once
//     per (block, neighbour) into a clamped 46x48 LDS window, after which
//     all reads are plain LDS reads.
//   * occlusion guard: skip the block when sad / (bw*bh) > 12
//     (src/arnr.rs:603-608).
//   * blend: acc += wt[|ref - ctr|] * ref, wsum += wt[..] with the 256-entry
//     table of src/arnr.rs:307-323; chroma uses (mv/2 truncated toward zero)
//     and the chroma-plane clamp (src/arnr.rs:402-459, 620-639).  Sums are
//     u32 and commutative, so LDS atomics across waves are exact and the
//     result does not depend on wave scheduling.
//   * resolve: (acc + ws/2) / ws (src/arnr.rs:670-675).
// Chroma output clamping (src/arnr.rs:442-444) never fires from the block
// grid: cbx + ceil(bw/2) = ceil(min(bx+16, w)/2) <= ceil(w/2).
#include "vp8g_device.h"
#include "vp8g_internal.h"

namespace vp8g {

constexpr int kWinRows = 46;        // by-15 .. by+30
constexpr int kWinCols = 48;        // bx-16 .. bx+31 (dword aligned)
constexpr int kWinDwords = kWinRows * kWinCols / 4;  // 552

// Stage the clamped luma window of `f` around block (bx, by) into `win`.
__device__ __forceinline__ void stage_window(uint32_t* win, const FrameDesc& f, int bx, int by,
                                             int w, int h, bool aligned4, int lane) {
  for (int s = lane; s < kWinDwords; s += 64) {
    const int wy = s / 12;
    const int dq = s - wy * 12;
    const int y = clampi(by - 15 + wy, 0, h - 1);
    const uint8_t* row = f.y + (size_t)y * f.y_stride;
    const int x0 = bx - 16 + 4 * dq;
    uint32_t v;
    if (aligned4 && x0 >= 0 && x0 + 4 <= w) {
      v = *reinterpret_cast<const uint32_t*>(row + x0);
    } else {
      v = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) v |= (uint32_t)row[clampi(x0 + j, 0, w - 1)] << (8 * j);
    }
    win[s] = v;
  }
}

__device__ __forceinline__ void lds_add(uint32_t* p, uint32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(1024) arnr_fused_kernel(const ArnrArgs a) {
  extern __shared__ uint32_t smem[];
  uint32_t* s_ctr = smem;             // 16 rows x 4 dwords, zero outside bw x bh
  uint32_t* s_acc = smem + 64;        // 256 luma + 64 U + 64 V
  uint32_t* s_ws = s_acc + 384;
  uint8_t* s_lut = reinterpret_cast<uint8_t*>(s_ws + 384);  // 256 bytes
  uint32_t* s_win = s_ws + 384 + 64;

  const int nw = blockDim.x >> 6;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;

  const int mb = blockIdx.x;
  const int bcol = mb % a.bcols;
  const int brow = mb / a.bcols;
  const int bx = bcol * 16, by = brow * 16;
  const int bw = min(16, a.w - bx), bh = min(16, a.h - by);
  const int cbx = bx >> 1, cby = by >> 1;
  const int cbw = (bw + 1) >> 1, cbh = (bh + 1) >> 1;
  const FrameDesc C = a.frames[a.center];

  // Centre block (masked), centre-weight seed of the accumulators, LUT.
  if (tid < 64) {
    const int r = tid >> 2, q = tid & 3;
    uint32_t v = 0;
    if (r < bh) {
      const uint8_t* row = C.y + (size_t)(by + r) * C.y_stride + bx;
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (4 * q + j < bw) v |= (uint32_t)row[4 * q + j] << (8 * j);
    }
    s_ctr[tid] = v;
  }
  for (int i = tid; i < 384; i += blockDim.x) {
    uint32_t p = 0;
    if (i < 256) {
      const int r = i >> 4, c = i & 15;
      if (r < bh && c < bw) p = C.y[(size_t)(by + r) * C.y_stride + bx + c];
    } else {
      const int j = (i - 256) & 63, r = j >> 3, c = j & 7;
      const uint8_t* pl = i < 320 ? C.u : C.v;
      if (r < cbh && c < cbw) p = pl[(size_t)(cby + r) * C.uv_stride + cbx + c];
    }
    s_acc[i] = kWMax * p;
    s_ws[i] = kWMax;
  }
  for (int i = tid; i < 256; i += blockDim.x) s_lut[i] = a.wt[i];
  __syncthreads();

  // Per-lane constants for the search (8 probe groups x 8 row pairs) and
  // the blend (4 luma pixels, 1 U and 1 V pixel per lane).
  const int g = lane >> 3, kk = lane & 7;
  uint32_t src[8];
#pragma unroll
  for (int i = 0; i < 8; i++) src[i] = s_ctr[(2 * kk + (i >> 2)) * 4 + (i & 3)];
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
  const int lr = lane >> 2, lq = lane & 3;      // luma blend pixel row / dword
  const int cr = lane >> 3, cc = lane & 7;      // chroma blend pixel
  const bool chroma_in = cr < cbh && cc < cbw;
  int cu8 = 0, cv8 = 0;
  if (chroma_in) {
    cu8 = C.u[(size_t)(cby + cr) * C.uv_stride + cbx + cc];
    cv8 = C.v[(size_t)(cby + cr) * C.uv_stride + cbx + cc];
  }
  const uint32_t ctr_l = s_ctr[lr * 4 + lq];
  const uint32_t area = (uint32_t)(bw * bh);

  uint32_t* win = s_win + wave * kWinDwords;
  const int nnb = a.filter_on ? a.n - 1 : 0;

  for (int k = wave; k < nnb; k += nw) {
    const int fi = k < a.center ? k : k + 1;
    const FrameDesc F = a.frames[fi];
    stage_window(win, F, bx, by, a.w, a.h, a.aligned4 != 0, lane);
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
          uint32_t w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
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
    if (a.mv_out) {
      if (lane == 0) {
        int32_t* o = a.mv_out + 3 * ((size_t)fi * a.brows * a.bcols + mb);
        o[0] = bestx; o[1] = besty; o[2] = (int32_t)best_sad;
      }
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
            lds_add(&s_acc[lr * 16 + 4 * lq + j], wgt * (uint32_t)rp);
            lds_add(&s_ws[lr * 16 + 4 * lq + j], wgt);
          }
        }
      }
      if (chroma_in) {
        const int cmvx = bestx / 2, cmvy = besty / 2;  // truncation toward zero
        const int x = clampi(cbx + cc + cmvx, 0, a.cw - 1);
        const int y = clampi(cby + cr + cmvy, 0, a.ch - 1);
        const size_t o = (size_t)y * F.uv_stride + x;
        const int ru = F.u[o], rv = F.v[o];
        const int pix = cr * 8 + cc;
        const uint32_t wu = s_lut[ru > cu8 ? ru - cu8 : cu8 - ru];
        const uint32_t wv = s_lut[rv > cv8 ? rv - cv8 : cv8 - rv];
        lds_add(&s_acc[256 + pix], wu * (uint32_t)ru);
        lds_add(&s_ws[256 + pix], wu);
        lds_add(&s_acc[320 + pix], wv * (uint32_t)rv);
        lds_add(&s_ws[320 + pix], wv);
      }
    }
    wave_lds_sync();  // window reads done before the next neighbour restages it
  }
  __syncthreads();

  for (int i = tid; i < 384; i += blockDim.x) {
    const uint32_t ws = s_ws[i];
    const uint8_t px = (uint8_t)((s_acc[i] + (ws >> 1)) / ws);
    if (i < 256) {
      const int r = i >> 4, c = i & 15;
      if (r < bh && c < bw) a.out_y[(size_t)(by + r) * a.w + bx + c] = px;
    } else {
      const int j = (i - 256) & 63, r = j >> 3, c = j & 7;
      uint8_t* pl = i < 320 ? a.out_u : a.out_v;
      if (r < cbh && c < cbw) pl[(size_t)(cby + r) * a.cw + cbx + c] = px;
    }
  }
}

size_t arnr_fused_lds_bytes(int nw) { return (size_t)(64 + 384 + 384 + 64 + nw * kWinDwords) * 4; }

hipError_t launch_arnr_fused(const ArnrArgs& a, int nw, hipStream_t stream) {
  const int blocks = a.bcols * a.brows;
  hipLaunchKernelGGL(arnr_fused_kernel, dim3(blocks), dim3(64 * nw), arnr_fused_lds_bytes(nw),
                     stream, a);
  return hipGetLastError();
}

}  // namespace vp8g
