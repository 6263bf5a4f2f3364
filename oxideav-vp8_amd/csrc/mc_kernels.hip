// VP8 motion compensation and sub-pel SAD scoring on gfx950: one wave per
// macroblock, the edge-clamped halo staged in LDS, the two section 20.14
// passes run out of LDS, results written as whole dwords.
//
// Semantics follow the reference crate:
//   * predict_inter_mb, src/motion_comp.rs:1986-2078: luma MV doubled to
//     eighth-pel (wrapping i16, :187-192), chroma MV = avg of four copies
//     (:209-239), optional full-pixel truncation (:256-261), clamp bounds
//     mb_cols*16 x mb_rows*16 (chroma *8) (:2005-2008).
//   * halo origin (mb_x + (mv >> 3) - 2, mb_y + (mv >> 3) - 2), 21x21 luma /
//     13x13 chroma, every read clamped to the plane (:837-880, :1153-1196).
//   * pass = clamp255((sum_k s_k * f_k + 64) >> 7) (:396-402); horizontal
//     pass over all halo rows into an 8-bit temp, then the vertical pass
//     (:1018-1057, :1311-1350).  A tap row equal to {0,0,128,0,0,0} is an
//     exact identity (:919-1007), so the whole-pixel copies
//     (:1513-1613) are the same bytes as both passes with row 0.
//   * mb_luma_sad_at_mv, src/motion_search.rs:470-529: luma prediction with
//     the six-tap set, then the 16x16 SAD (:153-161).
//   * half/quarter_pixel_refine_luma, src/motion_search.rs:564-728: centre
//     plus 8 offsets, components clamped to +-1023, strict '<' in order.
#include "vp8g_device.h"
#include "vp8g_internal.h"

namespace vp8g {

__constant__ int32_t kSixtap[8][6] = {
    {0, 0, 128, 0, 0, 0},     {0, -6, 123, 12, -1, 0}, {2, -11, 108, 36, -8, 1},
    {0, -9, 93, 50, -6, 0},   {3, -16, 77, 77, -16, 3}, {0, -6, 50, 93, -9, 0},
    {1, -8, 36, 108, -11, 2}, {0, -1, 12, 123, -6, 0}};

struct Taps {
  int32_t t[6];
  bool ident;
};

__device__ __forceinline__ Taps load_taps(const int32_t* row) {
  Taps k;
#pragma unroll
  for (int i = 0; i < 6; i++) k.t[i] = row[i];
  k.ident = k.t[0] == 0 && k.t[1] == 0 && k.t[2] == 128 && k.t[3] == 0 && k.t[4] == 0 && k.t[5] == 0;
  return k;
}

// One section 18.3 output sample; i32 arithmetic wraps as the reference's
// release build does.
__device__ __forceinline__ uint32_t interp6(const Taps& k, int s0, int s1, int s2, int s3, int s4,
                                            int s5) {
  uint32_t a = (uint32_t)s0 * (uint32_t)k.t[0] + (uint32_t)s1 * (uint32_t)k.t[1] +
               (uint32_t)s2 * (uint32_t)k.t[2] + (uint32_t)s3 * (uint32_t)k.t[3] +
               (uint32_t)s4 * (uint32_t)k.t[4] + (uint32_t)s5 * (uint32_t)k.t[5];
  // clamp255(x >> 7) written as (clamp(x, 0, 32767)) >> 7 (equal for every
  // x): the shift-then-saturate form gets selected to v_ashr_pk_u8_i32,
  // whose result's upper half is then assumed zero by the code that packs
  // the bytes, which corrupts bytes 2-3 once a tap sum reaches 2^16.
  const int x = (int)(a + 64u);
  return (uint32_t)min(max(x, 0), 32767) >> 7;
}

// Clamped gather of a ROWS x COLS block at (x0, y0), spread over the wave:
// load() issues every byte load into registers, store() writes them to LDS
// (row pitch `pitch`), so all loads of a macroblock are in flight at once.
template <int ROWS, int COLS>
struct Gather {
  static constexpr int kTotal = ROWS * COLS;
  static constexpr int kIters = (kTotal + 63) / 64;
  uint8_t v[kIters];

  __device__ __forceinline__ void load(const uint8_t* plane, uint32_t stride, int w, int h, int x0,
                                       int y0, int lane) {
#pragma unroll
    for (int it = 0; it < kIters; it++) {
      const int s = lane + 64 * it;
      if (s < kTotal) {
        const int r = s / COLS, c = s - r * COLS;
        const int y = clampi(y0 + r, 0, h - 1), x = clampi(x0 + c, 0, w - 1);
        v[it] = plane[(size_t)y * stride + x];
      }
    }
  }

  __device__ __forceinline__ void store(uint8_t* dst, int pitch, int lane) const {
#pragma unroll
    for (int it = 0; it < kIters; it++) {
      const int s = lane + 64 * it;
      if (s < kTotal) {
        const int r = s / COLS, c = s - r * COLS;
        dst[r * pitch + c] = v[it];
      }
    }
  }
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Two-pass filter of an OW x OW block from its (OW+5)^2 halo held in LDS
// (pitch hp); the horizontal result goes to `tmp` (pitch OW).  Returns, for
// the calling lane, the 4 output samples of unit `u` packed little-endian:
// row u / (OW/4), columns 4*(u % (OW/4)) .. +3.  Units >= OW*OW/4 return 0.
template <int OW>
__device__ __forceinline__ void hpass(const uint8_t* halo, int hp, uint8_t* tmp, const Taps& fh,
                                      int lane) {
  constexpr int HW = OW + 5;
  constexpr int UPR = OW / 4;           // 4-sample units per row
  constexpr int UNITS = HW * UPR;
  for (int u = lane; u < UNITS; u += 64) {
    const int r = u / UPR, c0 = 4 * (u % UPR);
    // Three naturally aligned dwords hold the 9 support samples (unaligned
    // LDS reads stall the LDS pipe).
    const uint32_t* s32 = reinterpret_cast<const uint32_t*>(halo + r * hp + c0);
    const uint32_t w0 = s32[0], w1 = s32[1], w2 = s32[2];
    uint32_t packed;
    if (fh.ident) {
      packed = funnel(w0, w1, 2);
    } else {
      int v[9];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        v[j] = (w0 >> (8 * j)) & 255;
        v[j + 4] = (w1 >> (8 * j)) & 255;
      }
      v[8] = w2 & 255;
      packed = 0;
#pragma unroll
      for (int j = 0; j < 4; j++)
        packed |= interp6(fh, v[j], v[j + 1], v[j + 2], v[j + 3], v[j + 4], v[j + 5]) << (8 * j);
    }
    *reinterpret_cast<uint32_t*>(tmp + r * OW + c0) = packed;
  }
}

template <int OW>
__device__ __forceinline__ uint32_t vpass_unit(const uint8_t* tmp, const Taps& fv, int u) {
  constexpr int UPR = OW / 4;
  const int r = u / UPR, c0 = 4 * (u % UPR);
  const uint32_t* t = reinterpret_cast<const uint32_t*>(tmp + r * OW + c0);
  constexpr int RD = OW / 4;  // dword pitch
  if (fv.ident) return t[2 * RD];
  uint32_t rows[6];
#pragma unroll
  for (int k = 0; k < 6; k++) rows[k] = t[k * RD];
  uint32_t packed = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int sh = 8 * j;
    packed |= interp6(fv, (rows[0] >> sh) & 255, (rows[1] >> sh) & 255, (rows[2] >> sh) & 255,
                      (rows[3] >> sh) & 255, (rows[4] >> sh) & 255, (rows[5] >> sh) & 255)
              << sh;
  }
  return packed;
}

__device__ __forceinline__ int16_t dbl16(int16_t v) { return (int16_t)(uint16_t)((uint16_t)v * 2u); }

__device__ __forceinline__ int avg4(int c) {
  const int s = 4 * c;
  return s >= 0 ? (s + 4) >> 3 : -((-s + 4) >> 3);
}

// LDS per wave: luma halo 21 x 24, luma temp 21 x 16, two chroma halos
// 13 x 16, two chroma temps 13 x 8.
constexpr int kMcLdsPerWave = 21 * 24 + 21 * 16 + 2 * 13 * 16 + 2 * 13 * 8;

__global__ void __launch_bounds__(256) predict_inter_mb_kernel(const McArgs a) {
  __shared__ __align__(16) uint8_t smem[4 * kMcLdsPerWave];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n) return;  // whole wave exits together
  uint8_t* lh = smem + wave * kMcLdsPerWave;
  uint8_t* lt = lh + 21 * 24;
  uint8_t* chalo = lt + 21 * 16;  // U at 0, V at 13*16
  uint8_t* ctmp = chalo + 2 * 13 * 16;

  const int mb_col = (int)a.mb_xy[2 * i], mb_row = (int)a.mb_xy[2 * i + 1];
  int16_t yr = dbl16(a.mv[2 * i]), yc = dbl16(a.mv[2 * i + 1]);
  int16_t ur = (int16_t)avg4(yr), uc = (int16_t)avg4(yc);
  if (a.full_pixel) {
    yr = (int16_t)(yr & ~7); yc = (int16_t)(yc & ~7);
    ur = (int16_t)(ur & ~7); uc = (int16_t)(uc & ~7);
  }
  // A whole-pixel vector is a plain copy whatever row 0 of the caller's
  // table holds (src/motion_comp.rs:2022-2028, 2044-2067).
  Taps fh = load_taps(a.filters[yc & 7]);
  Taps fv = load_taps(a.filters[yr & 7]);
  Taps cfh = load_taps(a.filters[uc & 7]);
  Taps cfv = load_taps(a.filters[ur & 7]);
  if ((yc & 7) == 0 && (yr & 7) == 0) fh.ident = fv.ident = true;
  if ((uc & 7) == 0 && (ur & 7) == 0) cfh.ident = cfv.ident = true;

  Gather<21, 21> gy;
  Gather<13, 13> gu, gv;
  gy.load(a.ref.y, a.ref.y_stride, a.ref.lw, a.ref.lh, mb_col * 16 + (yc >> 3) - 2,
          mb_row * 16 + (yr >> 3) - 2, lane);
  const int cx0 = mb_col * 8 + (uc >> 3) - 2, cy0 = mb_row * 8 + (ur >> 3) - 2;
  gu.load(a.ref.u, a.ref.uv_stride, a.ref.cw, a.ref.ch, cx0, cy0, lane);
  gv.load(a.ref.v, a.ref.uv_stride, a.ref.cw, a.ref.ch, cx0, cy0, lane);
  gy.store(lh, 24, lane);
  gu.store(chalo, 16, lane);
  gv.store(chalo + 13 * 16, 16, lane);
  wave_lds_sync();
  hpass<16>(lh, 24, lt, fh, lane);
  if (lane < 26) {
    hpass<8>(chalo, 16, ctmp, cfh, lane);  // lanes 0..25 -> U
  } else if (lane < 52) {
    hpass<8>(chalo + 13 * 16, 16, ctmp + 13 * 8, cfh, lane - 26);  // V
  }
  wave_lds_sync();
  uint8_t* out = a.out + (size_t)i * 384;
  *reinterpret_cast<uint32_t*>(out + 4 * lane) = vpass_unit<16>(lt, fv, lane);
  if (lane < 16) {
    *reinterpret_cast<uint32_t*>(out + 256 + 4 * lane) = vpass_unit<8>(ctmp, cfv, lane);
  } else if (lane < 32) {
    *reinterpret_cast<uint32_t*>(out + 320 + 4 * (lane - 16)) =
        vpass_unit<8>(ctmp + 13 * 8, cfv, lane - 16);
  }
}

// 16x16 luma prediction for eighth-pel MV (r8, c8) whose halo lies inside
// `win` (pitch 24) at offset (ox, oy); returns this lane's 4 samples.
__device__ __forceinline__ uint32_t luma_pred_from_window(const uint8_t* win, int ox, int oy,
                                                          int r8, int c8, uint8_t* lt, int lane) {
  const Taps fh = load_taps(kSixtap[c8 & 7]);
  const Taps fv = load_taps(kSixtap[r8 & 7]);
  hpass<16>(win + oy * 24 + ox, 24, lt, fh, lane);
  wave_lds_sync();
  const uint32_t p = vpass_unit<16>(lt, fv, lane);
  wave_lds_sync();
  return p;
}

__global__ void __launch_bounds__(256) mb_luma_sad_at_mv_kernel(const SadArgs a) {
  __shared__ __align__(16) uint8_t smem[4 * (21 * 24 + 21 * 16)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n) return;
  uint8_t* lh = smem + wave * (21 * 24 + 21 * 16);
  uint8_t* lt = lh + 21 * 24;
  const int mb_col = (int)a.mb_xy[2 * i], mb_row = (int)a.mb_xy[2 * i + 1];
  const int16_t r8 = dbl16(a.mv[2 * i]), c8 = dbl16(a.mv[2 * i + 1]);
  Gather<21, 21> gy;
  gy.load(a.plane, a.stride, a.w, a.h, mb_col * 16 + (c8 >> 3) - 2, mb_row * 16 + (r8 >> 3) - 2,
          lane);
  const uint32_t s = *reinterpret_cast<const uint32_t*>(a.src + (size_t)i * 256 + 4 * lane);
  gy.store(lh, 24, lane);
  wave_lds_sync();
  const uint32_t p = luma_pred_from_window(lh, 0, 0, r8, c8, lt, lane);
  const uint32_t tot = wave_sum(sad4(s, p, 0));
  if (lane == 0) a.sad[i] = tot;
}

// Centre + 8 offsets: every candidate's integer part lies within one
// pixel of the centre's, so one clamped 24x24 window at the smallest
// integer origin minus 2 holds all nine 21x21 halos (clamping depends only
// on the absolute coordinate, so a sub-window equals a direct fetch).
__global__ void __launch_bounds__(256) refine_ladder_kernel(const RefineArgs a) {
  __shared__ __align__(16) uint8_t smem[4 * (24 * 24 + 21 * 16)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n) return;
  uint8_t* win = smem + wave * (24 * 24 + 21 * 16);
  uint8_t* lt = win + 24 * 24;
  const int mb_col = (int)a.mb_xy[2 * i], mb_row = (int)a.mb_xy[2 * i + 1];
  // Centres outside +-1023 are outside the reference's contract (it
  // debug-asserts); clamping keeps every candidate inside the staged window.
  const int crow = clampi(a.center[2 * i], -1023, 1023), ccol = clampi(a.center[2 * i + 1], -1023, 1023);
  const int st = a.step;
  const int offr[8] = {-st, -st, -st, 0, 0, st, st, st};
  const int offc[8] = {-st, 0, st, -st, st, -st, 0, st};
  int cr[9], ccl[9];
  cr[0] = crow; ccl[0] = ccol;
  int minr = 1 << 30, minc = 1 << 30;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    if (k > 0) {
      cr[k] = clampi(crow + offr[k - 1], -1023, 1023);
      ccl[k] = clampi(ccol + offc[k - 1], -1023, 1023);
    }
    minr = min(minr, (int)(dbl16((int16_t)cr[k]) >> 3));
    minc = min(minc, (int)(dbl16((int16_t)ccl[k]) >> 3));
  }
  Gather<24, 24> gw;
  gw.load(a.plane, a.stride, a.w, a.h, mb_col * 16 + minc - 2, mb_row * 16 + minr - 2, lane);
  const uint32_t s = *reinterpret_cast<const uint32_t*>(a.src + (size_t)i * 256 + 4 * lane);
  gw.store(win, 24, lane);
  wave_lds_sync();
  int best_r = cr[0], best_c = ccl[0];
  uint32_t best_sad = 0;
#pragma unroll 1
  for (int k = 0; k < 9; k++) {
    const int16_t r8 = dbl16((int16_t)cr[k]), c8 = dbl16((int16_t)ccl[k]);
    if (k > 0 && cr[k] == best_r && ccl[k] == best_c) continue;  // collapsed onto the best
    const uint32_t p = luma_pred_from_window(win, (c8 >> 3) - minc, (r8 >> 3) - minr, r8, c8, lt, lane);
    const uint32_t tot = wave_sum(sad4(s, p, 0));
    if (k == 0 || tot < best_sad) {
      best_sad = tot;
      best_r = cr[k];
      best_c = ccl[k];
    }
  }
  if (lane == 0) {
    a.out[2 * i] = ((uint32_t)best_r & 0xFFFFu) | ((uint32_t)best_c << 16);
    a.out[2 * i + 1] = best_sad;
  }
}

hipError_t launch_predict_inter_mb(const McArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(predict_inter_mb_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_mb_luma_sad_at_mv(const SadArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mb_luma_sad_at_mv_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

// HalfPlanes (src/encoder/inter_search.rs:256-346): the 22x22 window at
// (x-3, y-3), edge-clamped (fetch_window :266-287); the horizontal plane
// h[r][c] = interp(win[r][c..c+5]) over all 22 rows (or rows 3..18 without
// diagonals), v[r][c] = interp(win[r..r+5][c+3]) 17x16, the diagonal plane
// d[r][c] = interp(h[r..r+5][c]) 17x17; the SAD for half-pel offset
// (dx, dy) reads plane position (dx > 0, dy > 0) (HalfPlanes::sad :334-345).
// One wave per block; outputs the 8 SADs in (dy, dx) raster order, the
// diagonal ones as 0xFFFFFFFF when `diagonals` is off.
constexpr int kHpLds = 22 * 24 + 22 * 17 + 17 * 16 + 17 * 17;

__global__ void __launch_bounds__(256) half_plane_sads_kernel(const HalfPlaneArgs a) {
  __shared__ __align__(16) uint8_t smem[4 * ((kHpLds + 15) & ~15)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n) return;
  uint8_t* win = smem + wave * ((kHpLds + 15) & ~15);
  uint8_t* hp = win + 22 * 24;
  uint8_t* vp = hp + 22 * 17;
  uint8_t* dp = vp + 17 * 16;
  const int x = a.xy[2 * i], y = a.xy[2 * i + 1];
  Gather<22, 22> gw;
  gw.load(a.plane, a.stride, a.w, a.h, x - 3, y - 3, lane);
  const uint32_t s4 = *reinterpret_cast<const uint32_t*>(a.src + (size_t)i * 256 + 4 * lane);
  gw.store(win, 24, lane);
  Taps t;
#pragma unroll
  for (int k = 0; k < 6; k++) t.t[k] = a.taps[k];
  t.ident = false;
  wave_lds_sync();
  const int r0 = a.diagonals ? 0 : 3, r1 = a.diagonals ? 22 : 19;
  for (int o = lane; o < (r1 - r0) * 17; o += 64) {
    const int r = r0 + o / 17, c = o % 17;
    const uint8_t* s = win + r * 24 + c;
    hp[r * 17 + c] = (uint8_t)interp6(t, s[0], s[1], s[2], s[3], s[4], s[5]);
  }
  for (int o = lane; o < 17 * 16; o += 64) {
    const int r = o >> 4, c = o & 15;
    const uint8_t* s = win + r * 24 + c + 3;
    vp[o] = (uint8_t)interp6(t, s[0], s[24], s[48], s[72], s[96], s[120]);
  }
  wave_lds_sync();
  if (a.diagonals) {
    for (int o = lane; o < 17 * 17; o += 64) {
      const uint8_t* s = hp + o;
      dp[o] = (uint8_t)interp6(t, s[0], s[17], s[34], s[51], s[68], s[85]);
    }
  }
  wave_lds_sync();
  const int pr = lane >> 2, pc = 4 * (lane & 3);
  int k = 0;
#pragma unroll
  for (int dy = -1; dy <= 1; dy++) {
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
      if (dx == 0 && dy == 0) continue;
      const int cx = dx > 0, cy = dy > 0;
      const uint8_t* p;
      if (dy == 0) p = hp + (3 + pr) * 17 + cx + pc;
      else if (dx == 0) p = vp + (cy + pr) * 16 + pc;
      else p = dp + (cy + pr) * 17 + cx + pc;
      const uint32_t pred = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
      const uint32_t tot = wave_sum(sad4(s4, pred, 0));
      if (lane == 0) a.sad[8 * i + k] = (dx != 0 && dy != 0 && !a.diagonals) ? 0xFFFFFFFFu : tot;
      k++;
    }
  }
}

hipError_t launch_half_plane_sads(const HalfPlaneArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(half_plane_sads_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_refine_ladder(const RefineArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(refine_ladder_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace vp8g
