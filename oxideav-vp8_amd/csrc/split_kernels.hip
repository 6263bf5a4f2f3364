// SPLITMV motion compensation and inter-macroblock reconstruction on
// gfx950: predict_split_mv (src/motion_comp.rs:2227-2320) with
// filter_block_4x4 (:1693-1712) and split_chroma_mvs (:2162-2203), and the
// residual of reconstruct_inter_mb / reconstruct_split_mv_mb (:2098-2150,
// :2340-2385): inverse WHT of the Y2 block (src/inverse_transform.rs:190-240)
// and the inverse DCT of every 4x4 block added to the prediction with
// clamping (:417-480, 609-678, 871-875).
//
// One wave per macroblock.  Prediction: lane l takes row l % 4 of luma
// sub-block l / 4 under that sub-block's own vector, then lanes 0..31 take
// the rows of the eight chroma sub-blocks; each lane filters its four
// samples from clamped reads of the plane (the 9x9 support of the 4x4
// block, H pass over the six rows it needs, then the V pass).  Residual:
// lanes 0..23 each run one 4x4 block's inverse DCT in the reference's
// wrapping 32-bit arithmetic and add it to the prediction in place.
#include "vp8g_device.h"
#include "vp8g_internal.h"

namespace vp8g {

namespace {

__device__ __forceinline__ uint8_t px_at(const uint8_t* plane, uint32_t stride, int w, int h, int x, int y) {
  return plane[(size_t)clampi(y, 0, h - 1) * stride + clampi(x, 0, w - 1)];
}

__device__ __forceinline__ int interp_taps(const int32_t* f, const int* s) {
  int acc = 64;
#pragma unroll
  for (int k = 0; k < 6; k++) acc += s[k] * f[k];
  return clampi(acc >> 7, 0, 255);
}

// Row rr (0..3) of filter_block_4x4 for the 4x4 block at (bx, by) under the
// eighth-pel vector (row8, col8), packed little-endian.
__device__ __forceinline__ uint32_t filter_row4(const uint8_t* plane, uint32_t stride, int w, int h, int bx, int by,
                                                int row8, int col8, int rr, const int32_t (*filters)[6]) {
  const int mx = col8 & 7, my = row8 & 7;
  const int ix = bx + (col8 >> 3), iy = by + (row8 >> 3);
  uint32_t out = 0;
  if (mx == 0 && my == 0) {
#pragma unroll
    for (int c = 0; c < 4; c++) out |= (uint32_t)px_at(plane, stride, w, h, ix + c, iy + rr) << (8 * c);
    return out;
  }
  // Horizontal pass over halo rows rr .. rr + 5 (halo row t is plane row
  // iy - 2 + t), columns 0..3 of the block; then the vertical pass.
  int hv[6][4];
#pragma unroll
  for (int t = 0; t < 6; t++) {
    int s[9];
#pragma unroll
    for (int k = 0; k < 9; k++) s[k] = px_at(plane, stride, w, h, ix - 2 + k, iy - 2 + rr + t);
#pragma unroll
    for (int c = 0; c < 4; c++) hv[t][c] = interp_taps(filters[mx], s + c);
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int s[6] = {hv[0][c], hv[1][c], hv[2][c], hv[3][c], hv[4][c], hv[5][c]};
    out |= (uint32_t)interp_taps(filters[my], s) << (8 * c);
  }
  return out;
}

// avg() of four doubled luma components (src/motion_comp.rs:209-216).
__device__ __forceinline__ int avg_doubled(int a, int b, int c, int d) {
  const int s = 2 * a + 2 * b + 2 * c + 2 * d;
  return s >= 0 ? (s + 4) >> 3 : -((-s + 4) >> 3);
}

}  // namespace

__global__ void __launch_bounds__(256) split_predict_kernel(const SplitArgs a) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n) return;
  const int mbx = (int)min(a.mb_xy[2 * i], 0xFFFFFu), mby = (int)min(a.mb_xy[2 * i + 1], 0xFFFFFu);
  const int16_t* mv = a.mvs + 32 * (size_t)i;
  uint8_t* out = a.out + 384 * (size_t)i;
  {
    const int b = l >> 2, rr = l & 3, sb = b >> 2, sc = b & 3;
    int r8 = (int16_t)(2 * mv[2 * b]), c8 = (int16_t)(2 * mv[2 * b + 1]);  // stored_luma_mv
    if (a.full_pixel) {
      r8 &= ~7;
      c8 &= ~7;
    }
    const uint32_t v = filter_row4(a.ref.y, a.ref.y_stride, a.ref.lw, a.ref.lh, mbx * 16 + sc * 4, mby * 16 + sb * 4,
                                   r8, c8, rr, a.filters);
    *reinterpret_cast<uint32_t*>(out + (sb * 4 + rr) * 16 + sc * 4) = v;
  }
  if (l < 32) {
    const int p = l >> 4, c = (l >> 2) & 3, rr = l & 3, sb = c >> 1, sc = c & 1;
    const int b0 = (2 * sb) * 4 + 2 * sc;
    const int bs[4] = {b0, b0 + 1, b0 + 4, b0 + 5};
    int r8 = (int16_t)avg_doubled(mv[2 * bs[0]], mv[2 * bs[1]], mv[2 * bs[2]], mv[2 * bs[3]]);
    int c8 = (int16_t)avg_doubled(mv[2 * bs[0] + 1], mv[2 * bs[1] + 1], mv[2 * bs[2] + 1], mv[2 * bs[3] + 1]);
    if (a.full_pixel) {
      r8 &= ~7;
      c8 &= ~7;
    }
    const uint32_t v = filter_row4(p ? a.ref.v : a.ref.u, a.ref.uv_stride, a.ref.cw, a.ref.ch, mbx * 8 + sc * 4,
                                   mby * 8 + sb * 4, r8, c8, rr, a.filters);
    *reinterpret_cast<uint32_t*>(out + 256 + 64 * p + (sb * 4 + rr) * 8 + sc * 4) = v;
  }
}

namespace {

__device__ __forceinline__ int32_t wadd(int32_t x, int32_t y) { return (int32_t)((uint32_t)x + (uint32_t)y); }
__device__ __forceinline__ int32_t wsub(int32_t x, int32_t y) { return (int32_t)((uint32_t)x - (uint32_t)y); }
__device__ __forceinline__ int32_t wmul(int32_t x, int32_t y) { return (int32_t)((uint32_t)x * (uint32_t)y); }

// One 4-point pass of the inverse DCT (SINPI8_SQRT2 = 35468,
// COSPI8_SQRT2_MINUS1 = 20091); outputs in rows 0, 1, 2, 3 order.
__device__ __forceinline__ void idct4(int32_t i0, int32_t i4, int32_t i8, int32_t i12, int32_t o[4]) {
  const int32_t a1 = wadd(i0, i8), b1 = wsub(i0, i8);
  const int32_t c1 = wsub(wmul(i4, 35468) >> 16, wadd(i12, wmul(i12, 20091) >> 16));
  const int32_t d1 = wadd(wadd(i4, wmul(i4, 20091) >> 16), wmul(i12, 35468) >> 16);
  o[0] = wadd(a1, d1);
  o[1] = wadd(b1, c1);
  o[2] = wsub(b1, c1);
  o[3] = wsub(a1, d1);
}

}  // namespace

__global__ void __launch_bounds__(256) mb_residual_kernel(const ResidualArgs a) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n || a.skip[i] || l >= 24) return;
  const int blocks = a.y2 ? 25 : 24;
  const int16_t* cf = a.coeffs + (size_t)i * blocks * 16;
  const int16_t* blk = cf + (a.y2 ? 16 : 0) + 16 * l;
  int32_t x[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = blk[k];
  if (a.y2 && l < 16) {
    // inverse_wht_4x4 of the Y2 block; output element l seeds block l's DC
    int32_t t[16];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int32_t a1 = cf[c] + cf[12 + c], b1 = cf[4 + c] + cf[8 + c];
      const int32_t c1 = cf[4 + c] - cf[8 + c], d1 = cf[c] - cf[12 + c];
      t[c] = a1 + b1;
      t[4 + c] = c1 + d1;
      t[8 + c] = a1 - b1;
      t[12 + c] = d1 - c1;
    }
    const int32_t* q = t + 4 * (l >> 2);
    const int32_t a1 = q[0] + q[3], b1 = q[1] + q[2], c1 = q[1] - q[2], d1 = q[0] - q[3];
    const int32_t v[4] = {a1 + b1, c1 + d1, a1 - b1, d1 - c1};
    x[0] = (int16_t)((v[l & 3] + 3) >> 3);
  }
  int32_t t[16], o[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    idct4(x[c], x[4 + c], x[8 + c], x[12 + c], o);
#pragma unroll
    for (int k = 0; k < 4; k++) t[4 * k + c] = o[k];
  }
  uint8_t* base;
  int stride;
  if (l < 16) {
    base = a.out + 384 * (size_t)i + (l >> 2) * 64 + (l & 3) * 4;
    stride = 16;
  } else {
    const int k = (l - 16) & 3;
    base = a.out + 384 * (size_t)i + (l < 20 ? 256 : 320) + (k >> 1) * 32 + (k & 1) * 4;
    stride = 8;
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    idct4(t[4 * r], t[4 * r + 1], t[4 * r + 2], t[4 * r + 3], o);
    uint32_t* p = reinterpret_cast<uint32_t*>(base + r * stride);
    const uint32_t pred = *p;
    uint32_t res = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int d = (int16_t)(wadd(o[k], 4) >> 3);
      res |= (uint32_t)clampi((int)((pred >> (8 * k)) & 255u) + d, 0, 255) << (8 * k);
    }
    *p = res;
  }
}

hipError_t launch_split_predict(const SplitArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(split_predict_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_mb_residual(const ResidualArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mb_residual_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace vp8g
