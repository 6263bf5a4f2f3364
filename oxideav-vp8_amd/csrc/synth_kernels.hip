// Synthetic region-motion luma frames on the GPU (training data for the
// learned motion proposer).  The content model is the one the test
// workloads use: frame t of a group samples a blurred-noise texture
// displaced by t times its region's velocity (8x8 regions), then
// rectangles are replaced by fresh blurred noise, then Gaussian noise is
// added and the result rounded and clipped.  Integer stages reproduce the
// host generator exactly; the Gaussian stage uses the device's fp64 log and
// cos, so a pixel may differ by one where the noisy value rounds on a
// half-way point.
#include "synth.h"

namespace vp8g {

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// blurred_noise(seed, h, w)[y][x]: 3x3 box blur (edge replicated, (s+4)/9)
// of the top bytes of splitmix64 draws 1 .. h*w.
__device__ __forceinline__ uint32_t blurred_at(uint64_t seed, int h, int w, int y, int x) {
  uint32_t s = 0;
#pragma unroll
  for (int dy = -1; dy <= 1; dy++) {
    const int yy = clampi(y + dy, 0, h - 1);
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
      const int xx = clampi(x + dx, 0, w - 1);
      s += (uint32_t)(splitmix_at(seed, (uint64_t)yy * w + xx + 1) >> 56);
    }
  }
  return (s + 4) / 9;
}

__global__ void synth_luma_kernel(const SynthArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long plane = (long long)a.w * a.h;
  if (i >= plane * a.n) return;
  const int t = (int)(i / plane);
  const int p = (int)(i - (long long)t * plane);
  const int y = p / a.w, x = p - y * a.w;
  const int ry = min(y * 8 / a.h, 7), rx = min(x * 8 / a.w, 7);
  const int reg = ry * 8 + rx;
  const int sx = clampi(x - t * a.vel[2 * reg], 0, a.w - 1);
  const int sy = clampi(y - t * a.vel[2 * reg + 1], 0, a.h - 1);
  uint32_t v = blurred_at(a.texture_seed, a.h, a.w, sy, sx);
  const SynthRect* rects = a.rects + (size_t)t * a.rects_per_frame;
  for (int k = 0; k < a.rects_per_frame; k++) {
    const SynthRect r = rects[k];
    if (x >= r.x0 && x < r.x0 + r.w && y >= r.y0 && y < r.y0 + r.h)
      v = blurred_at(r.seed, r.h, r.w, y - r.y0, x - r.x0);
  }
  if (a.clean) a.clean[i] = (uint8_t)v;
  // gaussian(seed_t, w*h, sigma)[p]: draws 2p+1, 2p+2 (Box-Muller).
  const uint64_t seed = a.noise_seeds[t];
  const uint64_t r1 = splitmix_at(seed, 2 * (uint64_t)p + 1), r2 = splitmix_at(seed, 2 * (uint64_t)p + 2);
  const double u1 = ((double)(r1 >> 11) + 1.0) / 9007199254740993.0;
  const double u2 = (double)(r2 >> 11) / 9007199254740992.0;
  const double g = a.sigma * sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
  const double o = rint((double)v + g);
  a.out[i] = (uint8_t)(o < 0.0 ? 0.0 : (o > 255.0 ? 255.0 : o));
}

hipError_t launch_synth_luma(const SynthArgs& a, hipStream_t s) {
  const long long n = (long long)a.w * a.h * a.n;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(synth_luma_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace vp8g
