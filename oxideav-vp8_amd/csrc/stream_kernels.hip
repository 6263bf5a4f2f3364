// The stream driver's scene-cut metric on gfx950: the luma SAD behind
// luma_mad (src/stream.rs:2654-2666) for many frame pairs in one launch.
//
// Grid (row bands, pairs); each thread sums |a - b| over 16-byte pieces of
// its rows with v_sad_u8 (4 bytes per op) when both rows are 4-byte
// aligned, byte by byte otherwise; a workgroup reduces its sums and adds
// them to the pair's 64-bit total with one atomic (integer addition
// commutes, so the total does not depend on the launch shape).
#include "vp8g_device.h"
#include "vp8g_internal.h"

namespace vp8g {

__global__ void __launch_bounds__(256) luma_sad_pairs_kernel(const LumaSadArgs a) {
  __shared__ uint32_t s_part[4];
  const LumaSadPair P = a.pairs[blockIdx.y];
  const int rows_per_band = (a.h + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * rows_per_band, r1 = min(a.h, r0 + rows_per_band);
  const int pieces = (a.w + 15) / 16;
  uint32_t acc = 0;
  for (int s = threadIdx.x; s < (r1 - r0) * pieces; s += 256) {
    const int r = r0 + s / pieces, x = 16 * (s % pieces);
    const uint8_t* pa = P.a + (size_t)r * P.a_stride + x;
    const uint8_t* pb = P.b + (size_t)r * P.b_stride + x;
    if (x + 16 <= a.w && (((uintptr_t)pa | (uintptr_t)pb) & 3) == 0) {
      const uint4 va = *reinterpret_cast<const uint4*>(pa);
      const uint4 vb = *reinterpret_cast<const uint4*>(pb);
      acc = sad4(va.x, vb.x, acc);
      acc = sad4(va.y, vb.y, acc);
      acc = sad4(va.z, vb.z, acc);
      acc = sad4(va.w, vb.w, acc);
    } else {
      const int e = min(16, a.w - x);
      for (int k = 0; k < e; k++) acc += (uint32_t)abs((int)pa[k] - (int)pb[k]);
    }
  }
  // per thread at most ceil(rows * pieces / 256) pieces of 16 bytes: far
  // below 2^32 / (255 * 16) for any frame the entry point accepts
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t t = (uint64_t)s_part[0] + s_part[1] + s_part[2] + s_part[3];
    if (t) atomicAdd(reinterpret_cast<unsigned long long*>(a.sad + blockIdx.y), (unsigned long long)t);
  }
}

__global__ void zero_u64_kernel(uint64_t* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

hipError_t launch_luma_sad_pairs(const LumaSadArgs& a, int pairs, hipStream_t stream) {
  if (pairs <= 0) return hipSuccess;
  if (pairs > kLumaSadPairsPerLaunch) return hipErrorInvalidValue;
  hipLaunchKernelGGL(zero_u64_kernel, dim3((pairs + 255) / 256), dim3(256), 0, stream, a.sad, pairs);
  // about 16 rows per band, at most 64 bands per pair
  const int bands = max(1, min(64, (a.h + 15) / 16));
  hipLaunchKernelGGL(luma_sad_pairs_kernel, dim3(bands, pairs), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace vp8g
