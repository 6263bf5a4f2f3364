// Device-side helpers shared by the gfx950 kernels: packed-u8 SAD,
// byte funnel shifts, DPP wave reductions.  CDNA4 (wave64) only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vp8g {

// Unaligned 32-bit window of the byte stream (lo:hi little-endian) starting
// `shift` bytes into `lo` (v_alignbyte_b32).
__device__ __forceinline__ uint32_t funnel(uint32_t lo, uint32_t hi, uint32_t shift) {
  return __builtin_amdgcn_alignbyte(hi, lo, shift);
}

// v_sad_u8: acc + sum over the 4 byte lanes of |a - b|.
__device__ __forceinline__ uint32_t sad4(uint32_t a, uint32_t b, uint32_t acc) {
  return __builtin_amdgcn_sad_u8(a, b, acc);
}

// DPP controls (gfx9 encoding).
constexpr int kDppQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141; // lane i <- lane 7-i within 8
constexpr int kDppRowMirror = 0x140;     // lane i <- lane 15-i within 16

template <int Ctrl>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, 0xF, 0xF, false);
}

// Sum over each aligned group of 8 lanes; every lane of the group gets it.
__device__ __forceinline__ uint32_t sum8(uint32_t v) {
  v += dpp<kDppQuadXor1>(v);
  v += dpp<kDppQuadXor2>(v);
  v += dpp<kDppRowHalfMirror>(v);
  return v;
}

// Sum over each aligned group of 16 lanes; every lane of the group gets it.
__device__ __forceinline__ uint32_t sum16(uint32_t v) {
  v = sum8(v);
  v += dpp<kDppRowMirror>(v);
  return v;
}

// Full wave64 sum, valid in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  v = sum16(v);
  v += (uint32_t)__shfl_xor((int)v, 16, 64);
  v += (uint32_t)__shfl_xor((int)v, 32, 64);
  return v;
}

// Minimum over the whole wave, returned wave-uniform: DPP within each row
// of 16 lanes, then the four row results.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  v = min(v, dpp<kDppQuadXor1>(v));
  v = min(v, dpp<kDppQuadXor2>(v));
  v = min(v, dpp<kDppRowHalfMirror>(v));
  v = min(v, dpp<kDppRowMirror>(v));
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
  const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
  const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
  const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
  return min(min(r0, r1), min(r2, r3));
}

// Row broadcasts (gfx9 DPP): lane 15 of each row to the next row
// (row_bcast:15), lane 31 to rows 2 and 3 (row_bcast:31).  Rows outside
// `RowMask` receive `old`.
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

template <int Ctrl, int RowMask>
__device__ __forceinline__ uint32_t dpp_rows(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, Ctrl, RowMask, 0xF, false);
}

// Given a value that is uniform within each row of 16 lanes: the sum over
// the four rows, wave-uniform (lane 63 after two row broadcasts).
__device__ __forceinline__ uint32_t wave_sum_to_sgpr(uint32_t v) {
  v += dpp_rows<kDppRowBcast15, 0xA>(0u, v);  // rows 1, 3: + previous row
  v += dpp_rows<kDppRowBcast31, 0xC>(0u, v);  // rows 2, 3: + rows 0..1
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Minimum over the whole wave, wave-uniform (any lane values).
__device__ __forceinline__ uint32_t wave_min_any(uint32_t v) {
  v = min(v, dpp<kDppQuadXor1>(v));
  v = min(v, dpp<kDppQuadXor2>(v));
  v = min(v, dpp<kDppRowHalfMirror>(v));
  v = min(v, dpp<kDppRowMirror>(v));
  v = min(v, dpp_rows<kDppRowBcast15, 0xA>(~0u, v));
  v = min(v, dpp_rows<kDppRowBcast31, 0xC>(~0u, v));
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Given a value that is uniform within each aligned group of 8 lanes: the
// minimum over the wave, wave-uniform.
__device__ __forceinline__ uint32_t wave_min_to_sgpr(uint32_t v) {
  v = min(v, dpp<0x128>(v));                       // row_ror:8: the row's other group
  v = min(v, dpp_rows<kDppRowBcast15, 0xA>(~0u, v));  // rows 1, 3: min with previous row
  v = min(v, dpp_rows<kDppRowBcast31, 0xC>(~0u, v));  // rows 2, 3: min with rows 0..1
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// LDS writes of the wave visible to the wave's later LDS reads (and no
// access moved across): lanes exchange data through LDS within one wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Maps a launch index onto work items so that each of the 8 XCDs (which
// receive workgroups round-robin) works on one contiguous band of items and
// keeps that band's frame rows in its own L2.
__device__ __forceinline__ int xcd_band(int i, int total) {
  const int full = (total >> 3) << 3;
  if (i >= full) return i;
  return (i & 7) * (total >> 3) + (i >> 3);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Wave-uniform reads of per-item inputs the kernel never writes, through the
// constant address space: scalar loads (s_load) into SGPRs, whose latency is
// below a vector load's.  `p` must be 4-byte aligned.
__device__ __forceinline__ uint32_t uniform_u32(const void* p) {
  return *(const __attribute__((address_space(4))) uint32_t*)p;
}
// A pair of int16 (little-endian: element 0 in the low half).
__device__ __forceinline__ void uniform_i16x2(const int16_t* p, int16_t& e0, int16_t& e1) {
  const uint32_t w = uniform_u32(p);
  e0 = (int16_t)(w & 0xFFFFu);
  e1 = (int16_t)(w >> 16);
}

}  // namespace vp8g
