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

// ---- Dot-product form of the six-tap passes --------------------------------
// A tap row as two signed-byte dwords for v_dot4_i32_i8 (taps 0..3, taps
// 4..5).  With every sample s given as s ^ 0x80 (= s - 128 as a signed
// byte), sum_k (s_k - 128) t_k + 128 * sum_k t_k is the unsigned sum; the
// form applies to rows whose taps lie in [-128, 127] and sum to 128, which
// is every non-identity row of the VP8 six-tap and bilinear tables
// (dot_table_ok checks a caller's table).  Identity rows are copies.
struct DotTaps {
  int32_t lo, hi;
  bool ident;
};

// A table row packed for the dot-product form: taps 0..3 and 4..5 as
// signed bytes (host and device).
constexpr uint32_t pack_lo(const int32_t* t) {
  return (uint32_t)(t[0] & 255) | ((uint32_t)(t[1] & 255) << 8) | ((uint32_t)(t[2] & 255) << 16) |
         ((uint32_t)(t[3] & 255) << 24);
}
constexpr uint32_t pack_hi(const int32_t* t) { return (uint32_t)(t[4] & 255) | ((uint32_t)(t[5] & 255) << 8); }
constexpr bool is_identity(const int32_t* t) {
  return t[0] == 0 && t[1] == 0 && t[2] == 128 && t[3] == 0 && t[4] == 0 && t[5] == 0;
}

// The six-tap set (RFC 6386 section 18.3), packed at compile time; row 0 is
// the identity.
constexpr int32_t kSixtapRows[8][6] = {
    {0, 0, 128, 0, 0, 0},     {0, -6, 123, 12, -1, 0}, {2, -11, 108, 36, -8, 1},
    {0, -9, 93, 50, -6, 0},   {3, -16, 77, 77, -16, 3}, {0, -6, 50, 93, -9, 0},
    {1, -8, 36, 108, -11, 2}, {0, -1, 12, 123, -6, 0}};
__constant__ uint32_t kSixtapLo[8] = {pack_lo(kSixtapRows[0]), pack_lo(kSixtapRows[1]), pack_lo(kSixtapRows[2]),
                                      pack_lo(kSixtapRows[3]), pack_lo(kSixtapRows[4]), pack_lo(kSixtapRows[5]),
                                      pack_lo(kSixtapRows[6]), pack_lo(kSixtapRows[7])};
__constant__ uint32_t kSixtapHi[8] = {pack_hi(kSixtapRows[0]), pack_hi(kSixtapRows[1]), pack_hi(kSixtapRows[2]),
                                      pack_hi(kSixtapRows[3]), pack_hi(kSixtapRows[4]), pack_hi(kSixtapRows[5]),
                                      pack_hi(kSixtapRows[6]), pack_hi(kSixtapRows[7])};

// Taps of fraction f of the six-tap set (only row 0 is an identity).
__device__ __forceinline__ DotTaps sixtap_dot(int f, bool ident) {
  DotTaps k;
  k.lo = (int32_t)kSixtapLo[f];
  k.hi = (int32_t)kSixtapHi[f];
  k.ident = ident || f == 0;
  return k;
}

// Taps of fraction f of a caller's table packed on the host (McArgs).
__device__ __forceinline__ DotTaps table_dot(const McArgs& a, int f, bool ident) {
  DotTaps k;
  k.lo = (int32_t)a.dot_lo[f];
  k.hi = (int32_t)a.dot_hi[f];
  k.ident = ident || ((a.ident_rows >> f) & 1);
  return k;
}

constexpr uint32_t kBias = 0x80808080u;

// clamp255((sum_k s_k t_k + 64) >> 7): `a` holds samples 0..3 and `b` samples
// 4, 5 at the byte positions of the set taps of `hi`, all as s ^ 0x80.
__device__ __forceinline__ uint32_t tap6(uint32_t a, int32_t lo, uint32_t b, int32_t hi) {
  const int x = __builtin_amdgcn_sdot4((int)a, lo, __builtin_amdgcn_sdot4((int)b, hi, 16384 + 64, false), false);
  return (uint32_t)min(max(x, 0), 32767) >> 7;
}

// Horizontal pass over ROWS rows (default OW + 5) of the halo whose row r
// starts `sh` bytes into the dword at halo + r * hp; writes the ROWS x OW
// result to `tmp` (pitch OW) as samples ^ 0x80.  A unit is 4 outputs of one row: its
// 9 support samples lie in 3 aligned dwords.
template <int OW, int ROWS = OW + 5>
__device__ __forceinline__ void hpass_dot(const uint8_t* halo, int hp, uint32_t sh, uint8_t* tmp, const DotTaps& f,
                                          int lane) {
  constexpr int UPR = OW / 4;
  constexpr int UNITS = ROWS * UPR;
  for (int u = lane; u < UNITS; u += 64) {
    const int r = u / UPR, c0 = 4 * (u % UPR);
    const uint32_t* s32 = reinterpret_cast<const uint32_t*>(halo + r * hp + c0);
    const uint32_t w0 = s32[0], w1 = s32[1], w2 = s32[2];
    const uint32_t u0 = funnel(w0, w1, sh) ^ kBias, u1 = funnel(w1, w2, sh) ^ kBias;  // samples 0..3, 4..7
    uint32_t packed;
    if (f.ident) {
      packed = funnel(u0, u1, 2);
    } else {
      const uint32_t u2 = funnel(w2, w2, sh) ^ kBias;  // byte 0: sample 8
      const uint32_t o0 = tap6(u0, f.lo, u1, f.hi);
      const uint32_t o1 = tap6(funnel(u0, u1, 1), f.lo, u1, f.hi << 8);
      const uint32_t o2 = tap6(funnel(u0, u1, 2), f.lo, u1, f.hi << 16);
      const uint32_t o3 = tap6(funnel(u0, u1, 3), f.lo, funnel(u1, u2, 3), f.hi);
      packed = (o0 | (o1 << 8) | (o2 << 16) | (o3 << 24)) ^ kBias;
    }
    *reinterpret_cast<uint32_t*>(tmp + r * OW + c0) = packed;
  }
}

// Vertical pass for unit u (row u / (OW/4), columns 4 (u % (OW/4)) .. +3)
// over `tmp` as hpass_dot leaves it; returns the 4 output samples.  The
// 6 x 4 support block is transposed with byte permutes so that each
// column's samples sit in one dword.
template <int OW>
__device__ __forceinline__ uint32_t vpass_dot(const uint8_t* tmp, const DotTaps& f, int u) {
  constexpr int UPR = OW / 4;
  constexpr int RD = OW / 4;  // dword pitch
  const int r = u / UPR, c0 = 4 * (u % UPR);
  const uint32_t* t = reinterpret_cast<const uint32_t*>(tmp + r * OW + c0);
  if (f.ident) return t[2 * RD] ^ kBias;
  const uint32_t r0 = t[0], r1 = t[RD], r2 = t[2 * RD], r3 = t[3 * RD], r4 = t[4 * RD], r5 = t[5 * RD];
  // v_perm_b32(s0, s1, sel): selector byte 0..3 picks s1's bytes, 4..7 s0's
  const uint32_t t0 = __builtin_amdgcn_perm(r1, r0, 0x05010400u);  // r0b0 r1b0 r0b1 r1b1
  const uint32_t t1 = __builtin_amdgcn_perm(r1, r0, 0x07030602u);  // r0b2 r1b2 r0b3 r1b3
  const uint32_t t2 = __builtin_amdgcn_perm(r3, r2, 0x05010400u);
  const uint32_t t3 = __builtin_amdgcn_perm(r3, r2, 0x07030602u);
  const uint32_t b01 = __builtin_amdgcn_perm(r5, r4, 0x05010400u);  // r4b0 r5b0 r4b1 r5b1
  const uint32_t b23 = __builtin_amdgcn_perm(r5, r4, 0x07030602u);
  const uint32_t o0 = tap6(__builtin_amdgcn_perm(t2, t0, 0x05040100u), f.lo, b01, f.hi);
  const uint32_t o1 = tap6(__builtin_amdgcn_perm(t2, t0, 0x07060302u), f.lo, b01, f.hi << 16);
  const uint32_t o2 = tap6(__builtin_amdgcn_perm(t3, t1, 0x05040100u), f.lo, b23, f.hi);
  const uint32_t o3 = tap6(__builtin_amdgcn_perm(t3, t1, 0x07060302u), f.lo, b23, f.hi << 16);
  return o0 | (o1 << 8) | (o2 << 16) | (o3 << 24);
}

// Unclamped gather of ROWS x DW dwords from the dword-aligned column xa of
// rows y0 .. y0+ROWS-1 (the caller checked the rectangle lies inside the
// clamp bounds and the plane is 4-byte aligned), stored to LDS at `pitch`.
template <int ROWS, int DW>
struct GatherDw {
  static constexpr int kTotal = ROWS * DW;
  static constexpr int kIters = (kTotal + 63) / 64;
  uint32_t v[kIters];

  __device__ __forceinline__ void load(const uint8_t* plane, uint32_t stride, int xa, int y0, int lane) {
#pragma unroll
    for (int it = 0; it < kIters; it++) {
      const int s = min(lane + 64 * it, kTotal - 1);
      const int r = s / DW, q = s - r * DW;
      v[it] = *reinterpret_cast<const uint32_t*>(plane + (size_t)(y0 + r) * stride + xa + 4 * q);
    }
  }

  __device__ __forceinline__ void store(uint8_t* dst, int pitch, int lane) const {
#pragma unroll
    for (int it = 0; it < kIters; it++) {
      const int s = lane + 64 * it;
      if (s < kTotal) {
        const int r = s / DW, q = s - r * DW;
        *reinterpret_cast<uint32_t*>(dst + r * pitch + 4 * q) = v[it];
      }
    }
  }
};

// The ROWS x COLS halo at (x0, y0) of a plane clamped to (w, h), staged in
// two phases so that several halos' loads are in flight together: load()
// issues the loads into registers (whole dwords from the dword-aligned
// column below x0 for interior halos of 4-byte aligned planes, clamped
// bytes otherwise), store() writes them to LDS (pitch `pitch`, a multiple of
// 4 with room for COLS + 3 bytes) and returns the halo's byte shift within
// each LDS row.
template <int ROWS, int COLS>
struct HaloStage {
  static constexpr int DW = (COLS + 3 + 3) / 4;  // dwords covering COLS bytes at any shift
  GatherDw<ROWS, DW> dw;
  Gather<ROWS, COLS> bytes;
  bool fast;
  uint32_t sh;

  __device__ __forceinline__ void load(const uint8_t* plane, uint32_t stride, int w, int h, int x0, int y0,
                                       bool aligned4, int lane) {
    const int xa = x0 & ~3;
    fast = aligned4 && xa >= 0 && xa + 4 * DW <= w && y0 >= 0 && y0 + ROWS <= h;
    if (fast) {
      dw.load(plane, stride, xa, y0, lane);
      sh = (uint32_t)(x0 - xa);
    } else {
      bytes.load(plane, stride, w, h, x0, y0, lane);
      sh = 0;
    }
  }

  __device__ __forceinline__ uint32_t store(uint8_t* dst, int pitch, int lane) const {
    if (fast)
      dw.store(dst, pitch, lane);
    else
      bytes.store(dst, pitch, lane);
    return sh;
  }
};

template <int ROWS, int COLS>
__device__ __forceinline__ uint32_t stage_halo(const uint8_t* plane, uint32_t stride, int w, int h, int x0, int y0,
                                               bool aligned4, uint8_t* dst, int pitch, int lane) {
  HaloStage<ROWS, COLS> st;
  st.load(plane, stride, w, h, x0, y0, aligned4, lane);
  return st.store(dst, pitch, lane);
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

template <bool DOT>
__global__ void __launch_bounds__(256) predict_inter_mb_kernel(const McArgs a) {
  __shared__ __align__(16) uint8_t smem[4 * kMcLdsPerWave];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // MBs are usually in raster order: an XCD takes a band of them
  const int i = xcd_band(blockIdx.x, gridDim.x) * 4 + wave;
  if (i >= a.n) return;  // whole wave exits together
  uint8_t* lh = smem + wave * kMcLdsPerWave;
  uint8_t* lt = lh + 21 * 24;
  uint8_t* chalo = lt + 21 * 16;  // U at 0, V at 13*16
  uint8_t* ctmp = chalo + 2 * 13 * 16;

  const int mb_col = (int)uniform_u32(a.mb_xy + 2 * i), mb_row = (int)uniform_u32(a.mb_xy + 2 * i + 1);
  int16_t mvr, mvc;
  uniform_i16x2(a.mv + 2 * i, mvr, mvc);
  int16_t yr = dbl16(mvr), yc = dbl16(mvc);
  int16_t ur = (int16_t)avg4(yr), uc = (int16_t)avg4(yc);
  if (a.full_pixel) {
    yr = (int16_t)(yr & ~7); yc = (int16_t)(yc & ~7);
    ur = (int16_t)(ur & ~7); uc = (int16_t)(uc & ~7);
  }
  // A whole-pixel vector is a plain copy whatever row 0 of the caller's
  // table holds (src/motion_comp.rs:2022-2028, 2044-2067).
  const bool ycopy = (yc & 7) == 0 && (yr & 7) == 0, ccopy = (uc & 7) == 0 && (ur & 7) == 0;
  const int lx0 = mb_col * 16 + (yc >> 3) - 2, ly0 = mb_row * 16 + (yr >> 3) - 2;
  const int cx0 = mb_col * 8 + (uc >> 3) - 2, cy0 = mb_row * 8 + (ur >> 3) - 2;
  uint8_t* out = a.out + (size_t)i * 384;
  if (DOT) {
    const DotTaps fh = table_dot(a, yc & 7, ycopy), fv = table_dot(a, yr & 7, ycopy);
    const DotTaps cfh = table_dot(a, uc & 7, ccopy), cfv = table_dot(a, ur & 7, ccopy);
    const bool al = a.aligned4 != 0;
    HaloStage<21, 21> hy;
    HaloStage<13, 13> hu, hv;
    hy.load(a.ref.y, a.ref.y_stride, a.ref.lw, a.ref.lh, lx0, ly0, al, lane);
    hu.load(a.ref.u, a.ref.uv_stride, a.ref.cw, a.ref.ch, cx0, cy0, al, lane);
    hv.load(a.ref.v, a.ref.uv_stride, a.ref.cw, a.ref.ch, cx0, cy0, al, lane);
    const uint32_t lsh = hy.store(lh, 24, lane);
    const uint32_t ush = hu.store(chalo, 16, lane);
    const uint32_t vsh = hv.store(chalo + 13 * 16, 16, lane);
    wave_lds_sync();
    hpass_dot<16>(lh, 24, lsh, lt, fh, lane);
    // The chroma planes in one pass each, lanes picking the plane (U for
    // lanes 0..25, V for 26..51; then U 0..15, V 16..31): the same
    // instructions with per-lane addresses and shifts, not two branches.
    {
      const bool v = lane >= 26;
      if (lane < 52)
        hpass_dot<8>(chalo + (v ? 13 * 16 : 0), 16, v ? vsh : ush, ctmp + (v ? 13 * 8 : 0), cfh, lane - (v ? 26 : 0));
    }
    wave_lds_sync();
    *reinterpret_cast<uint32_t*>(out + 4 * lane) = vpass_dot<16>(lt, fv, lane);
    if (lane < 32) {
      const bool v = lane >= 16;
      *reinterpret_cast<uint32_t*>(out + (v ? 320 : 256) + 4 * (lane & 15)) =
          vpass_dot<8>(ctmp + (v ? 13 * 8 : 0), cfv, lane & 15);
    }
    return;
  }
  Taps fh = load_taps(a.filters[yc & 7]);
  Taps fv = load_taps(a.filters[yr & 7]);
  Taps cfh = load_taps(a.filters[uc & 7]);
  Taps cfv = load_taps(a.filters[ur & 7]);
  if (ycopy) fh.ident = fv.ident = true;
  if (ccopy) cfh.ident = cfv.ident = true;

  Gather<21, 21> gy;
  Gather<13, 13> gu, gv;
  gy.load(a.ref.y, a.ref.y_stride, a.ref.lw, a.ref.lh, lx0, ly0, lane);
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
  *reinterpret_cast<uint32_t*>(out + 4 * lane) = vpass_unit<16>(lt, fv, lane);
  if (lane < 16) {
    *reinterpret_cast<uint32_t*>(out + 256 + 4 * lane) = vpass_unit<8>(ctmp, cfv, lane);
  } else if (lane < 32) {
    *reinterpret_cast<uint32_t*>(out + 320 + 4 * (lane - 16)) =
        vpass_unit<8>(ctmp + 13 * 8, cfv, lane - 16);
  }
}

// 16x16 luma prediction (six-tap) for eighth-pel MV (r8, c8) whose halo
// lies inside `win` (pitch 24, at most 3 columns of slack) at offset
// (ox, oy) plus a per-row byte shift `sh`; returns this lane's 4 samples.
__device__ __forceinline__ uint32_t luma_pred_from_window(const uint8_t* win, int ox, int oy, uint32_t sh, int r8,
                                                          int c8, uint8_t* lt, int lane) {
  const bool copy = (r8 & 7) == 0 && (c8 & 7) == 0;
  const DotTaps fh = sixtap_dot(c8 & 7, copy), fv = sixtap_dot(r8 & 7, copy);
  const uint32_t x = (uint32_t)ox + sh;
  hpass_dot<16>(win + oy * 24 + (x & ~3u), 24, x & 3u, lt, fh, lane);
  wave_lds_sync();
  const uint32_t p = vpass_dot<16>(lt, fv, lane);
  wave_lds_sync();
  return p;
}

__global__ void __launch_bounds__(256) mb_luma_sad_at_mv_kernel(const SadArgs a) {
  __shared__ __align__(16) uint8_t smem[4 * (21 * 24 + 21 * 16)];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // MBs are usually in raster order: an XCD takes a band of them
  const int i = xcd_band(blockIdx.x, gridDim.x) * 4 + wave;
  if (i >= a.n) return;
  uint8_t* lh = smem + wave * (21 * 24 + 21 * 16);
  uint8_t* lt = lh + 21 * 24;
  const int mb_col = (int)uniform_u32(a.mb_xy + 2 * i), mb_row = (int)uniform_u32(a.mb_xy + 2 * i + 1);
  int16_t mvr, mvc;
  uniform_i16x2(a.mv + 2 * i, mvr, mvc);
  const int16_t r8 = dbl16(mvr), c8 = dbl16(mvc);
  const uint32_t s = *reinterpret_cast<const uint32_t*>(a.src + (size_t)i * 256 + 4 * lane);
  const uint32_t sh = stage_halo<21, 21>(a.plane, a.stride, a.w, a.h, mb_col * 16 + (c8 >> 3) - 2,
                                         mb_row * 16 + (r8 >> 3) - 2, a.aligned4 != 0, lh, 24, lane);
  wave_lds_sync();
  const uint32_t p = luma_pred_from_window(lh, 0, 0, sh, r8, c8, lt, lane);
  const uint32_t tot = wave_sum(sad4(s, p, 0));
  if (lane == 0) a.sad[i] = tot;
}

// Centre + 8 offsets: every candidate's integer part lies within one
// pixel of the centre's, so one clamped 24x24 window at the smallest
// integer origin minus 2 holds all nine 21x21 halos (clamping depends only
// on the absolute coordinate, so a sub-window equals a direct fetch).  The
// nine candidates take three column positions and three row positions: one
// horizontal pass per column position over all 24 window rows, then one
// vertical pass per candidate.  All nine SADs are computed, then the ladder
// decides in the reference's order (strict '<', a candidate equal to the
// current best skipped).
__global__ void __launch_bounds__(256) refine_ladder_kernel(const RefineArgs a) {
  constexpr int WP = 28;  // window pitch: 24 columns at any dword shift
  constexpr int kWave = 24 * WP + 3 * 24 * 16;
  __shared__ __align__(16) uint8_t smem[4 * kWave];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // MBs are usually in raster order: an XCD takes a band of them
  const int i = xcd_band(blockIdx.x, gridDim.x) * 4 + wave;
  if (i >= a.n) return;
  uint8_t* win = smem + wave * kWave;
  uint8_t* lt = win + 24 * WP;  // three 24 x 16 horizontal-pass results
  const int mb_col = (int)uniform_u32(a.mb_xy + 2 * i), mb_row = (int)uniform_u32(a.mb_xy + 2 * i + 1);
  // Centres outside +-1023 are outside the reference's contract (it
  // debug-asserts); clamping keeps every candidate inside the staged window.
  int16_t cr16, cc16;
  uniform_i16x2(a.center + 2 * i, cr16, cc16);
  const int crow = clampi(cr16, -1023, 1023), ccol = clampi(cc16, -1023, 1023);
  const int st = a.step;
  // candidate k: row index rk(k) (0..2 -> -st, 0, +st), column index ck(k)
  const int rowv[3] = {clampi(crow - st, -1023, 1023), crow, clampi(crow + st, -1023, 1023)};
  const int colv[3] = {clampi(ccol - st, -1023, 1023), ccol, clampi(ccol + st, -1023, 1023)};
  int minr = 1 << 30, minc = 1 << 30;
#pragma unroll
  for (int t = 0; t < 3; t++) {
    minr = min(minr, (int)(dbl16((int16_t)rowv[t]) >> 3));
    minc = min(minc, (int)(dbl16((int16_t)colv[t]) >> 3));
  }
  const uint32_t s = *reinterpret_cast<const uint32_t*>(a.src + (size_t)i * 256 + 4 * lane);
  const uint32_t wsh = stage_halo<24, 24>(a.plane, a.stride, a.w, a.h, mb_col * 16 + minc - 2, mb_row * 16 + minr - 2,
                                          a.aligned4 != 0, win, WP, lane);
  wave_lds_sync();
  // horizontal passes: column position t at window column (c8 >> 3) - minc
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const int16_t c8 = dbl16((int16_t)colv[t]);
    const uint32_t x = (uint32_t)((c8 >> 3) - minc) + wsh;
    // a whole-pel column position is a copy in both passes when the row is
    // whole-pel too; the vertical pass below makes that choice per
    // candidate, so the horizontal identity is exact here either way
    const DotTaps fh = sixtap_dot(c8 & 7, false);
    hpass_dot<16, 24>(win + (x & ~3u), WP, x & 3u, lt + t * 24 * 16, fh, lane);
  }
  wave_lds_sync();
  // Per-lane SADs of the 9 candidates; each whole-wave SAD is at most
  // 64 * 1020 < 2^16, so two candidates share one wave reduction (low and
  // high halves).
  uint32_t part[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    // k = 0: centre; k = 1..8: (-st,-st), (-st,0), (-st,+st), (0,-st),
    // (0,+st), (+st,-st), (+st,0), (+st,+st)
    const int rk = k == 0 ? 1 : (k <= 3 ? 0 : (k <= 5 ? 1 : 2));
    const int ck = k == 0 ? 1 : (k <= 3 ? k - 1 : (k == 4 ? 0 : (k == 5 ? 2 : k - 6)));
    const int16_t r8 = dbl16((int16_t)rowv[rk]);
    const DotTaps fv = sixtap_dot(r8 & 7, false);
    const uint8_t* tmp = lt + ck * 24 * 16 + ((r8 >> 3) - minr) * 16;
    part[k] = sad4(s, vpass_dot<16>(tmp, fv, lane), 0);
  }
  uint32_t sads[9];
#pragma unroll
  for (int k = 0; k < 8; k += 2) {
    const uint32_t both = wave_sum_to_sgpr(sum16(part[k] | (part[k + 1] << 16)));
    sads[k] = both & 0xFFFFu;
    sads[k + 1] = both >> 16;
  }
  sads[8] = wave_sum_to_sgpr(sum16(part[8]));
  int best_r = crow, best_c = ccol;
  uint32_t best_sad = sads[0];
#pragma unroll
  for (int k = 1; k < 9; k++) {
    const int rk = k <= 3 ? 0 : (k <= 5 ? 1 : 2);
    const int ck = k <= 3 ? k - 1 : (k == 4 ? 0 : (k == 5 ? 2 : k - 6));
    const int r = rowv[rk], c = colv[ck];
    if (r == best_r && c == best_c) continue;  // collapsed onto the best
    if (sads[k] < best_sad) {
      best_sad = sads[k];
      best_r = r;
      best_c = c;
    }
  }
  if (lane == 0) {
    a.out[2 * i] = ((uint32_t)best_r & 0xFFFFu) | ((uint32_t)best_c << 16);
    a.out[2 * i + 1] = best_sad;
  }
}

// Every row either the identity {0, 0, 128, 0, 0, 0} or taps in [-128, 127]
// summing to 128 (the dot-product form's precondition).
static bool dot_table_ok(const int32_t f[8][6]) {
  for (int r = 0; r < 8; r++) {
    const int32_t* t = f[r];
    if (t[0] == 0 && t[1] == 0 && t[2] == 128 && t[3] == 0 && t[4] == 0 && t[5] == 0) continue;
    int64_t sum = 0;
    for (int k = 0; k < 6; k++) {
      if (t[k] < -128 || t[k] > 127) return false;
      sum += t[k];
    }
    if (sum != 128) return false;
  }
  return true;
}

static bool aligned4(const void* p, uint32_t stride) { return (((uintptr_t)p | stride) & 3) == 0; }

hipError_t launch_predict_inter_mb(const McArgs& args, hipStream_t stream) {
  if (args.n <= 0) return hipSuccess;
  McArgs a = args;
  a.aligned4 = aligned4(a.ref.y, a.ref.y_stride) && aligned4(a.ref.u, a.ref.uv_stride) &&
               aligned4(a.ref.v, a.ref.uv_stride);
  a.dot = dot_table_ok(a.filters);
  a.ident_rows = 0;
  for (int r = 0; r < 8; r++) {
    a.dot_lo[r] = pack_lo(a.filters[r]);
    a.dot_hi[r] = pack_hi(a.filters[r]);
    if (is_identity(a.filters[r])) a.ident_rows |= 1u << r;
  }
  if (a.dot)
    hipLaunchKernelGGL(predict_inter_mb_kernel<true>, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL(predict_inter_mb_kernel<false>, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_mb_luma_sad_at_mv(const SadArgs& args, hipStream_t stream) {
  if (args.n <= 0) return hipSuccess;
  SadArgs a = args;
  a.aligned4 = aligned4(a.plane, a.stride);
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
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // MBs are usually in raster order: an XCD takes a band of them
  const int i = xcd_band(blockIdx.x, gridDim.x) * 4 + wave;
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

hipError_t launch_refine_ladder(const RefineArgs& args, hipStream_t stream) {
  if (args.n <= 0) return hipSuccess;
  RefineArgs a = args;
  a.aligned4 = aligned4(a.plane, a.stride);
  hipLaunchKernelGGL(refine_ladder_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace vp8g
