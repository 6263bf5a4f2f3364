// RFC 6386 section 15 loop filter over whole frames (src/loop_filter.rs).
//
// The filter is sequential by definition: macroblocks in raster order and,
// inside one, left edge, inner vertical edges, top edge, inner horizontal
// edges (src/loop_filter.rs:1393-1494).  MB (r, c) depends on (r, c-1),
// (r-1, c) and (r-1, c+1), so rows may run concurrently with row r two MBs
// behind row r-1 (src/loop_filter/bands.rs:1-35).
//
// One workgroup filters one instance: a frame, or a band of MB rows scored
// as a frame of its own, with one (level, sharpness) candidate.  Each wave
// owns a pair of MB rows: lanes 0-31 run row 2p at column c while lanes
// 32-63 run row 2p+1 at column c-2, so one wave covers both sides of the row
// boundary inside the pair.  Waves take pairs p = wave, wave + NW, ...
//
// Lane roles inside a half-wave (32 lanes): lanes 0-15 are the 16 luma rows
// (vertical edges) or columns (horizontal edges) of the MB, lanes 16-23 the
// 8 U rows/columns, lanes 24-31 the V ones.
//
// Data of one half-wave lives in an LDS strip per plane: n + 4 rows (the 4
// rows above the MB, then the MB's n) by S = 2n bytes, plane column x at
// byte x mod S.  The 4 columns left of the MB (modified by its left edge,
// final only afterwards) are therefore still in the strip from the previous
// step.  A step:
//   1. the 4 rows above the MB come from the row above's ring slot c;
//   2. row lanes load the MB's n pristine bytes of their row, take the 4
//      carried bytes on the left from the strip, run the vertical edges in
//      registers and write the row back;
//   3. the row's ring slot c-1 (its bottom 4 rows at columns of MB c-1,
//      final once this MB's left edge ran) is published;
//   4. column lanes read their n + 4 bytes down the strip, run the
//      horizontal edges and write them back;
//   5. the pixels that no later MB touches are written out (or scored):
//      strip rows 1..n by columns x0-4 .. x0+n-5, plus the right and bottom
//      margins at the last column / row of the instance.
// The row above's ring is in LDS for the boundary inside a pair and in
// global memory (one slot per MB column, so no producer ever waits for a
// consumer) between pairs.  Pair p publishes progress p*65536 + steps done
// in LDS with workgroup-scope release; pair p+1 acquires it before reading
// the ring.  Every wait is on the previous pair only, so the waits form a
// chain ending at pair 0, which never waits: the schedule cannot deadlock.
// A wait still gives up after a bounded number of polls (flagging the
// instance) so that a broken invariant cannot hang the device.
//
// Pixels are written in place: a step writes only pixels no later step
// reads from the frame (every later read of them comes from a strip or a
// ring), so the output may alias the reconstruction.
#include "vp8g_internal.h"

namespace vp8g {
namespace {

constexpr int kRingA = 4;            // LDS ring slots inside a pair (lag is 2)
constexpr int kSlotBytes = 128;      // luma 4x16 + two chroma 4x8
constexpr int kStripBytes = 640 + 2 * 192;  // luma 20 x 32, chroma 12 x 16 (per half-wave)
constexpr int kWaveLds = 2 * kStripBytes + kRingA * kSlotBytes;

__device__ __forceinline__ int c8(int v) { return min(max(v, -128), 127); }

__device__ __forceinline__ uint32_t ad(uint32_t a, uint32_t b) { return __builtin_amdgcn_sad_u16(a, b, 0u); }
__device__ __forceinline__ uint32_t px(int v) { return (uint32_t)min(max(v, 0), 255); }

// Section 15.3 MB-edge (MB) or sub-block-edge filter on the 8 taps
// p3 p2 p1 p0 | q0 q1 q2 q3 = bytes of lo then hi (src/loop_filter.rs:292-393),
// both outcomes computed and selected.  In unsigned pixels: the signed
// differences of the section 15.2 arithmetic are the unsigned ones, and
// converting u2s(p) + a back to unsigned (c8, then + 128) is
// clamp(p + a, 0, 255).
template <bool MB>
__device__ __forceinline__ void normal_edge(uint32_t& lo, uint32_t& hi, int I, int E, int T, bool on) {
  const uint32_t p3 = lo & 255, p2 = (lo >> 8) & 255, p1 = (lo >> 16) & 255, p0 = lo >> 24;
  const uint32_t q0 = hi & 255, q1 = (hi >> 8) & 255, q2 = (hi >> 16) & 255, q3 = hi >> 24;
  const uint32_t dp = ad(p1, p0), dq = ad(q1, q0);
  const uint32_t m = max(max(max(ad(p3, p2), ad(p2, p1)), dp), max(max(ad(q3, q2), ad(q2, q1)), dq));
  const bool yes = on && ad(p0, q0) * 2 + (ad(p1, q1) >> 1) <= (uint32_t)E && m <= (uint32_t)I;
  const bool hv = max(dp, dq) > (uint32_t)T;
  const int P0 = p0, P1 = p1, P2 = p2, Q0 = q0, Q1 = q1, Q2 = q2;
  const int outer = c8(P1 - Q1), step = 3 * (Q0 - P0);
  uint32_t np0, nq0, np1, nq1, np2 = p2, nq2 = q2;
  if (MB) {
    const int w = c8(outer + step);
    const int a0 = c8((27 * w + 63) >> 7), a1 = c8((18 * w + 63) >> 7), a2 = c8((9 * w + 63) >> 7);
    const int fa = c8(w + 4) >> 3, fb = c8(w + 3) >> 3;
    np0 = px(P0 + (hv ? fb : a0));
    nq0 = px(Q0 - (hv ? fa : a0));
    np1 = hv ? p1 : px(P1 + a1);
    nq1 = hv ? q1 : px(Q1 - a1);
    np2 = hv ? p2 : px(P2 + a2);
    nq2 = hv ? q2 : px(Q2 - a2);
  } else {
    const int a = c8((hv ? outer : 0) + step);
    const int f1 = c8(a + 4) >> 3, f2 = c8(a + 3) >> 3, h = (f1 + 1) >> 1;
    np0 = px(P0 + f2);
    nq0 = px(Q0 - f1);
    np1 = hv ? p1 : px(P1 + h);
    nq1 = hv ? q1 : px(Q1 - h);
  }
  const uint32_t nlo = p3 | np2 << 8 | np1 << 16 | np0 << 24, nhi = nq0 | nq1 << 8 | nq2 << 16 | q3 << 24;
  lo = yes ? nlo : lo;
  hi = yes ? nhi : hi;
}

// Section 15.2 simple filter on p1 p0 | q0 q1 = the top two bytes of lo and
// the bottom two of hi (src/loop_filter.rs:200-245).
__device__ __forceinline__ void simple_edge(uint32_t& lo, uint32_t& hi, int E, bool on) {
  const uint32_t p1 = (lo >> 16) & 255, p0 = lo >> 24, q0 = hi & 255, q1 = (hi >> 8) & 255;
  const bool yes = on && ad(p0, q0) * 2 + (ad(p1, q1) >> 1) <= (uint32_t)E;
  const int a = c8(c8((int)p1 - (int)q1) + 3 * ((int)q0 - (int)p0));
  const uint32_t np0 = px((int)p0 + (c8(a + 3) >> 3)), nq0 = px((int)q0 - (c8(a + 4) >> 3));
  lo = yes ? (lo & 0x00FFFFFFu) | np0 << 24 : lo;
  hi = yes ? (hi & 0xFFFFFF00u) | nq0 : hi;
}

// calculate_mb_filter_level_inter, src/loop_filter.rs:1280-1334 (with an
// intra MB, calculate_mb_filter_level, src/loop_filter.rs:1222-1265).
__device__ __forceinline__ int mb_level(const LfArgs& a, int base, uint32_t mb) {
  const int seg = mb & 255, ref = (mb >> 8) & 255, mode = (mb >> 16) & 255, flags = mb >> 24;
  int level = base;
  if (a.segmentation_enabled) {
    const int s = seg == 255 ? 0 : min(seg, 3);
    level = a.segment_abs ? a.segment_lf_level[s] : level + a.segment_lf_level[s];
  }
  level = min(max(level, 0), 63);
  if (a.delta_enabled) {
    level += a.ref_delta[min(ref, 3)];
    if (ref == 0)
      level += (flags & 1) ? a.mode_delta[0] : 0;
    else
      level += a.mode_delta[mode == 1 ? 1 : (mode == 2 ? 3 : 2)];
  }
  return min(max(level, 0), 63);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t ssd4(uint32_t a, uint32_t b, uint32_t keep) {
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int d = (int)((a >> (8 * j)) & 255) - (int)((b >> (8 * j)) & 255);
    s += (keep >> j & 1) ? (uint32_t)(d * d) : 0u;
  }
  return s;
}

// Output pass q of a lane (one dword of the final pixels of a step):
// q = 0, 1: luma strip rows 1..16 by dwords 0..3; q = 2: chroma strip rows
// 1..8 by dwords 0..1 of U (lanes 0-15) or V (lanes 16-31).
__device__ __forceinline__ void out_entry(int li, int q, int& epl, int& row, int& dw) {
  if (q < 2) {
    epl = 0, row = 1 + (li >> 2) + 8 * q, dw = li & 3;
  } else {
    epl = li < 16 ? 1 : 2, row = 1 + ((li & 15) >> 1), dw = li & 1;
  }
}

__device__ __forceinline__ int strip_off(int pl) { return pl == 0 ? 0 : (pl == 1 ? 640 : 832); }

// Plane base pointers as uniform values: the kernels address planes as
// base + 32-bit offset (a plane is below 4 GiB), with one instruction per
// plane under lane masks so that the base stays in scalar registers.
__device__ __forceinline__ const uint8_t* src_plane(const LfArgs& a, int pl) { return pl == 0 ? a.sy : (pl == 1 ? a.su : a.sv); }

// The source dword at plane (x, y) for scoring, 0 outside the visible
// window; bytes past its right edge are not read.
__device__ __forceinline__ uint32_t src_word(const LfArgs& a, int pl, int x, int y) {
  const int vw = pl == 0 ? a.vis_w : a.vis_cw, vh = pl == 0 ? a.vis_h : a.vis_ch;
  if (y >= vh || x >= vw) return 0;
  const uint32_t off = (uint32_t)y * (uint32_t)(pl == 0 ? a.sy_stride : a.suv_stride) + (uint32_t)x;
  const bool whole = x + 4 <= vw && a.src_aligned4;
  uint32_t s = 0;
  if (whole) {
    if (pl == 0) s = *reinterpret_cast<const uint32_t*>(a.sy + off);
    if (pl == 1) s = *reinterpret_cast<const uint32_t*>(a.su + off);
    if (pl == 2) s = *reinterpret_cast<const uint32_t*>(a.sv + off);
  } else {
    const uint8_t* sp = src_plane(a, pl) + off;
    for (int j = 0; j < 4 && x + j < vw; j++) s |= (uint32_t)sp[j] << (8 * j);
  }
  return s;
}

// Squared error of the 4 bytes of v against the source word, inside the
// visible window.
__device__ __forceinline__ uint32_t src_ssd(const LfArgs& a, int pl, int x, int y, uint32_t v, uint32_t sw) {
  const int vw = pl == 0 ? a.vis_w : a.vis_cw, vh = pl == 0 ? a.vis_h : a.vis_ch;
  if (y >= vh || x >= vw) return 0;
  return ssd4(v, sw, x + 4 <= vw ? 0xFu : ((1u << (vw - x)) - 1u));
}

__device__ __forceinline__ void store_out(const LfArgs& a, int pl, int x, int y, uint32_t v) {
  const uint32_t off = (uint32_t)y * (uint32_t)(pl == 0 ? a.y_stride : a.uv_stride) + (uint32_t)x;
  if (pl == 0) *reinterpret_cast<uint32_t*>(a.oy + off) = v;
  if (pl == 1) *reinterpret_cast<uint32_t*>(a.ou + off) = v;
  if (pl == 2) *reinterpret_cast<uint32_t*>(a.ov + off) = v;
}

// One final dword straight from the strip (the rare right / bottom margins).
template <bool SCORE>
__device__ __forceinline__ void emit(const LfArgs& a, const uint8_t* hs, int pl, int row, int dw, int x0, int y0,
                                     uint32_t& acc) {
  const int S = pl == 0 ? 32 : 16;
  const int x = x0 - 4 + 4 * dw, y = y0 - 4 + row;
  const uint32_t v = *reinterpret_cast<const uint32_t*>(hs + strip_off(pl) + row * S + (x & (S - 1)));
  if (!SCORE)
    store_out(a, pl, x, y, v);
  else
    acc += src_ssd(a, pl, x, y, v, src_word(a, pl, x, y));
}

// The n pristine bytes of one MB row of plane pl at offset `off` (chroma
// in .x .y).
__device__ __forceinline__ uint4 load_mb_row(const LfArgs& a, int pl, uint32_t off) {
  uint4 m = make_uint4(0, 0, 0, 0);
  if (pl == 0) m = *reinterpret_cast<const uint4*>(a.y + off);
  if (pl == 1) {
    const uint2 v = *reinterpret_cast<const uint2*>(a.u + off);
    m.x = v.x, m.y = v.y;
  }
  if (pl == 2) {
    const uint2 v = *reinterpret_cast<const uint2*>(a.v + off);
    m.x = v.x, m.y = v.y;
  }
  return m;
}

template <bool SIMPLE, bool SCORE>
__global__ void __launch_bounds__(64 * kLfMaxWaves) lf_kernel(LfArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int nw = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int half = lane >> 5, li = lane & 31;
  const int pl = li < 16 ? 0 : (li < 24 ? 1 : 2);
  const int k = pl == 0 ? li : (li & 7);      // row / column of the MB this lane handles
  const int n = pl == 0 ? 16 : 8, S = 2 * n;
  const int C = a.mb_cols;

  // instance: candidate x band
  const int cand = blockIdx.x / a.n_bands, band = blockIdx.x - cand * a.n_bands;
  const int row0 = band * a.every_rows;
  const int rows = min(a.band_rows, a.mb_rows - row0);
  const int base_level = a.cand_level[cand], sharp = a.cand_sharp[cand];
  const bool skip = SCORE && base_level == 0;  // level 0 scores the unfiltered picture
  const int P = (rows + 1) >> 1;
  // interior limit of a level (src/loop_filter.rs:99-112) depends on it and the sharpness only
  const int il_shift = sharp ? (sharp > 4 ? 2 : 1) : 0, il_cap = sharp ? (sharp >= 9 ? 0 : 9 - sharp) : 255;

  int* prog = reinterpret_cast<int*>(lds + nw * kWaveLds);
  int* abort_flag = prog + (nw + 1);
  unsigned long long* wsum =
      reinterpret_cast<unsigned long long*>(lds + nw * kWaveLds + ((nw + 2) * 4 + 7) / 8 * 8);
  if (threadIdx.x <= (unsigned)nw) prog[threadIdx.x] = 0;
  if (threadIdx.x == 0) *abort_flag = 0;
  __syncthreads();

  uint8_t* wl = lds + wave * kWaveLds;
  uint8_t* hs = wl + half * kStripBytes;
  uint8_t* strip = hs + strip_off(pl);
  uint8_t* ringa = wl + 2 * kStripBytes;
  uint8_t* ringb = a.ring + (size_t)blockIdx.x * a.ring_instance_bytes;
  const uint32_t stride = (uint32_t)(pl == 0 ? a.y_stride : a.uv_stride);
  // ring dword li <-> (plane, row m in 0..3, dword j) of a slot
  const int rm = pl == 0 ? li >> 2 : ((li & 7) >> 1), rj = pl == 0 ? (li & 3) : (li & 1);
  // the strip row read down a column, clamped into this plane's strip
  const int hrow_last = n + 3;

  unsigned long long total = 0;
  bool aborted = false;
  for (int p = wave; p < P && !aborted; p += nw) {
    const int r = 2 * p + half;              // instance-relative MB row of this half
    const bool row_ok = r < rows;
    const bool has_below = r + 1 < rows;     // this row's ring has a consumer
    const int y0 = (row0 + r) * n;
    const uint32_t* mbrow = a.mbs + (size_t)(row0 + r) * C;
    const uint32_t grow = (uint32_t)(y0 + k) * stride;  // offset of this lane's pixel row
    uint32_t acc = 0;
    // this half's current MB: pristine row k and the MB record, one column ahead
    uint4 cur = make_uint4(0, 0, 0, 0), nxt = cur;
    uint32_t mbc = 0, mbn = 0;
    if (row_ok && half == 0) {
      cur = load_mb_row(a, pl, grow);
      mbc = mbrow[0];
    }
    for (int s = 0; s < C + 2; s++) {
      // -- wait for the row above when it lives in another wave
      if (p > 0 && s < C) {
        const int need = (p - 1) * 65536 + min(s + 4, C + 2);
        int* slot = prog + (p - 1) % (nw + 1);
        int spins = 0;
        while (__hip_atomic_load(slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need) {
          if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) ||
              ++spins > a.spin_limit) {
            aborted = true;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        if (aborted) {
          __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (lane == 0) *a.err = 1;
          break;
        }
      }
      const int c = s - 2 * half;
      const bool act = row_ok && c >= 0 && c < C;
      const int x0 = c * n;
      wave_sync();

      // -- A. loads for later in this step and for the next one
      uint32_t ringv = 0;
      if (act && r > 0 && half == 0)  // the row above is the previous pair's lower half
        ringv = *reinterpret_cast<const uint32_t*>(ringb + (uint32_t)((((p - 1) % (nw + 1)) * C + c) * kSlotBytes +
                                                                   4 * li));
      if (row_ok && c + 1 >= 0 && c + 1 < C) {
        nxt = load_mb_row(a, pl, grow + (uint32_t)(x0 + n));
        mbn = mbrow[c + 1];
      }
      uint32_t sw[3] = {0, 0, 0};
      if (SCORE && act) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
          int epl, row, dw;
          out_entry(li, q, epl, row, dw);
          const int en = epl == 0 ? 16 : 8;
          if ((row >= 4 || r > 0) && (dw > 0 || c > 0))
            sw[q] = src_word(a, epl, c * en - 4 + 4 * dw, (row0 + r) * en - 4 + row);
        }
      }

      // -- B. this half's MB: level and limits
      const int level = act && !skip ? mb_level(a, base_level, mbc) : 0;
      const bool sub = (mbc >> 24) & 3;
      const int il = max(min(level >> il_shift, il_cap), 1);
      const int hev = a.key_frame ? (level >= 40 ? 2 : (level >= 15 ? 1 : 0))
                                  : (level >= 40 ? 3 : (level >= 20 ? 2 : (level >= 15 ? 1 : 0)));
      const int mbe = min((level + 2) * 2 + il, 255), sbe = min(level * 2 + il, 255);
      const bool lanes_on = level > 0 && (!SIMPLE || pl == 0);

      // -- C. vertical edges along this lane's row (row k of the MB)
      if (act) {
        uint32_t d[5];
        uint8_t* rowp = strip + (4 + k) * S;
        d[0] = *reinterpret_cast<const uint32_t*>(rowp + ((x0 - 4) & (S - 1)));
        d[1] = cur.x, d[2] = cur.y, d[3] = cur.z, d[4] = cur.w;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const bool on = lanes_on && (e == 0 ? c > 0 : sub) && (e < n / 4);
          if (SIMPLE)
            simple_edge(d[e], d[e + 1], e == 0 ? mbe : sbe, on);
          else if (e == 0)
            normal_edge<true>(d[0], d[1], il, mbe, hev, on);
          else
            normal_edge<false>(d[e], d[e + 1], il, sbe, hev, on);
        }
        *reinterpret_cast<uint32_t*>(rowp + ((x0 - 4) & (S - 1))) = d[0];
        if (pl == 0)
          *reinterpret_cast<uint4*>(rowp + (x0 & (S - 1))) = make_uint4(d[1], d[2], d[3], d[4]);
        else
          *reinterpret_cast<uint2*>(rowp + (x0 & (S - 1))) = make_uint2(d[1], d[2]);
        // -- D. the 4 rows above from the ring of the row above
        if (r > 0) {
          if (half)
            ringv = *reinterpret_cast<const uint32_t*>(ringa + (c % kRingA) * kSlotBytes + 4 * li);
          *reinterpret_cast<uint32_t*>(strip + rm * S + ((x0 + 4 * rj) & (S - 1))) = ringv;
        }
      }
      wave_sync();

      // -- E. publish this row's ring slot c-1 (bottom 4 rows, MB c-1)
      if (act && has_below && c > 0) {
        const uint32_t v =
            *reinterpret_cast<const uint32_t*>(strip + (n + rm) * S + ((x0 - n + 4 * rj) & (S - 1)));
        if (half == 0)
          *reinterpret_cast<uint32_t*>(ringa + ((c - 1) % kRingA) * kSlotBytes + 4 * li) = v;
        else
          *reinterpret_cast<uint32_t*>(ringb + (uint32_t)(((p % (nw + 1)) * C + (c - 1)) * kSlotBytes + 4 * li)) = v;
      }

      // -- F. horizontal edges down this lane's column (column k of the MB)
      if (act) {
        uint8_t* colp = strip + ((x0 + k) & (S - 1));
        uint32_t d[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 20; i++) d[i >> 2] |= (uint32_t)colp[min(i, hrow_last) * S] << (8 * (i & 3));
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const bool on = lanes_on && (e == 0 ? r > 0 : sub) && (e < n / 4);
          if (SIMPLE)
            simple_edge(d[e], d[e + 1], e == 0 ? mbe : sbe, on);
          else if (e == 0)
            normal_edge<true>(d[0], d[1], il, mbe, hev, on);
          else
            normal_edge<false>(d[e], d[e + 1], il, sbe, hev, on);
        }
        // rows 1..n+1 can change; chroma lanes repeat their row 9 past it
#pragma unroll
        for (int i = 1; i < 18; i++) {
          const uint8_t vb = (i < 10 || pl == 0) ? (uint8_t)(d[i >> 2] >> (8 * (i & 3))) : (uint8_t)(d[2] >> 8);
          colp[((i < 10 || pl == 0) ? i : 9) * S] = vb;
        }
      }
      wave_sync();

      // -- G. the last column's ring slot is final now
      const bool last_col = c == C - 1;
      if (act && last_col && has_below) {
        const uint32_t v = *reinterpret_cast<const uint32_t*>(strip + (n + rm) * S + ((x0 + 4 * rj) & (S - 1)));
        if (half == 0)
          *reinterpret_cast<uint32_t*>(ringa + (c % kRingA) * kSlotBytes + 4 * li) = v;
        else
          *reinterpret_cast<uint32_t*>(ringb + (uint32_t)(((p % (nw + 1)) * C + c) * kSlotBytes + 4 * li)) = v;
      }
      // -- H. progress of this pair (every lane stores the same value)
      __hip_atomic_store(prog + p % (nw + 1), p * 65536 + s + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);

      // -- I. final pixels: strip rows 1..n by dwords 0..n/4-1, plus the
      //    right / bottom margins at the instance's last column / row
      if (act) {
        const bool last_row = r == rows - 1;
#pragma unroll
        for (int q = 0; q < 3; q++) {
          int epl, row, dw;
          out_entry(li, q, epl, row, dw);
          if ((row >= 4 || r > 0) && (dw > 0 || c > 0)) {
            const int en = epl == 0 ? 16 : 8, Se = 2 * en;
            const int x = c * en - 4 + 4 * dw, y = (row0 + r) * en - 4 + row;
            const uint32_t v = *reinterpret_cast<const uint32_t*>(hs + strip_off(epl) + row * Se + (x & (Se - 1)));
            if (!SCORE) {
              store_out(a, epl, x, y, v);
            } else {
              acc += src_ssd(a, epl, x, y, v, sw[q]);
            }
          }
        }
        if (last_col) {  // right margin: dword n/4 of rows 1..n
          const int row = 1 + k;
          if (row >= 4 || r > 0) emit<SCORE>(a, hs, pl, row, n / 4, x0, y0, acc);
        }
        if (last_row) {  // bottom margin: rows n+1..n+3
          // luma 3 rows x 4 dwords (lanes 0-11), chroma 3 x 2 per plane (12-17, 18-23)
          if (li < 24) {
            const int epl = li < 12 ? 0 : (li < 18 ? 1 : 2);
            const int e = li < 12 ? li : (li < 18 ? li - 12 : li - 18);
            const int en = epl == 0 ? 16 : 8, nd = en / 4;
            const int row = en + 1 + e / nd, dw = e % nd;
            if (dw > 0 || c > 0) emit<SCORE>(a, hs, epl, row, dw, c * en, (row0 + r) * en, acc);
          }
          if (last_col && li < 9) {  // bottom-right corner dword
            const int epl = li / 3, en = epl == 0 ? 16 : 8;
            emit<SCORE>(a, hs, epl, en + 1 + li % 3, en / 4, c * en, (row0 + r) * en, acc);
          }
        }
      }
      cur = nxt;
      mbc = mbn;
    }
    total += acc;
  }

  if (SCORE) {
    for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o);
    if (lane == 0) wsum[wave] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int w = 0; w < nw; w++) t += wsum[w];
      a.inst_ssd[blockIdx.x] = t;
    }
  }
}

// Per-candidate SSD: the sum of its bands' instance SSDs.
__global__ void lf_sum_bands_kernel(const unsigned long long* inst, int n_cands, int n_bands,
                                    unsigned long long* out) {
  const int cnd = blockIdx.x * blockDim.x + threadIdx.x;
  if (cnd >= n_cands) return;
  unsigned long long t = 0;
  for (int b = 0; b < n_bands; b++) t += inst[(size_t)cnd * n_bands + b];
  out[cnd] = t;
}

}  // namespace

size_t lf_lds_bytes(int waves) {
  return (size_t)waves * kWaveLds + ((size_t)(waves + 2) * 4 + 7) / 8 * 8 + 8 * (size_t)waves;
}
size_t lf_ring_bytes(int waves, int mb_cols) { return (size_t)(waves + 1) * (size_t)mb_cols * kSlotBytes; }

hipError_t launch_loop_filter(const LfArgs& a, int instances, int waves, bool simple, bool score, hipStream_t stream) {
  const size_t lds = lf_lds_bytes(waves);
  const dim3 grid(instances), block(64 * waves);
  if (simple) {
    if (score)
      hipLaunchKernelGGL((lf_kernel<true, true>), grid, block, lds, stream, a);
    else
      hipLaunchKernelGGL((lf_kernel<true, false>), grid, block, lds, stream, a);
  } else {
    if (score)
      hipLaunchKernelGGL((lf_kernel<false, true>), grid, block, lds, stream, a);
    else
      hipLaunchKernelGGL((lf_kernel<false, false>), grid, block, lds, stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_lf_sum_bands(const unsigned long long* inst, int n_cands, int n_bands, unsigned long long* out,
                               hipStream_t stream) {
  hipLaunchKernelGGL(lf_sum_bands_kernel, dim3((n_cands + 63) / 64), dim3(64), 0, stream, inst, n_cands, n_bands,
                     out);
  return hipGetLastError();
}

}  // namespace vp8g
