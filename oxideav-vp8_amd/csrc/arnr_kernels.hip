// ARNR altref synthesis on gfx950, search + occlusion guard + blend + resolve
// fused into one launch.
//
// Work decomposition: one wave per anchor macroblock (times K frame
// splits), M horizontally adjacent MBs of one MB row per workgroup.  A wave
// loops over its share of the neighbour frames; for each it stages the
// frame's clamped luma and chroma window around the MB (+-15 px luma, +-7 px
// chroma) into its own LDS slice, runs the block search there and adds the
// blended pixels into per-lane register accumulators.  The next frame's
// window is fetched into registers while the current one is searched.  With
// K > 1 the split waves of one MB meet once, in LDS, before the resolve.
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
//     are packed into one u32 (acc < 2^20, wsum < 2^12 for <= 64 frames):
//     integer addition is exact and commutative, so neither the frame order
//     of a wave nor the split of frames over waves changes the result.
//   * resolve: (acc + ws/2) / ws with the centre's weight-16 seed
//     (src/arnr.rs:662-675).
// Chroma output clamping (src/arnr.rs:442-444) never fires from the block
// grid: cbx + ceil(bw/2) = ceil(min(bx+16, w)/2) <= ceil(w/2).
#include "vp8g_device.h"
#include <type_traits>
#include "vp8g_internal.h"

namespace vp8g {

constexpr int kWinRows = 46;   // by-15 .. by+30
constexpr int kAccShift = 20;  // packed accumulator: wsum << 20 | acc
constexpr uint32_t kAccOne = 1u << kAccShift;

// One wave's LDS windows: luma rows by-15..by+30, cols bx-16..bx+31 (three
// aligned 16-byte pieces per row) at pitch 48, twice: the frame being
// searched and the next one, which lands by DMA meanwhile.  A search lane
// reads rows kk and kk + 8; at 12 dwords per row the 8 lanes of a probe
// group then read 8 distinct banks.  Chroma is not staged: the blend reads
// one U and one V byte per lane at the chosen vector straight from memory.
constexpr int kLumaPitch = 48;
constexpr int kLumaWin = kWinRows * kLumaPitch;  // 2208
constexpr int kWaveWinBytes = 2 * kLumaWin;      // 4416
constexpr int kLumaPieces = kWinRows * 3;        // 138 16-byte pieces

struct MbGeom {
  static constexpr int lp = kLumaPitch;
};

__device__ __forceinline__ uint32_t load_clamped4(const uint8_t* row, int x0, int w, bool aligned4) {
  if (aligned4 && x0 >= 0 && x0 + 4 <= w) return *reinterpret_cast<const uint32_t*>(row + x0);
  uint32_t v = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) v |= (uint32_t)row[clampi(x0 + j, 0, w - 1)] << (8 * j);
  return v;
}

using LdsBytes = __attribute__((address_space(3))) const uint8_t*;
using LdsWords = __attribute__((address_space(3))) const uint32_t*;

__device__ __forceinline__ void lds_add(uint32_t* p, uint32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Probe g of a step sits at centre + s*(dx_g, dy_g) with the reference's
// order (-1,-1),(0,-1),(1,-1),(-1,0),(1,0),(-1,1),(0,1),(1,1)
// (src/arnr.rs:274-283); (d + 1) is packed 2 bits per probe.
constexpr uint32_t kProbeDx1 = (0u << 0) | (1u << 2) | (2u << 4) | (0u << 6) | (2u << 8) | (0u << 10) |
                               (1u << 12) | (2u << 14);
constexpr uint32_t kProbeDy1 = (0u << 0) | (0u << 2) | (0u << 4) | (1u << 6) | (1u << 8) | (2u << 10) |
                               (2u << 12) | (2u << 14);

// Window reads and LDS banks.  A read instruction is served 32 lanes at a
// time by 32 banks of one dword; lanes of one 32-lane half that hit one bank
// at different dwords cost a cycle each.  The window is built from 16-byte
// pieces, so a dword's bank is its column dword mod 4 plus 4 times a row
// term, and 8 lanes of a probe (8 rows) cover all 8 row terms of one bank
// class: two probes of one half conflict whenever their dword columns agree
// mod 4.  Lane group j (lanes 8j .. 8j + 7) holds probe (kLaneProbe >> 4j)
// & 7, which puts probes 0, 1, 2, 6 in the first half and 3, 4, 5, 7 in the
// second, and probes 0..4 read their two dword pairs in the order (2, 3),
// (0, 1) (the source dwords are held in that order too).  With that, steps 2
// and 1 read without conflicts and step 4 with a quarter of them, against
// 20, 20 and 16 cycles per block for plain lane order (a 32-bank model,
// which also reproduces the conflict counter of the plain order).
constexpr uint32_t kLaneProbe = 0x75436210u;
__device__ __forceinline__ int lane_probe(int lane) { return (kLaneProbe >> (4 * (lane >> 3))) & 7; }

// Frame index of neighbour k (0 .. n - 2), skipping the centre, as plain
// integer arithmetic (a select here becomes vector code).
__device__ __forceinline__ int skip_centre(int k, int centre) { return k + (int)((uint32_t)(centre - 1 - k) >> 31); }
__device__ __forceinline__ int probe_pair_order(int probe) { return probe < 5 ? 1 : 0; }

// refine_search (src/arnr.rs:264-300) for one block against the staged
// window.  Lanes are 8 probe groups x 8 row pairs (rows kk and kk + 8);
// `win` is the window's LDS address (wave-uniform) and `L` the lane's
// offsets in it; `rh` = 1 when the lane reads its dword pairs swapped
// (src0 / src1 are held in the same order).  The probes of
// a step never leave +-15 (8+4+2+1), so the reference's range check never
// fires and is not evaluated.  MASK: which of columns >= bw (kMaskCols) and
// rows >= bh (kMaskRows) a partial block masks out of every SAD (the centre
// strip holds zeros there; the reference bytes are masked to zero, rows on
// the lane's row sums).  WAVE_ZERO (16x16 only): the (0, 0) SAD as one
// wave-wide sum over the window at `L.blend` / `zsrc` (row lane / 4, dword
// lane % 4) instead of the probe layout's eight group sums.
enum { kMaskNone = 0, kMaskRows = 1, kMaskCols = 2, kMaskBoth = 3 };

// A lane's window byte offsets, computed once per wave and held opaque, so
// that each probe address is one add of a wave-uniform term (window base
// plus the best vector's offset) to a lane constant:
//   at[s]  the lane's first row (kk) at the probe of step 2^s, pair A
//          ((kk + 15) lp + 16 + 8 rh + gd_off 2^s); at0: the (0, 0) probe;
//   dab    pair B's offset from pair A (8 - 16 rh);
//   blend  the lane's blend dword: row lane / 4, dword lane % 4 of the
//          block at (0, 0) ((lane / 4 + 15) lp + 16 + 4 (lane % 4)).
struct SearchLanes {
  uint32_t at8, at4, at2, at1, at0, dab, blend;
};

__device__ __forceinline__ LdsWords lds_words(uint32_t addr) { return (LdsWords)(uintptr_t)addr; }
__device__ __forceinline__ uint32_t lds_addr(const void* p) { return (uint32_t)(uintptr_t)(LdsBytes)p; }

template <int MASK, bool WAVE_ZERO, class G>
__device__ __forceinline__ void descend(uint32_t win, const SearchLanes& L, const uint4 src0, const uint4 src1,
                                        int gd_off, int gq, int kk, int rh, int bw, int bh, int& bestx,
                                        int& besty, uint32_t& best_sad, uint32_t zsrc) {
  static_assert((G::lp & 3) == 0, "window rows are dword aligned");
  constexpr bool kCols = (MASK & kMaskCols) != 0, kRows = (MASK & kMaskRows) != 0;
  uint32_t colmask[4] = {~0u, ~0u, ~0u, ~0u};
  uint32_t rowmask0 = ~0u, rowmask1 = ~0u;
  if (kCols) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int left = bw - 4 * (q ^ (2 * rh));  // the lane's q-th dword is column dword q ^ 2 rh
      colmask[q] = left >= 4 ? 0xFFFFFFFFu : (left <= 0 ? 0u : (0xFFFFFFFFu >> (32 - 8 * left)));
    }
  }
  if (kRows) {
    // row kk + 8 lies inside the block only if row kk does
    rowmask0 = kk < bh ? ~0u : 0u;
    rowmask1 = kk + 8 < bh ? ~0u : 0u;
  }
  // The lane's two rows at byte offset `off`, whose alignment (in bytes) is
  // the tag: 8, two 8-byte reads per row; 4, four dword reads per row;
  // otherwise one alignment for both rows (the row pitch is a dword
  // multiple), six naturally aligned dword reads per row and byte funnel
  // shifts (unaligned LDS reads stall the LDS pipe): the pair A = dwords
  // (2 rh, 2 rh + 1) and the dword after it, then the pair B = (2 - 2 rh,
  // 3 - 2 rh) and the dword after it, which gives the four funnels in the
  // lane's dword order for either rh.  The window rows start
  // 16-byte aligned and the pitch is a multiple of 16, so step 8's probes
  // (dx a multiple of 8 columns, from the centre) are 8-byte aligned and
  // step 4's and the partial block's (0, 0) dword aligned.  An 8-byte read
  // puts a lane on two adjacent banks in one access, which spreads step 8's
  // probes (whose dword offsets otherwise share one bank class for every
  // horizontal direction) over more banks: fewer conflict cycles.
  // `at`: the lane's pair A offset at the probe, `off` the wave-uniform
  // best-vector offset.
  auto sad_at = [&](uint32_t at, int off, auto align_tag) -> uint32_t {
    constexpr int kAlign = decltype(align_tag)::value;
    constexpr bool ALIGNED = kAlign >= 4;
    // Pair A starts at the probe rounded down to a dword (8 rh keeps the
    // byte shift), pair B at dab from it.  The bases are opaque: all read
    // offsets (at most 400 bytes) then fit the 8-bit offsets of ds_read2.
    const uint32_t p = (win + (uint32_t)off) + at;
    const uint32_t sh = ALIGNED ? 0u : p & 3;
    uint32_t pa = p - sh;
    asm volatile("" : "+v"(pa));
    uint32_t pb = pa + L.dab;
    asm volatile("" : "+v"(pb));
    LdsWords wa = lds_words(pa), wb = lds_words(pb);
    constexpr int P = 8 * G::lp >> 2;  // row kk + 8
    uint4 q0, q1;
    if (kAlign >= 8) {
      typedef uint32_t u32x2a8 __attribute__((ext_vector_type(2), aligned(8)));
      using LdsV = __attribute__((address_space(3))) const u32x2a8*;
      const u32x2a8 a0 = *(LdsV)wa, b0 = *(LdsV)wb, a1 = *(LdsV)(wa + P), b1 = *(LdsV)(wb + P);
      q0 = make_uint4(a0.x, a0.y, b0.x, b0.y);
      q1 = make_uint4(a1.x, a1.y, b1.x, b1.y);
    } else if (ALIGNED) {
      q0 = make_uint4(wa[0], wa[1], wb[0], wb[1]);
      q1 = make_uint4(wa[P], wa[P + 1], wb[P], wb[P + 1]);
    } else {
      const uint32_t a0 = wa[0], a1 = wa[1], ax = wa[2], b0 = wb[0], b1 = wb[1], bx = wb[2];
      const uint32_t c0 = wa[P], c1 = wa[P + 1], cx = wa[P + 2], d0 = wb[P], d1 = wb[P + 1], dx = wb[P + 2];
      q0 = make_uint4(funnel(a0, a1, sh), funnel(a1, ax, sh), funnel(b0, b1, sh), funnel(b1, bx, sh));
      q1 = make_uint4(funnel(c0, c1, sh), funnel(c1, cx, sh), funnel(d0, d1, sh), funnel(d1, dx, sh));
    }
    if (kCols) {
      q0.x &= colmask[0]; q0.y &= colmask[1]; q0.z &= colmask[2]; q0.w &= colmask[3];
      q1.x &= colmask[0]; q1.y &= colmask[1]; q1.z &= colmask[2]; q1.w &= colmask[3];
    }
    uint32_t acc = sad4(src1.x, q1.x, 0);
    acc = sad4(src1.y, q1.y, acc);
    acc = sad4(src1.z, q1.z, acc);
    acc = sad4(src1.w, q1.w, acc);
    if (kRows) acc &= rowmask1;
    acc = sad4(src0.x, q0.x, acc);
    acc = sad4(src0.y, q0.y, acc);
    acc = sad4(src0.z, q0.z, acc);
    acc = sad4(src0.w, q0.w, acc);
    if (kRows) acc &= rowmask0;
    return sum8(acc);
  };
  // boff: window offset of the current best vector (dy * lp + dx)
  int boff = 0;
  uint32_t bsad;
  static_assert(MASK == kMaskNone || !WAVE_ZERO, "the wave-wide (0, 0) SAD covers a whole block");
  if constexpr (WAVE_ZERO) {
    // SAD at (0, 0) over the whole wave: lane = (row l / 4, dword l % 4),
    // both operands dword aligned.
    uint32_t z = sad4(zsrc, *lds_words(win + L.blend), 0);
    bsad = wave_sum_to_sgpr(sum16(z));
  } else {
    bsad = (uint32_t)__builtin_amdgcn_readfirstlane((int)sad_at(L.at0, 0, std::integral_constant<int, 4>{}));
  }
  // Probe order is the group index and the incumbent keeps ties, so the key
  // sad << 11 | group << 7 | (probe offset + 64) against best << 11 yields
  // exactly the reference's winner as one unsigned minimum over the wave,
  // and carries the winner's window offset (|offset| <= lp + 1 < 64).
  static_assert(G::lp + 1 < 64, "probe offsets fit the key's 7-bit field");
  static_assert((G::lp & 3) == 0, "aligned steps need a dword-multiple pitch");
  const uint32_t key_lo = ((uint32_t)gq << 7) | (uint32_t)(gd_off + 64);
  auto step = [&](auto sh_tag) {
    constexpr int sh = decltype(sh_tag)::value;
    // alignment of the probe offsets: step 8 from the centre (boff = 0),
    // step 4 by whole dwords, steps 2 and 1 by bytes
    constexpr int kStepAlign = sh == 3 ? 8 : (sh == 2 ? 4 : 1);
    const uint32_t at = sh == 3 ? L.at8 : (sh == 2 ? L.at4 : (sh == 1 ? L.at2 : L.at1));
    const uint32_t s = sad_at(at, boff, std::integral_constant<int, kStepAlign>{});
    const uint32_t kmin = wave_min_to_sgpr((s << 11) | key_lo);
    if (kmin < (bsad << 11)) {
      bsad = kmin >> 11;
      boff += ((int)(kmin & 127) - 64) * (1 << sh);
    }
  };
  step(std::integral_constant<int, 3>{});
  step(std::integral_constant<int, 2>{});
  step(std::integral_constant<int, 1>{});
  step(std::integral_constant<int, 0>{});
  // |dx| <= 15 < lp / 2: dy is boff / lp rounded to nearest, dx the rest
  besty = (boff + 16 * G::lp + G::lp / 2) / G::lp - 16;
  bestx = boff - besty * G::lp;
  best_sad = bsad;
}



// A plane as a raw buffer: base in scalar registers, loads take a 32-bit
// lane offset (plus an optional wave-uniform one) and read zero past
// `bytes` instead of leaving the plane.  The DMA paths need every plane
// below 2^31 bytes (the host clears uniform_strides otherwise).
constexpr int kRawBufferWord3 = 0x00020000;  // 32-bit data format, no swizzle
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const uint8_t* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, kRawBufferWord3);
}

// DMA of one frame's luma window into `win` (buffer_load_dwordx4 ... lds:
// lane l of an instruction lands at win + 16 l): piece p = lane + 64 i is
// window row p / 3, 16-byte column p % 3, at byte 16 p of the window.
// off_i is lane piece i's offset in the plane.  The data is in LDS once
// vmcnt has drained (wait_vm).
using LdsPtr = __attribute__((address_space(3))) void*;
// MASKED: lane piece i is skipped when bit i of `outside` is set (a piece
// wholly outside the frame's columns; the caller fills it).
template <bool MASKED>
__device__ __forceinline__ void dma_window(__amdgpu_buffer_rsrc_t plane, uint32_t off0, uint32_t off1, uint32_t off2,
                                           uint8_t* win, int lane, uint32_t outside) {
  if (!MASKED || !(outside & 1)) __builtin_amdgcn_raw_ptr_buffer_load_lds(plane, (LdsPtr)(win), 16, off0, 0, 0, 0);
  if (!MASKED || !(outside & 2))
    __builtin_amdgcn_raw_ptr_buffer_load_lds(plane, (LdsPtr)(win + 1024), 16, off1, 0, 0, 0);
  if (lane < kLumaPieces - 128 && (!MASKED || !(outside & 4)))
    __builtin_amdgcn_raw_ptr_buffer_load_lds(plane, (LdsPtr)(win + 2048), 16, off2, 0, 0, 0);
}

// Every vector memory operation of the wave done (DMA writes included), and
// no memory access moved across.
__device__ __forceinline__ void wait_vm() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Windows the DMA cannot take (planes not 16-byte aligned or without one
// stride, or a window piece straddling the frame's right edge): clamped
// dword gathers into LDS.
__device__ __forceinline__ void stage_gather(uint8_t* win, const FrameDesc& f, const ArnrGroup& a, int bx, int by,
                                             int lane) {
  constexpr int ldw = kLumaPitch / 4;  // dwords per window row
  for (int s = lane; s < kWinRows * ldw; s += 64) {
    const int wy = s / ldw, dq = s - wy * ldw;
    const int y = clampi(by - 15 + wy, 0, a.h - 1);
    reinterpret_cast<uint32_t*>(win)[s] =
        load_clamped4(f.y + (size_t)y * f.y_stride, bx - 16 + 4 * dq, a.w, a.aligned4 != 0);
  }
}

// Row r of the anchor MB as 16 bytes (zero beyond (bw, bh)) from the lanes'
// blend pixels, in the search lane's dword order (pairs swapped when rh):
// lane 4r + q holds dword q of row r in `ctr4`.  A crossbar exchange (no LDS
// memory, whole wave active) instead of a second global load of the
// centre, which every wave would issue at launch.
__device__ __forceinline__ uint4 centre_row16(uint32_t ctr4, int r, int rh) {
  auto from = [&](int l) { return (uint32_t)__builtin_amdgcn_ds_bpermute(4 * l, (int)ctr4); };
  const int a = 4 * r + 2 * rh, b = 4 * r + 2 - 2 * rh;
  return make_uint4(from(a), from(a + 1), from(b), from(b + 1));
}

// Issue priority of a wave at iteration `it` of `nf` frames: 3 down to 0 as
// it progresses, so that the SIMD's arbiter (highest priority first, then
// oldest) lets waves that are behind catch up.  Waves then finish together
// instead of leaving the SIMD to a few latency-bound stragglers.
__device__ __forceinline__ void pace(int it, int nf) {
  const int p = 3 - (it * 4) / nf;
  if (p >= 3)
    __builtin_amdgcn_s_setprio(3);
  else if (p == 2)
    __builtin_amdgcn_s_setprio(2);
  else if (p == 1)
    __builtin_amdgcn_s_setprio(1);
  else
    __builtin_amdgcn_s_setprio(0);
}

// One anchor MB as a wave sees it, and the lane's packed accumulators
// (wsum << 20 | acc) of luma pixels 4*lq .. 4*lq+3 of row lr and of chroma
// pixel (cr, cc) of U and V.
struct MbState {
  int bx, by, bw, bh, cbx, cby, brow;
  uint32_t ctr4, cu, cv;  // centre pixels of the lane's blend positions
  uint32_t ay0, ay1, ay2, ay3, au, av;
};

// x = ref | 1 << 20: sad(x, ctr | 1 << 20) = |ref - ctr|, and wt * x adds wt
// to the weight field and wt * ref to the sum.
__device__ __forceinline__ void blend_px(uint32_t& acc, uint32_t ref, uint32_t ctr, LdsBytes lut) {
  const uint32_t x = ref | kAccOne;
  acc += lut[sad4(x, ctr | kAccOne, 0)] * x;
}

// Frame loops of one wave, one per kind of MB, so that each loop keeps only
// its own lane constants live:
//   kPathFast:    interior 16x16 MB (window and chroma reach inside the
//                 frame): the unmasked search, chroma unclamped.
//   kPathDma:     any other MB on the DMA path (top and bottom rows, left
//                 and right columns; always 16 pixels wide): the search
//                 with rows >= bh masked on the row-clamped, edge-filled
//                 window, chroma clamped.
//   kPathGather:  planes not 16-byte aligned or without one stride, or a
//                 window piece straddling the frame's right edge (width not
//                 a multiple of 16): clamped dword gathers.
// DMA MBs (16-byte aligned luma, one stride for all frames, every window
// piece wholly inside or wholly outside the frame's columns) take their
// window by DMA with row-clamped piece offsets; pieces outside the columns
// are filled with the row's edge pixel.  The next frame's window arrives
// during this frame's search.  A frame's chroma bytes are loaded after its
// search and blended after the next frame's wait, so that their latency
// overlaps the next search.
enum { kPathFast = 1, kPathDma = 2, kPathGather = 3 };

// `lut_sync` fills the weight LUT and synchronises the workgroup; every wave
// calls it exactly once, here after its first window DMA is in flight.
template <int PATH, class FrameAt, class LutSync>
__device__ __forceinline__ void frame_loop(const ArnrGroup& a, const FrameAt& frame_at, MbState& s, int split,
                                           int K, int lane, uint8_t* win, LdsBytes s_lut, const LutSync& lut_sync) {
  const int lr = lane >> 2, lq = lane & 3;  // luma: row, dword
  const int cr = lane >> 3, cc = lane & 7;  // chroma: pixel
  const int nnb = a.filter_on ? a.n - 1 : 0;
  // search: lane = probe group x row pair kk (rows kk, kk + 8); gq is the
  // group's probe and rh its dword pair order
  const int gq = lane_probe(lane), kk = lane & 7, rh = probe_pair_order(gq);
  const int gd_off = (((kProbeDy1 >> (2 * gq)) & 3) - 1) * kLumaPitch + (((kProbeDx1 >> (2 * gq)) & 3) - 1);
  SearchLanes L;
  {
    const uint32_t at0 = (uint32_t)((kk + 15) * kLumaPitch + 16 + 8 * rh);
    L.at0 = at0;
    L.at8 = at0 + (uint32_t)(gd_off * 8);
    L.at4 = at0 + (uint32_t)(gd_off * 4);
    L.at2 = at0 + (uint32_t)(gd_off * 2);
    L.at1 = at0 + (uint32_t)gd_off;
    L.dab = (uint32_t)(8 - 16 * rh);
    L.blend = (uint32_t)((lr + 15) * kLumaPitch + 16 + 4 * lq);
    asm volatile("" : "+v"(L.at8), "+v"(L.at4), "+v"(L.at2), "+v"(L.at1), "+v"(L.at0), "+v"(L.dab), "+v"(L.blend));
  }
  constexpr bool kDma = PATH == kPathFast || PATH == kPathDma;
  // window pieces may lie outside the frame's columns (filled with the edge pixel)
  constexpr bool kMasked = PATH == kPathDma;
  uint4 src0, src1;
  // DMA paths: plane offset of lane piece i (window row p / 3 clamped to the
  // frame, 16-byte column p % 3, p = lane + 64 i); kPathDma: bit i of
  // `outside` set when the piece lies wholly outside the frame's columns,
  // its offset then that of the row's edge pixel, which is repeated.
  uint32_t off0 = 0, off1 = 0, off2 = 0, coff = 0, outside = 0;
  {
    const FrameDesc C = frame_at(a.center);
    if (kDma) {
      auto piece = [&](int p, int i) -> uint32_t {
        const int row = p / 3;
        const uint32_t y = (uint32_t)clampi(s.by - 15 + row, 0, a.h - 1) * C.y_stride;
        const int x0 = s.bx - 16 + 16 * (p - 3 * row);
        if (kMasked && (x0 < 0 || x0 + 16 > a.w)) {
          outside |= 1u << i;
          return y + (uint32_t)(x0 < 0 ? 0 : a.w - 1);
        }
        return y + (uint32_t)x0;
      };
      off0 = piece(lane, 0);
      off1 = piece(lane + 64, 1);
      off2 = piece(min(lane + 128, kLumaPieces - 1), 2);
    }
    if (PATH == kPathFast) coff = (uint32_t)cr * C.uv_stride + (uint32_t)cc;
  }
  // edge pixels of the outside pieces of the frame in flight (kMasked)
  uint32_t e0 = 0, e1 = 0, e2 = 0;
  auto fetch_edges = [&](const uint8_t* plane) {
    if (outside & 1) e0 = plane[off0];
    if (outside & 2) e1 = plane[off1];
    if (outside & 4) e2 = plane[off2];
  };
  auto rep16 = [](uint32_t b) {
    const uint32_t r = b * 0x01010101u;
    return make_uint4(r, r, r, r);
  };
  auto fill_edges = [&](uint8_t* wb) {
    if (outside & 1) reinterpret_cast<uint4*>(wb)[lane] = rep16(e0);
    if (outside & 2) reinterpret_cast<uint4*>(wb)[lane + 64] = rep16(e1);
    if ((outside & 4) && lane < kLumaPieces - 128) reinterpret_cast<uint4*>(wb)[lane + 128] = rep16(e2);
  };

  int k = split;
  int cur = 0;  // window buffer of the frame being searched
  FrameDesc F{};
  if (k < nnb) {
    F = frame_at(skip_centre(k, a.center));
    if (kDma) dma_window<kMasked>(plane_rsrc(F.y, (uint32_t)a.h * F.y_stride), off0, off1, off2, win, lane, outside);
    if (kMasked) fetch_edges(F.y);
  }
  lut_sync();
  src0 = centre_row16(s.ctr4, kk, rh);
  src1 = centre_row16(s.ctr4, kk + 8, rh);
  bool pend = false;  // chroma bytes of the previous frame await their blend
  uint32_t ru = 0, rv = 0;
  const int nf = (nnb - split + K - 1) / K;
  for (int it = 0; k < nnb; k += K, it++) {
    const int fi = skip_centre(k, a.center);
    const int kn = k + K;
    FrameDesc Fn{};
    if (kn < nnb) Fn = frame_at(skip_centre(kn, a.center));
    uint8_t* w = win + cur * kLumaWin;
    pace(it, nf);
    if (kDma) wait_vm();  // this frame's window and the previous frame's chroma have landed
    // Registers loaded for this frame are read here (unconditionally: the
    // compiler's own waits for them then sit here), before the next DMA goes
    // out, so none of those waits covers the DMA.
    asm volatile("" ::"v"(ru), "v"(rv), "v"(e0), "v"(e1), "v"(e2));
    if (kMasked) fill_edges(w);
    if (pend) {
      blend_px(s.au, ru, s.cu, s_lut);
      blend_px(s.av, rv, s.cv, s_lut);
    }
    if (kDma) {
      if (kn < nnb) {
        dma_window<kMasked>(plane_rsrc(Fn.y, (uint32_t)a.h * Fn.y_stride), off0, off1, off2, win + (cur ^ 1) * kLumaWin,
                            lane, outside);
        if (kMasked) fetch_edges(Fn.y);
      }
    } else if (PATH == kPathGather) {
      wave_lds_sync();  // the previous frame's window reads are done
      stage_gather(w, F, a, s.bx, s.by, lane);
      wave_lds_sync();
    }

    int bestx, besty;
    uint32_t best_sad;
    {
      const uint32_t wl = lds_addr(w);
      if constexpr (PATH == kPathFast) {
        // (0, 0) operands of the whole-wave SAD: the lane's blend dword
        descend<kMaskNone, true, MbGeom>(wl, L, src0, src1, gd_off, gq, kk, rh, 16, 16, bestx, besty, best_sad,
                                         s.ctr4);
      } else if constexpr (PATH == kPathDma) {
        // a DMA MB is 16 pixels wide: a partial-width MB has a window piece
        // straddling the right edge, which sends it to the gather path
        descend<kMaskRows, false, MbGeom>(wl, L, src0, src1, gd_off, gq, kk, rh, 16, s.bh, bestx, besty, best_sad, 0);
      } else {
        descend<kMaskBoth, false, MbGeom>(wl, L, src0, src1, gd_off, gq, kk, rh, s.bw, s.bh, bestx, besty, best_sad,
                                          0);
      }
      if (a.mv_out && lane == 0) {
        int32_t* o = a.mv_out + 3 * (((size_t)fi * a.brows + s.brow) * a.bcols + (s.bx >> 4));
        o[0] = bestx;
        o[1] = besty;
        o[2] = (int32_t)best_sad;
      }
    }

    // sad / (bw*bh) <= 12  <=>  sad < 13 * bw * bh (integer division).
    // Positions outside the block accumulate too (their window bytes are
    // valid clamped samples); the resolve never writes them out.
    pend = best_sad < (kSadPerPixelCutoff + 1) * (uint32_t)(s.bw * s.bh);
    if (pend) {
      const int cmvx = bestx / 2, cmvy = besty / 2;  // truncation toward zero
      // the lane's 4 reference pixels at the vector: the lane's blend offset
      // plus the wave-uniform window base and vector offset
      const uint32_t rp4 = (lds_addr(w) + (uint32_t)(besty * kLumaPitch + bestx)) + L.blend;
      const uint32_t rsh = rp4 & 3;
      const LdsWords rw = lds_words(rp4 - rsh);
      const uint32_t ref4 = funnel(rw[0], rw[1], rsh);
      if (PATH == kPathFast) {
        // the block's chroma origin at the vector (inside the plane here) as
        // the wave-uniform offset, the lane's pixel as the lane offset
        // (the descriptor is read again here, so that its pointers stay
        // wave-uniform scalars instead of being carried over the search)
        const FrameDesc Fc = frame_at(fi);
        const uint32_t cb = (uint32_t)(s.cby + cmvy) * Fc.uv_stride + (uint32_t)(s.cbx + cmvx);
        const uint32_t cbytes = (uint32_t)a.ch * Fc.uv_stride;
        ru = __builtin_amdgcn_raw_buffer_load_b8(plane_rsrc(Fc.u, cbytes), coff, cb, 0);
        rv = __builtin_amdgcn_raw_buffer_load_b8(plane_rsrc(Fc.v, cbytes), coff, cb, 0);
      } else {
        const size_t o = (size_t)clampi(s.cby + cr + cmvy, 0, a.ch - 1) * F.uv_stride +
                         clampi(s.cbx + cc + cmvx, 0, a.cw - 1);
        ru = F.u[o];
        rv = F.v[o];
      }
      blend_px(s.ay0, ref4 & 255, s.ctr4 & 255, s_lut);
      blend_px(s.ay1, (ref4 >> 8) & 255, (s.ctr4 >> 8) & 255, s_lut);
      blend_px(s.ay2, (ref4 >> 16) & 255, (s.ctr4 >> 16) & 255, s_lut);
      blend_px(s.ay3, ref4 >> 24, s.ctr4 >> 24, s_lut);
    }
    F = Fn;
    cur ^= 1;
  }
  if (pend) {
    blend_px(s.au, ru, s.cu, s_lut);
    blend_px(s.av, rv, s.cv, s_lut);
  }
}

struct TableEntry {
  int x = 0, y = 0;
  uint32_t sad = 0xFFFFFFFFu;
};

// Lane j: the table entry of the wave's j-th frame (sad ~0 past the last).
__device__ __forceinline__ TableEntry load_table_entry(const ArnrGroup& a, const MbState& s, int split, int K, int lane,
                                                      const int32_t* table) {
  const int nnb = a.filter_on ? a.n - 1 : 0;
  const int count = split < nnb ? (nnb - split + K - 1) / K : 0;
  TableEntry e;
  if (lane < count) {
    const int k = split + lane * K;
    const int fi = skip_centre(k, a.center);
    const int32_t* t = table + 3 * (((size_t)fi * a.brows + s.brow) * a.bcols + (s.bx >> 4));
    e.x = t[0];
    e.y = t[1];
    e.sad = (uint32_t)t[2];
  }
  return e;
}

// The altref blend from a vector table (n x brows x bcols x {mvx, mvy,
// sad}).  The wave's frames are k = split, split + K, ... of the nnb
// neighbours; lane j loads the table entry of its j-th frame, and a ballot
// of the entries under the SAD cutoff gives the frames that blend.  For each
// of those, every lane gathers its 4 luma and 2 chroma neighbour bytes with
// clamped coordinates (the vectors are not bounded by a staged window).
// The kernel is bound by the count of memory instructions, so with 4-byte
// aligned luma planes and strides and rows of 8 pixels or more (A4) the 4
// luma bytes come from one dword-aligned 8-byte read whose start is the
// clamped first column rounded down to a dword and moved left, when needed,
// so that both dwords start inside the row; one byte permute then applies
// the column clamp.
template <bool A4, class FrameAt, class LutSync>
__device__ __forceinline__ void table_loop(const ArnrGroup& a, const FrameAt& frame_at, MbState& s, int split, int K,
                                           int lane, LdsBytes s_lut, const TableEntry& e, const LutSync& lut_sync) {
  lut_sync();
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2), aligned(4)));
  const int lr = lane >> 2, lq = lane & 3;  // luma: row, dword
  const int cr = lane >> 3, cc = lane & 7;  // chroma: pixel
  const int last4 = ((a.w - 1) & ~3) - 4;   // last start of an 8-byte read
  // sad / (bw*bh) <= 12  <=>  sad < 13 * bw * bh (integer division)
  uint64_t todo = __builtin_amdgcn_ballot_w64(e.sad < (kSadPerPixelCutoff + 1) * (uint32_t)(s.bw * s.bh));
  while (todo) {
    const int j = __builtin_ctzll(todo);
    todo &= todo - 1;
    const int k = split + j * K;
    const FrameDesc F = frame_at(skip_centre(k, a.center));
    const int mvx = __builtin_amdgcn_readlane(e.x, j), mvy = __builtin_amdgcn_readlane(e.y, j);
    const uint8_t* rrow = F.y + (size_t)clampi(s.by + lr + mvy, 0, a.h - 1) * F.y_stride;
    const int xr = s.bx + 4 * lq + mvx;
    uint32_t r4;
    if (A4) {
      const int p0 = clampi(xr, 0, a.w - 1);
      const int a0 = min(p0 & ~3, last4);
      const uint32_t sel = (uint32_t)(p0 - a0) | (uint32_t)(clampi(xr + 1, 0, a.w - 1) - a0) << 8 |
                           (uint32_t)(clampi(xr + 2, 0, a.w - 1) - a0) << 16 |
                           (uint32_t)(clampi(xr + 3, 0, a.w - 1) - a0) << 24;
      const u32x2 d = *reinterpret_cast<const u32x2*>(rrow + a0);
      r4 = __builtin_amdgcn_perm(d.y, d.x, sel);
    } else {
      r4 = (uint32_t)rrow[clampi(xr, 0, a.w - 1)] | (uint32_t)rrow[clampi(xr + 1, 0, a.w - 1)] << 8 |
           (uint32_t)rrow[clampi(xr + 2, 0, a.w - 1)] << 16 | (uint32_t)rrow[clampi(xr + 3, 0, a.w - 1)] << 24;
    }
    const int cmvx = mvx / 2, cmvy = mvy / 2;  // truncation toward zero
    const size_t o =
        (size_t)clampi(s.cby + cr + cmvy, 0, a.ch - 1) * F.uv_stride + clampi(s.cbx + cc + cmvx, 0, a.cw - 1);
    const uint32_t ru = F.u[o], rv = F.v[o];
    blend_px(s.ay0, r4 & 255, s.ctr4 & 255, s_lut);
    blend_px(s.ay1, (r4 >> 8) & 255, (s.ctr4 >> 8) & 255, s_lut);
    blend_px(s.ay2, (r4 >> 16) & 255, (s.ctr4 >> 16) & 255, s_lut);
    blend_px(s.ay3, r4 >> 24, s.ctr4 >> 24, s_lut);
    blend_px(s.au, ru, s.cu, s_lut);
    blend_px(s.av, rv, s.cv, s_lut);
  }
}

// Anchor MB `mi` of strip (brow, scol) of group `a`, frames split `split` of
// K: the wave's frames are k = split, split + K, ... of the nnb neighbours.
// `frame_at(i)` returns frame i of the group, `wt4` the weight LUT.
// TABLE: the per-block vectors and SADs come from `table` (layout
// n x brows x bcols x {mvx, mvy, sad}) instead of the descent, and the
// blend reads the neighbour straight from memory with clamped coordinates
// (the vectors are not bounded by the +-15 window).
// LDS: weight LUT 256 B | (K > 1) M x 384 packed accumulators | per-wave
// window pairs.  Every wave of the workgroup calls this (inactive MBs past the
// right frame edge included: they take part in the barriers); the caller
// synchronises the workgroup before reusing LDS.
// The weight LUT sits in static LDS: its address is a compile-time constant
// that folds into the ds_read offsets of the blend.  The dynamic LDS (split
// accumulators, windows) follows it.
__shared__ __align__(16) uint32_t s_arnr_lut[64];

template <bool TABLE, class FrameAt>
__device__ __forceinline__ void process_mbs(const ArnrGroup& a, const uint32_t* wt4, const FrameAt& frame_at,
                                            int brow, int scol, int M, uint8_t* smem,
                                            const int32_t* table = nullptr) {
  const int nwaves = blockDim.x >> 6;
  const int K = nwaves / M;
  LdsBytes s_lut = (LdsBytes)s_arnr_lut;
  uint32_t* s_acc = reinterpret_cast<uint32_t*>(smem);
  uint8_t* s_win = smem + (K > 1 ? M * 384 * 4 : 0);

  // Opaque to the compiler, so that in the persistent batch loop the lane
  // constants derived from it are recomputed per unit instead of being
  // hoisted and held across units (96 -> 72 VGPRs there).
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // M is 1, 2, 4 or 8
  const int mi = wave & (M - 1), split = wave >> __builtin_ctz((unsigned)M);

  const int lr = lane >> 2, lq = lane & 3;  // luma: row, dword
  const int cr = lane >> 3, cc = lane & 7;  // chroma: pixel
  MbState s;
  s.bx = (scol * M + mi) * 16;
  s.by = brow * 16;
  s.brow = brow;
  const bool active = s.bx < a.w;
  s.bw = min(16, a.w - s.bx);
  s.bh = min(16, a.h - s.by);
  s.cbx = s.bx >> 1;
  s.cby = s.by >> 1;
  const int cbw = (s.bw + 1) >> 1, cbh = (s.bh + 1) >> 1;
  s.ay0 = s.ay1 = s.ay2 = s.ay3 = s.au = s.av = 0;
  s.ctr4 = s.cu = s.cv = 0;

  // The table entries and centre pixels are loaded before the LUT and the
  // barrier, so that their latency overlaps the LUT's.
  TableEntry e;
  if (TABLE && active) e = load_table_entry(a, s, split, K, lane, table);
  if (active) {
    const FrameDesc C = frame_at(a.center);
    // centre pixels of this lane's blend positions (zero outside the frame)
    if (s.bw == 16 && a.aligned4) {
      if (lr < s.bh)
        s.ctr4 = *reinterpret_cast<const uint32_t*>(C.y + (size_t)(s.by + lr) * C.y_stride + s.bx + 4 * lq);
    } else {
      // all four loads issued (clamped into the block), outside bytes zeroed
      const uint8_t* row = C.y + (size_t)(s.by + min(lr, s.bh - 1)) * C.y_stride + s.bx;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int c = 4 * lq + j;
        const uint32_t keep = c < s.bw && lr < s.bh ? 0xFFu : 0u;
        s.ctr4 |= ((uint32_t)row[min(c, s.bw - 1)] & keep) << (8 * j);
      }
    }
    if (cr < cbh && cc < cbw) {
      const size_t o = (size_t)(s.cby + cr) * C.uv_stride + s.cbx + cc;
      s.cu = C.u[o];
      s.cv = C.v[o];
    }
  }
  // The LUT and the split accumulators are written, and the workgroup
  // synchronised, by each wave's frame loop once its first window DMA is
  // issued, so that the DMA's latency overlaps the centre and LUT loads.
  // With K = 1 no wave reads another's LDS: each wave writes the whole LUT
  // itself (every wave the same values) and needs only its own writes done,
  // so no wave waits at a barrier for the workgroup's slowest prologue.
  auto lut_sync = [&]() {
    if (K > 1) {
      for (int i = tid; i < 64; i += blockDim.x) s_arnr_lut[i] = wt4[i];
      for (int i = tid; i < 384 * M; i += blockDim.x) s_acc[i] = 0;
      __syncthreads();
    } else {
      s_arnr_lut[lane] = wt4[lane];
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    }
  };

  if (active) {
    uint8_t* win = s_win + wave * kWaveWinBytes;
    // DMA: every window piece wholly inside or wholly outside the frame's
    // columns; fast: the whole window and every chroma read at +-7 inside
    const bool cols_inside = s.bx >= 16 && s.bx + 32 <= a.w;
    const bool dma = a.aligned16 && a.uniform_strides && (cols_inside || (a.w & 15) == 0);
    const bool fast = dma && cols_inside && s.by >= 15 && s.by + 31 <= a.h && s.cbx >= 7 && s.cbx + 15 <= a.cw &&
                      s.cby >= 7 && s.cby + 15 <= a.ch;
    if (TABLE && a.aligned4 && a.w >= 8)
      table_loop<true>(a, frame_at, s, split, K, lane, s_lut, e, lut_sync);
    else if (TABLE)
      table_loop<false>(a, frame_at, s, split, K, lane, s_lut, e, lut_sync);
    else if (fast)
      frame_loop<kPathFast>(a, frame_at, s, split, K, lane, win, s_lut, lut_sync);
    else if (dma)
      frame_loop<kPathDma>(a, frame_at, s, split, K, lane, win, s_lut, lut_sync);
    else
      frame_loop<kPathGather>(a, frame_at, s, split, K, lane, win, s_lut, lut_sync);
  } else {
    lut_sync();
  }

  if (K > 1) {
    // lane offsets derived again from the lane id, so that none is held
    // across the frame loops
    const int lane2 = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const int lr = lane2 >> 2, lq = lane2 & 3;
    const int cr = lane2 >> 3, cc = lane2 & 7;
    uint32_t* acc = s_acc + 384 * mi;
    if (active) {
      lds_add(&acc[lr * 16 + 4 * lq + 0], s.ay0);
      lds_add(&acc[lr * 16 + 4 * lq + 1], s.ay1);
      lds_add(&acc[lr * 16 + 4 * lq + 2], s.ay2);
      lds_add(&acc[lr * 16 + 4 * lq + 3], s.ay3);
      lds_add(&acc[256 + cr * 8 + cc], s.au);
      lds_add(&acc[320 + cr * 8 + cc], s.av);
    }
    __syncthreads();
    if (active && split == 0) {
      const uint4 v = *reinterpret_cast<const uint4*>(&acc[lr * 16 + 4 * lq]);
      s.ay0 = v.x;
      s.ay1 = v.y;
      s.ay2 = v.z;
      s.ay3 = v.w;
      s.au = acc[256 + cr * 8 + cc];
      s.av = acc[320 + cr * 8 + cc];
    }
  }
  if (!active || split != 0) return;

  // The lane's pixel offsets are derived again here from the lane id (which
  // the compiler cannot match with the ones computed before the frame loop),
  // so that no address term is held across the loop.
  {
    const int lane2 = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const int lr = lane2 >> 2, lq = lane2 & 3;
    const int cr = lane2 >> 3, cc = lane2 & 7;

  // Resolve: (16 * ctr + acc + ws / 2) / ws with ws = 16 + wsum.  The
  // numerator n is below 256 * ws and ws lies in [16, 1024] (at most 64
  // frames), so the quotient is a pixel and n / ws has a fractional part of
  // 0 or at least 1/1024.  n * rcp(ws) (v_rcp_f32, within 1 ulp) is then
  // within 2^-14 of n / ws, and adding 2^-12 before truncating gives the
  // exact floor: an integer quotient cannot fall below itself, a fractional
  // one cannot reach the next integer (checked for every ws and both ends
  // of every quotient's interval, with the reciprocal off by up to 2 ulp).
  auto resolve = [](uint32_t packed, uint32_t ctr) -> uint32_t {
    const uint32_t sum = kWMax * ctr + (packed & (kAccOne - 1));
    const uint32_t ws = kWMax + (packed >> kAccShift);
    const float r = __builtin_amdgcn_rcpf((float)ws);
    return (uint32_t)__builtin_fmaf((float)(sum + (ws >> 1)), r, 0x1p-12f);
  };
  if (lr < s.bh) {
    const uint32_t c = s.ctr4;
    const uint32_t px4 = resolve(s.ay0, c & 255) | (resolve(s.ay1, (c >> 8) & 255) << 8) |
                         (resolve(s.ay2, (c >> 16) & 255) << 16) | (resolve(s.ay3, c >> 24) << 24);
    uint8_t* dst = a.out_y + (size_t)(s.by + lr) * a.w + s.bx + 4 * lq;
    if (4 * lq + 4 <= s.bw && ((((uintptr_t)dst) & 3) == 0)) {
      *reinterpret_cast<uint32_t*>(dst) = px4;
    } else {
      for (int e = 0; e < 4; e++)
        if (4 * lq + e < s.bw) dst[e] = (uint8_t)(px4 >> (8 * e));
    }
  }
  if (cr < cbh && cc < cbw) {
    const size_t o = (size_t)(s.cby + cr) * a.cw + s.cbx + cc;
    a.out_u[o] = (uint8_t)resolve(s.au, s.cu);
    a.out_v[o] = (uint8_t)resolve(s.av, s.cv);
  }
  }
}

// At most 64 VGPRs: 8 waves per SIMD.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))
arnr_fused_kernel(const ArnrArgs a) {
  __builtin_amdgcn_s_setprio(3);  // the prologue's loads go out ahead of running searches
  extern __shared__ __align__(16) uint8_t smem[];
  const int total = a.strip_count;
  // Strips go out bottom row first: the last MB row, partial when the height
  // is not a multiple of 16, takes the masked search (the slowest path) and
  // so starts with the launch instead of closing it.
  const int strip = a.strip_base + total - 1 - (a.xcd_remap ? xcd_band(blockIdx.x, total) : (int)blockIdx.x);
  const int brow = a.spr_mul ? (int)__umulhi((uint32_t)strip, a.spr_mul) : strip;
  const int scol = strip - brow * a.g.strips_per_row;
  process_mbs<false>(a.g, a.wt4, [&](int i) { return a.frames[i]; }, brow, scol, a.mbs_per_wg, smem);
}

// The altref from a given alignment table (guard, blend, resolve).
__global__ void __launch_bounds__(1024) arnr_table_kernel(const ArnrArgs a, const int32_t* table) {
  extern __shared__ __align__(16) uint8_t smem[];
  const int total = a.g.strips_per_row * a.g.brows;
  const int strip = a.xcd_remap ? xcd_band(blockIdx.x, total) : (int)blockIdx.x;
  const int brow = a.spr_mul ? (int)__umulhi((uint32_t)strip, a.spr_mul) : strip;
  const int scol = strip - brow * a.g.strips_per_row;
  process_mbs<true>(a.g, a.wt4, [&](int i) { return a.frames[i]; }, brow, scol, a.mbs_per_wg, smem, table);
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

// A frame descriptor through the constant address space: one scalar load
// into SGPRs (the batch's descriptors are written before the launch and
// only read during it), instead of vector loads made uniform lane by lane.
__device__ __forceinline__ FrameDesc load_frame(const FrameDesc* p) {
  static_assert(sizeof(FrameDesc) == 32, "three pointers and two strides");
  using C4 = const __attribute__((address_space(4))) uint32_t*;
  const C4 q = (C4)p;
  auto ptr = [&](int i) { return reinterpret_cast<const uint8_t*>(((uint64_t)q[i + 1] << 32) | q[i]); };
  FrameDesc f;
  f.y = ptr(0);
  f.u = ptr(2);
  f.v = ptr(4);
  f.y_stride = q[6];
  f.uv_stride = q[7];
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
  g.uniform_strides = uload(&p->uniform_strides);
  g.strips_per_row = uload(&p->strips_per_row);
  return g;
}


// Persistent batch kernel: every workgroup loops taking units (one strip of
// one group) until all chunks are drained; a unit is processed exactly as
// the single-group kernel processes that strip.
// At most 64 VGPRs: 8 waves per SIMD.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))
arnr_batch_kernel(const ArnrBatchArgs b) {
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
      process_mbs<false>(g, job->wt4, [&](int i) { return load_frame(b.frames + fb + i); }, brow, scol,
                         b.mbs_per_wg, smem);
    }
    __syncthreads();  // LDS and s_unit are free for the next unit
  }
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}


// Dynamic LDS of one workgroup: the split accumulators when K > 1, and one
// window per wave unless the vectors come from a table.
static size_t arnr_dynamic_lds_bytes(int k, int m, bool table) {
  return (k > 1 ? (size_t)m * 384 * 4 : 0) + (table ? 0 : (size_t)m * k * kWaveWinBytes);
}

// All LDS of one workgroup: the static LUT and the dynamic part.
size_t arnr_lds_bytes(int k, int m, bool table) {
  return sizeof(s_arnr_lut) + arnr_dynamic_lds_bytes(k, m, table);
}

int arnr_supported_mbs(int m) { return m >= 8 ? 8 : (m >= 4 ? 4 : (m >= 2 ? 2 : 1)); }

hipError_t launch_arnr_fused(const ArnrArgs& a, int k, hipStream_t stream) {
  const int blocks = a.strip_count;
  if (blocks < 1 || a.strip_base < 0 || a.strip_base + blocks > a.g.strips_per_row * a.g.brows)
    return hipErrorInvalidValue;
  const int m = a.mbs_per_wg;
  if (m < 1 || k < 1 || m * k > 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(arnr_fused_kernel, dim3(blocks), dim3(64 * m * k), arnr_dynamic_lds_bytes(k, m, false), stream, a);
  return hipGetLastError();
}

int arnr_batch_workgroups_per_cu(int k, int m) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, arnr_batch_kernel, 64 * m * k,
                                                   arnr_dynamic_lds_bytes(k, m, false)) != hipSuccess)
    return 0;
  return per_cu;
}

hipError_t launch_arnr_batch(const ArnrBatchArgs& b, int k, int workgroups, hipStream_t stream) {
  const int m = b.mbs_per_wg;
  if (m < 1 || k < 1 || m * k > 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(arnr_batch_kernel, dim3(workgroups), dim3(64 * m * k), arnr_dynamic_lds_bytes(k, m, false), stream,
                     b);
  return hipGetLastError();
}

hipError_t launch_arnr_table(const ArnrArgs& a, const int32_t* table, int k, hipStream_t stream) {
  const int blocks = a.g.strips_per_row * a.g.brows;
  const int m = a.mbs_per_wg;
  if (m < 1 || k < 1 || m * k > 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(arnr_table_kernel, dim3(blocks), dim3(64 * m * k), arnr_dynamic_lds_bytes(k, m, true), stream, a,
                     table);
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace vp8g
