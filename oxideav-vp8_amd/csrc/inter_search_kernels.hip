// The encoder's inter motion search on gfx950: whole_pixel_search and
// sub_pixel_refine (src/encoder/inter_search.rs:380-552) with MvCostTable
// rate terms (:166-194), and small_diamond_search_luma
// (src/motion_search.rs:381-442).
//
// One wave per macroblock.  Each search is a chain of rounds whose
// candidate positions are fixed when the round starts, so a round's SADs
// are computed together: lane l holds source row l/4, bytes 4*(l%4)..+3,
// gathers the four reference bytes under each candidate (clamped, as
// whole_sad / mb_luma_sad_at_whole_mv clamp) and the wave sums each
// candidate's v_sad_u8 partials.  The accept decisions then run in the
// reference's order on the wave-uniform sums.  The reference abandons a
// SAD once it reaches the bound the best cost allows; such a candidate is
// rejected whatever its full SAD, so full SADs give the same decisions.
//
// Rates are taken with the reference's floating-point operations: the two
// f32 table entries are added in f32 and widened, multiplied by lambda_sad
// and added to the SAD in f64, each with round-to-nearest and no fused
// multiply-add (__fadd_rn, __dmul_rn, __dadd_rn).
#include "vp8g_device.h"
#include "vp8g_internal.h"

namespace vp8g {

namespace {

// SIXTAP_FILTERS (RFC 6386 section 18.3, src/motion_comp.rs:294-303): the
// taps mb_luma_sad_at_mv always uses.
__constant__ int32_t kSixtapTaps[8][6] = {{0, 0, 128, 0, 0, 0},   {0, -6, 123, 12, -1, 0}, {2, -11, 108, 36, -8, 1},
                                          {0, -9, 93, 50, -6, 0},  {3, -16, 77, 77, -16, 3}, {0, -6, 50, 93, -9, 0},
                                          {1, -8, 36, 108, -11, 2}, {0, -1, 12, 123, -6, 0}};

// The lane's 4 reference bytes of the 16x16 block at (x, y), coordinates
// clamped to w x h.
__device__ __forceinline__ uint32_t block_bytes(const uint8_t* plane, uint32_t stride, int w, int h, int x, int y,
                                                int l) {
  const uint8_t* row = plane + (size_t)clampi(y + (l >> 2), 0, h - 1) * stride;
  const int c0 = x + 4 * (l & 3);
  if (c0 >= 0 && c0 + 4 <= w) {
    return (uint32_t)row[c0] | ((uint32_t)row[c0 + 1] << 8) | ((uint32_t)row[c0 + 2] << 16) |
           ((uint32_t)row[c0 + 3] << 24);
  }
  uint32_t v = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) v |= (uint32_t)row[clampi(c0 + b, 0, w - 1)] << (8 * b);
  return v;
}

__device__ __forceinline__ uint32_t wave_total(uint32_t v) { return wave_sum_to_sgpr(sum16(v)); }

// MvCostTable::bits (src/encoder/inter_search.rs:180-193): false when the
// differential leaves [-1023, 1023].
__device__ __forceinline__ bool cost_bits(const float* bits, int row, int col, int brow, int bcol, double* out) {
  const int dr = row - brow, dc = col - bcol;
  if (dr < -1023 || dr > 1023 || dc < -1023 || dc > 1023) return false;
  *out = (double)__fadd_rn(bits[dr + 1023], bits[2047 + dc + 1023]);
  return true;
}

struct Best {
  int row, col;  // quarter-pel
  uint32_t sad;
  double cost;
};

__device__ __forceinline__ void store_found(vp8g_found* o, const Best& b) {
  o->mv.row = (int16_t)b.row;
  o->mv.col = (int16_t)b.col;
  o->sad = b.sad;
  o->cost = b.cost;
}

// The job as the search reads it: device data is sanitised so that no
// input can make the kernel misbehave (the host entry points reject these
// values before any launch).
struct Job {
  int ox, oy;          // macroblock origin in pixels
  int brow, bcol;      // census.best
  int rx0, rx1, ry0, ry1;
  int n_seeds;
  double lambda;
  uint32_t good_enough;
  bool descend;
};

__device__ __forceinline__ Job load_job(const vp8g_mb_search* jp) {
  Job j;
  j.ox = (int)min(jp->mb_col, 0xFFFFFu) * 16;
  j.oy = (int)min(jp->mb_row, 0xFFFFFu) * 16;
  j.brow = jp->best.row;
  j.bcol = jp->best.col;
  j.rx0 = clampi(jp->range_x[0], -8191, 8191);
  j.rx1 = clampi(jp->range_x[1], j.rx0, 8191);
  j.ry0 = clampi(jp->range_y[0], -8191, 8191);
  j.ry1 = clampi(jp->range_y[1], j.ry0, 8191);
  j.n_seeds = (int)min(jp->n_seeds, 8u);
  j.lambda = jp->lambda_sad;
  j.good_enough = min(jp->good_enough, 0x3FFFFFFFu);
  j.descend = jp->descend != 0;
  return j;
}

// One probe of whole_pixel_search's `eval` closure (src/encoder/inter_search.rs:394-414)
// with the candidate's full SAD.
__device__ __forceinline__ bool wps_accept(const float* bits, const Job& j, int px, int py, uint32_t sad, Best* best) {
  const int row = (int16_t)(py * 4), col = (int16_t)(px * 4);
  double b;
  if (!cost_bits(bits, row, col, j.brow, j.bcol, &b)) return false;
  const double rate = __dmul_rn(j.lambda, b);
  if (rate >= best->cost) return false;
  const double cost = __dadd_rn((double)sad, rate);
  if (!(cost < best->cost)) return false;
  *best = Best{row, col, sad, cost};
  return true;
}

}  // namespace

// whole_pixel_search, one wave per macroblock.
__global__ void __launch_bounds__(256) whole_pixel_search_kernel(const WpsArgs a) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n) return;
  const Job j = load_job(a.jobs + i);
  const uint32_t sv = reinterpret_cast<const uint32_t*>(a.src + 256 * (size_t)i)[l];
  auto sad_at = [&](int px, int py) -> uint32_t {
    return sad4(sv, block_bytes(a.plane, a.stride, a.w, a.h, j.ox + px, j.oy + py, l), 0);
  };
  Best best{0, 0, 0xFFFFFFFFu, __builtin_inf()};
  // Seeds: rounded to whole pixels and clamped into the range, each
  // distinct position probed once, in order.
  {
    int px[8], py[8], m = 0;
    for (int k = 0; k < j.n_seeds; k++) {
      const vp8g_mv sd = a.jobs[i].seeds[k];
      const int x = clampi(((int)sd.col + 2) >> 2, j.rx0, j.rx1), y = clampi(((int)sd.row + 2) >> 2, j.ry0, j.ry1);
      bool seen = false;
      for (int t = 0; t < m; t++) seen |= (px[t] == x && py[t] == y);
      if (!seen) {
        px[m] = x;
        py[m] = y;
        m++;
      }
    }
    uint32_t part[8];
#pragma unroll
    for (int t = 0; t < 8; t++) part[t] = t < m ? sad_at(px[t], py[t]) : 0;
    for (int t = 0; t < m; t++) wps_accept(a.bits, j, px[t], py[t], wave_total(part[t]), &best);
  }
  if (!j.descend || best.sad <= j.good_enough) {
    if (l == 0) store_found(a.out + i, best);
    return;
  }
  const int wdx[8] = {-2, 2, 0, 0, -1, 1, -1, 1}, wdy[8] = {0, 0, -2, 2, -1, -1, 1, 1};
  const int ndx[4] = {-1, 1, 0, 0}, ndy[4] = {0, 0, -1, 1};
  auto in_range = [&](int x, int y) { return x >= j.rx0 && x <= j.rx1 && y >= j.ry0 && y <= j.ry1; };
  if (a.wide_pattern && best.sad > 4 * j.good_enough) {
    for (uint32_t round = 0; round < a.descent_rounds; round++) {
      const int cx = best.col / 4, cy = best.row / 4;
      uint32_t part[8];
#pragma unroll
      for (int t = 0; t < 8; t++) part[t] = in_range(cx + wdx[t], cy + wdy[t]) ? sad_at(cx + wdx[t], cy + wdy[t]) : 0;
      bool improved = false;
#pragma unroll
      for (int t = 0; t < 8; t++)
        if (in_range(cx + wdx[t], cy + wdy[t]))
          improved |= wps_accept(a.bits, j, cx + wdx[t], cy + wdy[t], wave_total(part[t]), &best);
      if (!improved || best.sad <= j.good_enough) break;
    }
  }
  for (uint32_t round = 0; round < a.descent_rounds; round++) {
    if (best.sad <= j.good_enough) break;
    const int cx = best.col / 4, cy = best.row / 4;
    uint32_t part[4];
#pragma unroll
    for (int t = 0; t < 4; t++) part[t] = in_range(cx + ndx[t], cy + ndy[t]) ? sad_at(cx + ndx[t], cy + ndy[t]) : 0;
    bool improved = false;
#pragma unroll
    for (int t = 0; t < 4; t++)
      if (in_range(cx + ndx[t], cy + ndy[t]))
        improved |= wps_accept(a.bits, j, cx + ndx[t], cy + ndy[t], wave_total(part[t]), &best);
    if (!improved) break;
  }
  if (l == 0) store_found(a.out + i, best);
}

namespace {

// interp (src/motion_comp.rs:396-402): clamp255((sum s_k f_k + 64) >> 7).
__device__ __forceinline__ uint32_t interp6(const int32_t* f, const uint8_t* s, int step) {
  int acc = 64;
#pragma unroll
  for (int k = 0; k < 6; k++) acc += (int)s[k * step] * f[k];
  return (uint32_t)clampi(acc >> 7, 0, 255);
}

// SAD of the lane's 4 source bytes against a plane held in LDS (row pitch
// `pitch`, the 16x16 block's top-left at `off`).
__device__ __forceinline__ uint32_t lds_part(uint32_t sv, const uint8_t* plane, int pitch, int off, int l) {
  const uint8_t* p = plane + off + (l >> 2) * pitch + 4 * (l & 3);
  const uint32_t v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
  return sad4(sv, v, 0);
}

__device__ __forceinline__ bool spr_consider(const float* bits, const Job& j, int row, int col, uint32_t sad,
                                             Best* best) {
  double b;
  if (!cost_bits(bits, row, col, j.brow, j.bcol, &b)) return false;
  const double cost = __dadd_rn((double)sad, __dmul_rn(j.lambda, b));
  if (!(cost < best->cost)) return false;
  *best = Best{row, col, sad, cost};
  return true;
}

}  // namespace

// sub_pixel_refine, one wave per macroblock, 4 waves per workgroup.  The
// wave stages a clamped 24x24 window at (x - 4, y - 4) around the start
// position (x, y): its inner 22x22 is fetch_window's (rows and columns
// -3..+18), and it holds the 21x21 six-tap support of every quarter-pel
// probe (their integer offsets lie within [-1, 0] of the start whatever the
// start's fraction, since start and probe differ by at most 3 quarter
// pixels per axis).  HalfPlanes' three planes are filtered into LDS, the
// half-pel probes are scored from them, then the quarter-pel probes from
// the two-pass six-tap prediction of the window.
__global__ void __launch_bounds__(256) sub_pixel_refine_kernel(const SprArgs a) {
  constexpr int WP = 24, HP = 17;
  __shared__ uint8_t s_win[4][WP * WP];
  __shared__ uint8_t s_h[4][22 * HP];
  __shared__ uint8_t s_v[4][HP * 16];
  __shared__ uint8_t s_d[4][HP * HP];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n) return;
  const Job j = load_job(a.jobs + i);
  const vp8g_found st = a.start[i];
  Best best{st.mv.row, st.mv.col, st.sad, st.cost};
  const int px = st.mv.col / 4, py = st.mv.row / 4;
  const int x0 = j.ox + px - 4, y0 = j.oy + py - 4;
  uint8_t* win = s_win[wave];
  for (int s = l; s < WP * WP; s += 64) {
    const int r = s / WP, c = s - r * WP;
    win[s] = a.plane[(size_t)clampi(y0 + r, 0, a.h - 1) * a.stride + clampi(x0 + c, 0, a.w - 1)];
  }
  const uint32_t sv = reinterpret_cast<const uint32_t*>(a.src + 256 * (size_t)i)[l];
  wave_lds_sync();
  // HalfPlanes::new (src/encoder/inter_search.rs:303-330) on the window
  // W[r][c] = win[(r + 1) * 24 + c + 1].
  const uint8_t* W = win + WP + 1;
  uint8_t *hpl = s_h[wave], *vpl = s_v[wave], *dpl = s_d[wave];
  const int h0 = a.half_diagonals ? 0 : 3, h1 = a.half_diagonals ? 22 : 19;
  for (int s = l; s < (h1 - h0) * HP; s += 64) {
    const int r = h0 + s / HP, c = s % HP;
    hpl[r * HP + c] = (uint8_t)interp6(a.taps, W + r * WP + c, 1);
  }
  for (int s = l; s < HP * 16; s += 64) {
    const int r = s / 16, c = s % 16;
    vpl[s] = (uint8_t)interp6(a.taps, W + r * WP + 3 + c, WP);
  }
  wave_lds_sync();
  if (a.half_diagonals) {
    for (int s = l; s < HP * HP; s += 64) {
      const int r = s / HP, c = s % HP;
      dpl[s] = (uint8_t)interp6(a.taps, hpl + r * HP + c, HP);
    }
    wave_lds_sync();
  }
  // Half-pel probes in (dy, dx) raster order; HalfPlanes::sad reads
  // position 0 for -1/2 and 1 for +1/2.
  uint32_t part[8];
  int hrow[8], hcol[8];
  bool use[8];
  {
    int k = 0;
    for (int dy = -1; dy <= 1; dy++) {
      for (int dx = -1; dx <= 1; dx++) {
        if (dx == 0 && dy == 0) continue;
        hrow[k] = (int16_t)(st.mv.row + 2 * dy);
        hcol[k] = (int16_t)(st.mv.col + 2 * dx);
        use[k] = (a.half_diagonals || dx == 0 || dy == 0) && abs(hrow[k]) <= 1020 && abs(hcol[k]) <= 1020;
        const int cx = dx > 0, cy = dy > 0;
        uint32_t p = 0;
        if (use[k]) {
          if (dy == 0) p = lds_part(sv, hpl, HP, 3 * HP + cx, l);
          else if (dx == 0) p = lds_part(sv, vpl, 16, cy * 16, l);
          else p = lds_part(sv, dpl, HP, cy * HP + cx, l);
        }
        part[k] = p;
        k++;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (use[k]) spr_consider(a.bits, j, hrow[k], hcol[k], wave_total(part[k]), &best);
  if (a.quarter) {
    // mb_luma_sad_at_mv (src/motion_search.rs:470-529) of the four
    // quarter-pel neighbours, always with the six-tap set.
    const int qdx[4] = {-1, 1, 0, 0}, qdy[4] = {0, 0, -1, 1};
    const int crow = best.row, ccol = best.col;
    uint32_t qpart[4];
    bool quse[4];
    int qrow[4], qcol[4];
    const int r = l >> 2, c0 = 4 * (l & 3);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      qrow[k] = (int16_t)(crow + qdy[k]);
      qcol[k] = (int16_t)(ccol + qdx[k]);
      quse[k] = abs(qrow[k]) <= 1022 && abs(qcol[k]) <= 1022;
      uint32_t p = 0;
      if (quse[k]) {
        const int r8 = (int16_t)(2 * qrow[k]), c8 = (int16_t)(2 * qcol[k]);  // stored_luma_mv
        const int mx = c8 & 7, my = r8 & 7;
        // halo origin (mb + (mv >> 3) - 2) in window coordinates
        const int hx = (c8 >> 3) - px + 2, hy = (r8 >> 3) - py + 2;
        const int32_t* fh = kSixtapTaps[mx];
        const int32_t* fv = kSixtapTaps[my];
        uint32_t pred = 0;
        uint8_t tmp[6][4];
#pragma unroll
        for (int t = 0; t < 6; t++) {
          const uint8_t* wr = win + (hy + r + t) * WP + hx + c0;
#pragma unroll
          for (int q = 0; q < 4; q++) tmp[t][q] = (uint8_t)interp6(fh, wr + q, 1);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          int acc = 64;
#pragma unroll
          for (int t = 0; t < 6; t++) acc += (int)tmp[t][q] * fv[t];
          pred |= (uint32_t)clampi(acc >> 7, 0, 255) << (8 * q);
        }
        p = sad4(sv, pred, 0);
      }
      qpart[k] = p;
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (quse[k]) spr_consider(a.bits, j, qrow[k], qcol[k], wave_total(qpart[k]), &best);
  }
  if (l == 0) store_found(a.out + i, best);
}

// small_diamond_search_luma, one wave per macroblock.
__global__ void __launch_bounds__(256) small_diamond_kernel(const DiamondArgs a) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= a.n) return;
  const int ox = (int)min(a.mb_xy[2 * i], 0xFFFFFu) * 16, oy = (int)min(a.mb_xy[2 * i + 1], 0xFFFFFu) * 16;
  const uint32_t sv = reinterpret_cast<const uint32_t*>(a.src + 256 * (size_t)i)[l];
  // snap_to_whole_pixel (truncating toward zero) after clamp_component
  auto snap = [](int v) { return (clampi(v, -1023, 1023) / 4) * 4; };
  // mb_luma_sad_at_whole_mv: integer offset (2 mv) >> 3 = mv / 4 for a whole-pel vector
  auto sad_part = [&](int row, int col) {
    return sad4(sv, block_bytes(a.plane, a.stride, a.w, a.h, ox + (col >> 2), oy + (row >> 2), l), 0);
  };
  int brow = snap(a.center[2 * i]), bcol = snap(a.center[2 * i + 1]);
  uint32_t bsad = wave_total(sad_part(brow, bcol));
  const int nr[4] = {-4, 4, 0, 0}, nc[4] = {0, 0, -4, 4};
  for (uint32_t it = 0; it < a.max_iters; it++) {
    // each probe is taken around the current best, which an improvement
    // moves for the probes after it (the reference's sequential scan)
    bool improved = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int cr = snap(brow + nr[k]), cc = snap(bcol + nc[k]);
      if (cr == brow && cc == bcol) continue;
      const uint32_t s = wave_total(sad_part(cr, cc));
      if (s < bsad) {
        bsad = s;
        brow = cr;
        bcol = cc;
        improved = true;
      }
    }
    if (!improved) break;
  }
  if (l == 0) {
    a.out[2 * i] = (uint32_t)(uint16_t)(int16_t)brow | ((uint32_t)(uint16_t)(int16_t)bcol << 16);
    a.out[2 * i + 1] = bsad;
  }
}

hipError_t launch_whole_pixel_search(const WpsArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(whole_pixel_search_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_sub_pixel_refine(const SprArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(sub_pixel_refine_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_small_diamond(const DiamondArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(small_diamond_kernel, dim3((a.n + 3) / 4), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace vp8g
