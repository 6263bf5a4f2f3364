// C ABI of the section 15 loop filter (include/vp8g.h): whole-frame
// filtering in place, candidate scoring, and the encoder's level /
// sharpness search.  Kernels in lf_kernels.hip.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/vp8g.h"
#include "vp8g_ctx.h"
#include "vp8g_internal.h"

using namespace vp8g;

namespace {

constexpr int kLfSsdSlots = 2 * 64;  // candidate SSDs of one search stage
constexpr int kSpinLimit = 1 << 22;  // polls (each ~2 x 64 cycles) before a row wait gives up

bool ok(hipError_t e) { return e == hipSuccess; }
size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

template <class F>
int run(const vp8g_ctx* c, F&& launch) {
  if (c->fault == VP8G_FAULT_LAUNCH) return VP8G_E_DEVICE;
  return launch() == hipSuccess ? VP8G_OK : VP8G_E_DEVICE;
}

int grow_device(vp8g_ctx* c, uint8_t** buf, size_t* have, size_t bytes) {
  if (c->fault == VP8G_FAULT_ALLOC) return VP8G_E_DEVICE;
  if (bytes <= *have) return VP8G_OK;
  if (*buf) {
    if (!ok(hipStreamSynchronize(c->stream))) return VP8G_E_DEVICE;
    (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
  }
  const size_t want = bytes + (bytes >> 2);
  if (!ok(hipMalloc(reinterpret_cast<void**>(buf), want))) return VP8G_E_DEVICE;
  *have = want;
  return VP8G_OK;
}

// The flag and the SSD slots, allocated once per context.
int ensure_small(vp8g_ctx* c) {
  if (c->fault == VP8G_FAULT_ALLOC) return VP8G_E_DEVICE;
  if (c->lf_err) return VP8G_OK;
  if (!ok(hipMalloc(reinterpret_cast<void**>(&c->lf_ssd), kLfSsdSlots * sizeof(unsigned long long))))
    return VP8G_E_DEVICE;
  if (!ok(hipMalloc(reinterpret_cast<void**>(&c->lf_err), 4))) return VP8G_E_DEVICE;
  return ok(launch_zero_u32(reinterpret_cast<uint32_t*>(c->lf_err), 1, c->stream)) ? VP8G_OK : VP8G_E_DEVICE;
}

// Device copy of host inputs plus its pinned mirror (same offsets).
int ensure_io(vp8g_ctx* c, size_t bytes) {
  if (c->fault == VP8G_FAULT_ALLOC) return VP8G_E_DEVICE;
  if (bytes <= c->lf_io_bytes) return VP8G_OK;
  if (c->lf_io) {
    if (!ok(hipStreamSynchronize(c->stream))) return VP8G_E_DEVICE;
    (void)hipFree(c->lf_io);
    (void)hipHostFree(c->lf_pinned);
    c->lf_io = c->lf_pinned = nullptr;
    c->lf_io_bytes = 0;
  }
  const size_t want = bytes + (bytes >> 2);
  if (!ok(hipMalloc(reinterpret_cast<void**>(&c->lf_io), want))) return VP8G_E_DEVICE;
  if (!ok(hipHostMalloc(reinterpret_cast<void**>(&c->lf_pinned), want, hipHostMallocDefault))) {
    (void)hipFree(c->lf_io);
    c->lf_io = nullptr;
    return VP8G_E_DEVICE;
  }
  c->lf_io_bytes = want;
  return VP8G_OK;
}

int check_planes(const vp8g_lf_planes* p, bool device) {
  if (!p) return VP8G_E_ARG;
  if (p->mb_cols == 0 || p->mb_rows == 0) return VP8G_OK;
  if (!p->y || !p->u || !p->v || p->mb_cols > 1024 || p->mb_rows > 1024) return VP8G_E_ARG;
  if (p->y_stride < p->mb_cols * 16 || p->uv_stride < p->mb_cols * 8) return VP8G_E_ARG;
  if (p->y_stride > 0x7FFFFFFF || p->uv_stride > 0x7FFFFFFF) return VP8G_E_ARG;
  // the kernels address a plane with 32-bit offsets
  if (p->y_stride * p->mb_rows * 16 > 0xFFFFFFFFull || p->uv_stride * p->mb_rows * 8 > 0xFFFFFFFFull) return VP8G_E_ARG;
  if (device && (((uintptr_t)p->y | p->y_stride) & 15)) return VP8G_E_ARG;
  if (device && (((uintptr_t)p->u | (uintptr_t)p->v | p->uv_stride) & 7)) return VP8G_E_ARG;
  return VP8G_OK;
}

int check_cfg(const vp8g_lf_config* g) {
  if (!g) return VP8G_E_ARG;
  if (g->loop_filter_level < 0 || g->loop_filter_level > 255) return VP8G_E_ARG;
  if (g->sharpness_level < 0 || g->sharpness_level > 255) return VP8G_E_ARG;
  return VP8G_OK;
}

int check_source(const vp8g_lf_source* s, const vp8g_lf_planes* p) {
  if (!s) return VP8G_E_ARG;
  if (s->width > p->mb_cols * 16 || s->height > p->mb_rows * 16) return VP8G_E_ARG;
  if (s->width == 0 || s->height == 0) return VP8G_OK;
  if (!s->y || !s->u || !s->v || s->y_stride < s->width || s->uv_stride < (s->width + 1) / 2) return VP8G_E_ARG;
  if (s->y_stride > 0x7FFFFFFF || s->uv_stride > 0x7FFFFFFF) return VP8G_E_ARG;
  if (s->y_stride * s->height > 0xFFFFFFFFull || s->uv_stride * ((s->height + 1) / 2) > 0xFFFFFFFFull)
    return VP8G_E_ARG;
  return VP8G_OK;
}

bool empty(const vp8g_lf_planes* p) { return p->mb_cols == 0 || p->mb_rows == 0; }

LfArgs base_args(vp8g_ctx* c, const vp8g_lf_planes* p, const vp8g_lf_mb* mbs, const vp8g_lf_config* g) {
  LfArgs a;
  memset(&a, 0, sizeof(a));
  a.y = p->y, a.u = p->u, a.v = p->v;
  a.y_stride = p->y_stride, a.uv_stride = p->uv_stride;
  a.mb_cols = (int)p->mb_cols, a.mb_rows = (int)p->mb_rows;
  a.mbs = reinterpret_cast<const uint32_t*>(mbs);
  a.key_frame = g->key_frame != 0;
  a.segmentation_enabled = g->segmentation_enabled != 0;
  a.segment_abs = g->segment_abs != 0;
  a.delta_enabled = g->delta_enabled != 0;
  for (int i = 0; i < 4; i++) {
    a.segment_lf_level[i] = g->segment_lf_level[i];
    a.ref_delta[i] = g->ref_delta[i];
    a.mode_delta[i] = g->mode_delta[i];
  }
  a.err = c->lf_err;
  a.spin_limit = kSpinLimit;
  return a;
}

int waves_for(int rows) { return std::min(kLfMaxWaves, (rows + 1) / 2); }

int enqueue_frame(vp8g_ctx* c, const vp8g_lf_planes* p, const vp8g_lf_mb* mbs, const vp8g_lf_config* g) {
  int st = ensure_small(c);
  if (st != VP8G_OK) return st;
  const int rows = (int)p->mb_rows, waves = waves_for(rows);
  const size_t ring = rows > 2 ? lf_ring_bytes(waves, (int)p->mb_cols) : 0;
  st = grow_device(c, &c->lf_work, &c->lf_work_bytes, std::max<size_t>(ring, 256));
  if (st != VP8G_OK) return st;
  LfArgs a = base_args(c, p, mbs, g);
  a.oy = p->y, a.ou = p->u, a.ov = p->v;
  a.n_bands = 1, a.every_rows = 0, a.band_rows = rows;
  a.cand_level[0] = (uint8_t)g->loop_filter_level;
  a.cand_sharp[0] = (uint8_t)g->sharpness_level;
  a.ring = c->lf_work;
  a.ring_instance_bytes = ring;
  return run(c, [&] { return launch_loop_filter(a, 1, waves, g->simple != 0, false, c->stream); });
}

int enqueue_score(vp8g_ctx* c, const vp8g_lf_planes* p, const vp8g_lf_mb* mbs, const vp8g_lf_config* g,
                  const vp8g_lf_source* s, const uint8_t* levels, const uint8_t* sharps, size_t n, size_t every_rows,
                  unsigned long long* ssd) {
  int st = ensure_small(c);
  if (st != VP8G_OK) return st;
  const int R = (int)p->mb_rows;
  const int every = every_rows == 0 ? 0 : (int)std::max<size_t>(std::min<size_t>(every_rows, 1024), 2);
  const int n_bands = every ? (R + every - 1) / every : 1, band_rows = every ? 2 : R;
  const int waves = waves_for(band_rows);
  const size_t ring = band_rows > 2 ? lf_ring_bytes(waves, (int)p->mb_cols) : 0;
  const size_t inst = n * (size_t)n_bands;
  const size_t o_ssd = up256(ring * inst), total = o_ssd + 8 * inst;
  st = grow_device(c, &c->lf_work, &c->lf_work_bytes, total);
  if (st != VP8G_OK) return st;
  LfArgs a = base_args(c, p, mbs, g);
  a.sy = s->y, a.su = s->u, a.sv = s->v;
  a.sy_stride = s->y_stride, a.suv_stride = s->uv_stride;
  a.vis_w = (int)s->width, a.vis_h = (int)s->height;
  a.vis_cw = (int)((s->width + 1) / 2), a.vis_ch = (int)((s->height + 1) / 2);
  a.src_aligned4 = !(((uintptr_t)s->y | (uintptr_t)s->u | (uintptr_t)s->v | s->y_stride | s->uv_stride) & 3);
  a.n_bands = n_bands, a.every_rows = every, a.band_rows = band_rows;
  for (size_t i = 0; i < n; i++) a.cand_level[i] = levels[i], a.cand_sharp[i] = sharps[i];
  a.ring = c->lf_work;
  a.ring_instance_bytes = ring;
  a.inst_ssd = reinterpret_cast<unsigned long long*>(c->lf_work + o_ssd);
  st = run(c, [&] { return launch_loop_filter(a, (int)inst, waves, g->simple != 0, true, c->stream); });
  if (st != VP8G_OK) return st;
  return run(c, [&] { return launch_lf_sum_bands(a.inst_ssd, (int)n, n_bands, ssd, c->stream); });
}

// Waits for the stream; a set row-wait flag fails the call (and is cleared).
int sync_and_check(vp8g_ctx* c, int st) {
  int flag = 0;
  if (st == VP8G_OK && !ok(hipMemcpyAsync(&flag, c->lf_err, 4, hipMemcpyDeviceToHost, c->stream))) st = VP8G_E_DEVICE;
  if (!ok(hipStreamSynchronize(c->stream)) && st == VP8G_OK) st = VP8G_E_DEVICE;
  if (st == VP8G_OK && flag) {
    (void)launch_zero_u32(reinterpret_cast<uint32_t*>(c->lf_err), 1, c->stream);
    (void)hipStreamSynchronize(c->stream);
    st = VP8G_E_DEVICE;
  }
  return st;
}

// Scores on the device, read back (synchronous).
int score_to_host(vp8g_ctx* c, const vp8g_lf_planes* p, const vp8g_lf_mb* mbs, const vp8g_lf_config* g,
                  const vp8g_lf_source* s, const uint8_t* levels, const uint8_t* sharps, size_t n, size_t every_rows,
                  unsigned long long* out) {
  int st = ensure_small(c);
  if (st == VP8G_OK) st = enqueue_score(c, p, mbs, g, s, levels, sharps, n, every_rows, c->lf_ssd);
  if (st == VP8G_OK && c->fault == VP8G_FAULT_COPY) st = VP8G_E_DEVICE;
  if (st == VP8G_OK && !ok(hipMemcpyAsync(out, c->lf_ssd, 8 * n, hipMemcpyDeviceToHost, c->stream)))
    st = VP8G_E_DEVICE;
  return sync_and_check(c, st);
}

// select_filter_params_threaded (src/loop_filter.rs:2133-2200) and
// select_filter_params_sampled (src/loop_filter.rs:2288-2412).  Every level
// the level stages can visit is scored in one launch (the picks only read
// the scores of the candidates the reference evaluates, in its order, so
// the extra ones change nothing); the sharpness sweep is a second launch.
int select_device(vp8g_ctx* c, const vp8g_lf_planes* p, const vp8g_lf_mb* mbs, const vp8g_lf_config* g,
                  const vp8g_lf_source* s, size_t every_rows, int sweep, uint8_t* level, uint8_t* sharpness) {
  const uint8_t bs = (uint8_t)std::min(g->sharpness_level, 7);
  if (empty(p)) {
    *level = 0;
    *sharpness = bs;
    return VP8G_OK;
  }
  uint8_t lv[64], sh[64];
  unsigned long long sc[64];
  for (int l = 0; l < 64; l++) lv[l] = (uint8_t)l, sh[l] = bs;
  int st = score_to_host(c, p, mbs, g, s, lv, sh, 64, every_rows, sc);
  if (st != VP8G_OK) return st;
  int best;
  unsigned long long best_ssd;
  if (every_rows == 0) {
    // coarse grid 0, 4, .., 60 plus 63, then +-3 around the winner
    best = 0, best_ssd = sc[0];
    for (int l = 4; l <= 60; l += 4)
      if (sc[l] < best_ssd) best_ssd = sc[l], best = l;
    if (sc[63] < best_ssd) best_ssd = sc[63], best = 63;
    const int lo = std::max(best - 3, 0), hi = std::min(best + 3, 63), centre = best;
    for (int l = lo; l <= hi; l++)
      if (l != centre && sc[l] < best_ssd) best_ssd = sc[l], best = l;
    sweep = best != 0;
  } else {
    // every eighth level plus 63, then +-4, +-2, +-1 around the running best
    best = 0, best_ssd = sc[0];
    for (int l = 8; l <= 56; l += 8)
      if (sc[l] < best_ssd) best_ssd = sc[l], best = l;
    if (sc[63] < best_ssd) best_ssd = sc[63], best = 63;
    for (int stp : {4, 2, 1}) {
      const int centre = best;
      if (centre >= stp && sc[centre - stp] < best_ssd) best_ssd = sc[centre - stp], best = centre - stp;
      if (centre + stp <= 63 && sc[centre + stp] < best_ssd) best_ssd = sc[centre + stp], best = centre + stp;
    }
    sweep = sweep && best != 0;
  }
  uint8_t best_sharp = bs;
  if (sweep) {
    int m = 0;
    for (int q = 0; q <= 7; q++)
      if (q != bs) lv[m] = (uint8_t)best, sh[m++] = (uint8_t)q;
    st = score_to_host(c, p, mbs, g, s, lv, sh, (size_t)m, every_rows, sc);
    if (st != VP8G_OK) return st;
    for (int i = 0; i < m; i++)
      if (sc[i] < best_ssd) best_ssd = sc[i], best_sharp = sh[i];
  }
  *level = (uint8_t)best;
  *sharpness = best_sharp;
  return VP8G_OK;
}

// Host planes -> lf_io: each plane packed at stride `width` (rounded up to
// 16), offsets 256-aligned.
struct IoPlan {
  size_t ys, cs, o_y, o_u, o_v, o_mb, o_sy, o_su, o_sv, sys, scs, total;
};

IoPlan plan_io(const vp8g_lf_planes* p, const vp8g_lf_source* s) {
  IoPlan q{};
  q.ys = p->mb_cols * 16, q.cs = p->mb_cols * 8;
  const size_t yb = q.ys * p->mb_rows * 16, cb = q.cs * p->mb_rows * 8;
  q.o_y = 0, q.o_u = up256(yb), q.o_v = q.o_u + up256(cb), q.o_mb = q.o_v + up256(cb);
  q.total = q.o_mb + up256(4 * p->mb_cols * p->mb_rows);
  if (s && s->width && s->height) {
    const size_t cw = (s->width + 1) / 2, ch = (s->height + 1) / 2;
    q.sys = (s->width + 15) & ~(size_t)15, q.scs = (cw + 15) & ~(size_t)15;
    q.o_sy = q.total, q.o_su = q.o_sy + up256(q.sys * s->height), q.o_sv = q.o_su + up256(q.scs * ch);
    q.total = q.o_sv + up256(q.scs * ch);
  }
  return q;
}

void pack(uint8_t* dst, size_t dst_stride, const uint8_t* src, size_t src_stride, size_t width, size_t rows) {
  for (size_t r = 0; r < rows; r++) memcpy(dst + r * dst_stride, src + r * src_stride, width);
}

// Stage host recon (+ source) and mbs into lf_io; device views in *dp / *ds.
int stage_in(vp8g_ctx* c, const vp8g_lf_planes* p, const vp8g_lf_mb* mbs, const vp8g_lf_source* s, IoPlan* q,
             vp8g_lf_planes* dp, vp8g_lf_source* ds) {
  *q = plan_io(p, s);
  int st = ensure_io(c, q->total);
  if (st != VP8G_OK) return st;
  if (c->fault == VP8G_FAULT_COPY) return VP8G_E_DEVICE;
  uint8_t* h = c->lf_pinned;
  const size_t R = p->mb_rows;
  pack(h + q->o_y, q->ys, p->y, p->y_stride, q->ys, R * 16);
  pack(h + q->o_u, q->cs, p->u, p->uv_stride, q->cs, R * 8);
  pack(h + q->o_v, q->cs, p->v, p->uv_stride, q->cs, R * 8);
  memcpy(h + q->o_mb, mbs, 4 * p->mb_cols * R);
  if (s && s->width && s->height) {
    const size_t cw = (s->width + 1) / 2, ch = (s->height + 1) / 2;
    pack(h + q->o_sy, q->sys, s->y, s->y_stride, s->width, s->height);
    pack(h + q->o_su, q->scs, s->u, s->uv_stride, cw, ch);
    pack(h + q->o_sv, q->scs, s->v, s->uv_stride, cw, ch);
  }
  if (!ok(hipMemcpyAsync(c->lf_io, h, q->total, hipMemcpyHostToDevice, c->stream))) return VP8G_E_DEVICE;
  uint8_t* d = c->lf_io;
  *dp = *p;
  dp->y = d + q->o_y, dp->u = d + q->o_u, dp->v = d + q->o_v;
  dp->y_stride = q->ys, dp->uv_stride = q->cs;
  if (ds && s) {
    *ds = *s;
    if (s->width && s->height) {
      ds->y = d + q->o_sy, ds->u = d + q->o_su, ds->v = d + q->o_sv;
      ds->y_stride = q->sys, ds->uv_stride = q->scs;
    }
  }
  return VP8G_OK;
}

}  // namespace

extern "C" {

int vp8g_loop_filter_frame_dev(vp8g_ctx* c, const vp8g_lf_planes* planes, const vp8g_lf_mb* mbs,
                               const vp8g_lf_config* cfg) {
  if (!c) return VP8G_E_ARG;
  int st = check_planes(planes, true);
  if (st == VP8G_OK) st = check_cfg(cfg);
  if (st != VP8G_OK) return st;
  if (empty(planes)) return VP8G_OK;
  if (!mbs || ((uintptr_t)mbs & 3)) return VP8G_E_ARG;
  if (!ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  return enqueue_frame(c, planes, mbs, cfg);
}

int vp8g_loop_filter_frame(vp8g_ctx* c, const vp8g_lf_planes* planes, const vp8g_lf_mb* mbs,
                           const vp8g_lf_config* cfg) {
  if (!c) return VP8G_E_ARG;
  int st = check_planes(planes, false);
  if (st == VP8G_OK) st = check_cfg(cfg);
  if (st != VP8G_OK) return st;
  if (empty(planes)) return VP8G_OK;
  if (!mbs) return VP8G_E_ARG;
  if (!ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  IoPlan q;
  vp8g_lf_planes dp;
  st = stage_in(c, planes, mbs, nullptr, &q, &dp, nullptr);
  if (st == VP8G_OK) st = enqueue_frame(c, &dp, reinterpret_cast<const vp8g_lf_mb*>(c->lf_io + q.o_mb), cfg);
  if (st == VP8G_OK && !ok(hipMemcpyAsync(c->lf_pinned, c->lf_io, q.o_mb, hipMemcpyDeviceToHost, c->stream)))
    st = VP8G_E_DEVICE;
  st = sync_and_check(c, st);
  if (st != VP8G_OK) return st;
  const size_t R = planes->mb_rows;
  pack(planes->y, planes->y_stride, c->lf_pinned + q.o_y, q.ys, q.ys, R * 16);
  pack(planes->u, planes->uv_stride, c->lf_pinned + q.o_u, q.cs, q.cs, R * 8);
  pack(planes->v, planes->uv_stride, c->lf_pinned + q.o_v, q.cs, q.cs, R * 8);
  return VP8G_OK;
}

int vp8g_loop_filter_score_dev(vp8g_ctx* c, const vp8g_lf_planes* recon, const vp8g_lf_mb* mbs,
                               const vp8g_lf_config* cfg, const vp8g_lf_source* src, const uint8_t* levels,
                               const uint8_t* sharps, size_t n, size_t every_rows, uint64_t* ssd) {
  if (!c) return VP8G_E_ARG;
  int st = check_planes(recon, true);
  if (st == VP8G_OK) st = check_cfg(cfg);
  if (st == VP8G_OK) st = check_source(src, recon);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (n > (size_t)kLfMaxCands || !levels || !sharps || !ssd || ((uintptr_t)ssd & 7)) return VP8G_E_ARG;
  if (!ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  if (empty(recon))
    return run(c, [&] { return launch_zero_u32(reinterpret_cast<uint32_t*>(ssd), (int)(2 * n), c->stream); });
  if (!mbs || ((uintptr_t)mbs & 3)) return VP8G_E_ARG;
  return enqueue_score(c, recon, mbs, cfg, src, levels, sharps, n, every_rows,
                       reinterpret_cast<unsigned long long*>(ssd));
}

int vp8g_loop_filter_select_dev(vp8g_ctx* c, const vp8g_lf_planes* recon, const vp8g_lf_mb* mbs,
                                const vp8g_lf_config* cfg, const vp8g_lf_source* src, size_t every_rows,
                                int sweep_sharpness, uint8_t* level, uint8_t* sharpness) {
  if (!c || !level || !sharpness) return VP8G_E_ARG;
  int st = check_planes(recon, true);
  if (st == VP8G_OK) st = check_cfg(cfg);
  if (st == VP8G_OK) st = check_source(src, recon);
  if (st != VP8G_OK) return st;
  if (!empty(recon) && (!mbs || ((uintptr_t)mbs & 3))) return VP8G_E_ARG;
  if (!ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  return select_device(c, recon, mbs, cfg, src, every_rows, sweep_sharpness, level, sharpness);
}

int vp8g_loop_filter_select(vp8g_ctx* c, const vp8g_lf_planes* recon, const vp8g_lf_mb* mbs,
                            const vp8g_lf_config* cfg, const vp8g_lf_source* src, size_t every_rows,
                            int sweep_sharpness, uint8_t* level, uint8_t* sharpness) {
  if (!c || !level || !sharpness) return VP8G_E_ARG;
  int st = check_planes(recon, false);
  if (st == VP8G_OK) st = check_cfg(cfg);
  if (st == VP8G_OK) st = check_source(src, recon);
  if (st != VP8G_OK) return st;
  if (empty(recon)) return select_device(c, recon, mbs, cfg, src, every_rows, sweep_sharpness, level, sharpness);
  if (!mbs) return VP8G_E_ARG;
  if (!ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  IoPlan q;
  vp8g_lf_planes dp;
  vp8g_lf_source ds;
  st = stage_in(c, recon, mbs, src, &q, &dp, &ds);
  if (st != VP8G_OK) {
    (void)hipStreamSynchronize(c->stream);
    return st;
  }
  return select_device(c, &dp, reinterpret_cast<const vp8g_lf_mb*>(c->lf_io + q.o_mb), cfg, &ds, every_rows,
                       sweep_sharpness, level, sharpness);
}

int vp8g_loop_filter_status(vp8g_ctx* c) {
  if (!c) return VP8G_E_ARG;
  if (!ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  if (!c->lf_err) return ok(hipStreamSynchronize(c->stream)) ? VP8G_OK : VP8G_E_DEVICE;
  int flag = 0;
  if (!ok(hipMemcpyAsync(&flag, c->lf_err, 4, hipMemcpyDeviceToHost, c->stream))) return VP8G_E_DEVICE;
  if (!ok(hipStreamSynchronize(c->stream))) return VP8G_E_DEVICE;
  if (flag && !ok(launch_zero_u32(reinterpret_cast<uint32_t*>(c->lf_err), 1, c->stream))) return VP8G_E_DEVICE;
  return flag ? VP8G_E_DEVICE : VP8G_OK;
}

}  // extern "C"
