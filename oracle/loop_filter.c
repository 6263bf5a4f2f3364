/*
 * CPU oracle for the oxideav-vp8 RFC 6386 section 15 loop filter and the
 * encoder's filter-level search.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The
 * product library never links or calls this file.  Each function restates
 * the reference crate's integer arithmetic and cites the file:line it
 * follows (paths relative to the oxideav-vp8 crate root).  The reference
 * evaluates the normal filter over whole edge runs without branches
 * (src/loop_filter.rs:407-583) or in SIMD lanes; both are proved
 * byte-exact with the per-segment scalar kernels restated here
 * (src/loop_filter.rs:2640-2698, 3833-3957), so the scalar form is the
 * oracle.
 */
#define _GNU_SOURCE
#include "vp8_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

static inline int lf_clamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int lf_abs(int v) { return v < 0 ? -v : v; }

/* LoopFilterParams::derive, src/loop_filter.rs:99-148. */
orc_lf_params orc_lf_derive(uint8_t level, uint8_t sharpness, int key_frame)
{
    int il = level;
    if (sharpness != 0) {
        il >>= sharpness > 4 ? 2 : 1;
        const int cap = sharpness >= 9 ? 0 : 9 - sharpness;
        if (il > cap) il = cap;
    }
    if (il == 0) il = 1;
    int hev;
    if (key_frame)
        hev = level >= 40 ? 2 : (level >= 15 ? 1 : 0);
    else
        hev = level >= 40 ? 3 : (level >= 20 ? 2 : (level >= 15 ? 1 : 0));
    orc_lf_params p;
    p.interior_limit = (uint8_t)il;
    p.hev_threshold = (uint8_t)hev;
    const int mbe = (level + 2) * 2 + il, sbe = level * 2 + il;
    p.mbedge_limit = (uint8_t)(mbe > 255 ? 255 : mbe);
    p.sub_bedge_limit = (uint8_t)(sbe > 255 ? 255 : sbe);
    return p;
}

/* clamp_s8 / u2s / s2u, src/loop_filter.rs:157-180. */
static inline int c8(int v) { return lf_clamp(v, -128, 127); }
static inline int u2s(uint8_t v) { return (int)v - 128; }
static inline uint8_t s2u(int v) { return (uint8_t)(c8(v) + 128); }

/* common_adjust on seg[0..4] = p1 p0 q0 q1, src/loop_filter.rs:200-228. */
int orc_lf_common_adjust(int use_outer_taps, uint8_t *seg)
{
    const int p1 = u2s(seg[0]), p0 = u2s(seg[1]), q0 = u2s(seg[2]), q1 = u2s(seg[3]);
    const int outer = use_outer_taps ? c8(p1 - q1) : 0;
    int a = c8(outer + 3 * (q0 - p0));
    const int b = c8(a + 3) >> 3;
    a = c8(a + 4) >> 3;
    seg[2] = s2u(q0 - a);
    seg[1] = s2u(p0 + b);
    return a;
}

/* simple_segment on seg[0..4], src/loop_filter.rs:235-245. */
void orc_lf_simple_segment(uint8_t edge_limit, uint8_t *seg)
{
    const int p1 = seg[0], p0 = seg[1], q0 = seg[2], q1 = seg[3];
    if (lf_abs(p0 - q0) * 2 + lf_abs(p1 - q1) / 2 <= (int)edge_limit) orc_lf_common_adjust(1, seg);
}

/* filter_yes and hev, src/loop_filter.rs:252-281. */
static int filter_yes(int i, int e, const int t[8])
{
    return lf_abs(t[3] - t[4]) * 2 + lf_abs(t[2] - t[5]) / 2 <= e && lf_abs(t[0] - t[1]) <= i &&
           lf_abs(t[1] - t[2]) <= i && lf_abs(t[2] - t[3]) <= i && lf_abs(t[7] - t[6]) <= i &&
           lf_abs(t[6] - t[5]) <= i && lf_abs(t[5] - t[4]) <= i;
}
static int hev(int thr, const int t[8]) { return lf_abs(t[2] - t[3]) > thr || lf_abs(t[5] - t[4]) > thr; }

/* subblock_filter on seg[0..8] = p3 .. q3, src/loop_filter.rs:292-325. */
void orc_lf_subblock_filter(uint8_t hev_threshold, uint8_t interior_limit, uint8_t edge_limit, uint8_t *seg)
{
    int t[8];
    for (int k = 0; k < 8; k++) t[k] = u2s(seg[k]);
    if (!filter_yes(interior_limit, edge_limit, t)) return;
    const int hv = hev(hev_threshold, t);
    const int a = (orc_lf_common_adjust(hv, seg + 2) + 1) >> 1;
    if (!hv) {
        seg[5] = s2u(t[5] - a);
        seg[2] = s2u(t[2] + a);
    }
}

/* mb_filter on seg[0..8] = p3 .. q3, src/loop_filter.rs:331-393. */
void orc_lf_mb_filter(uint8_t hev_threshold, uint8_t interior_limit, uint8_t edge_limit, uint8_t *seg)
{
    int t[8];
    for (int k = 0; k < 8; k++) t[k] = u2s(seg[k]);
    if (!filter_yes(interior_limit, edge_limit, t)) return;
    if (!hev(hev_threshold, t)) {
        const int w = c8(c8(t[2] - t[5]) + 3 * (t[4] - t[3]));
        int a = c8((27 * w + 63) >> 7);
        seg[4] = s2u(t[4] - a);
        seg[3] = s2u(t[3] + a);
        a = c8((18 * w + 63) >> 7);
        seg[5] = s2u(t[5] - a);
        seg[2] = s2u(t[2] + a);
        a = c8((9 * w + 63) >> 7);
        seg[6] = s2u(t[6] - a);
        seg[1] = s2u(t[1] + a);
    } else {
        orc_lf_common_adjust(1, seg + 2);
    }
}

/* calculate_mb_filter_level_inter, src/loop_filter.rs:1280-1334; with
 * ref_frame 0 it is calculate_mb_filter_level, src/loop_filter.rs:1222-1265. */
uint8_t orc_lf_mb_level(const orc_lf_config *cfg, orc_lf_mb mb)
{
    int level = cfg->loop_filter_level;
    if (cfg->segmentation_enabled) {
        const int seg = mb.segment_id == ORC_LF_NO_SEGMENT ? 0 : (mb.segment_id > 3 ? 3 : mb.segment_id);
        if (cfg->segment_abs)
            level = cfg->segment_lf_level[seg];
        else
            level += cfg->segment_lf_level[seg];
    }
    level = lf_clamp(level, 0, 63);
    if (cfg->delta_enabled) {
        level += cfg->ref_delta[mb.ref_frame > 3 ? 3 : mb.ref_frame];
        if (mb.ref_frame == 0) {
            if (mb.flags & ORC_LF_BPRED) level += cfg->mode_delta[0];
        } else if (mb.inter_mode == ORC_LF_ZEROMV) {
            level += cfg->mode_delta[1];
        } else if (mb.inter_mode == ORC_LF_SPLITMV) {
            level += cfg->mode_delta[3];
        } else {
            level += cfg->mode_delta[2];
        }
    }
    return (uint8_t)lf_clamp(level, 0, 63);
}

/* filter_v_edge_normal / filter_h_edge_normal, src/loop_filter.rs:1515-1591
 * and 1650-1725 (scalar bodies): the 8-tap segment straddling the edge, p
 * side left of / above position (x, y). */
static void edge_normal(uint8_t *plane, size_t stride, size_t x, size_t y, int vertical, int len,
                        const orc_lf_params *p, int mb_edge)
{
    const uint8_t el = mb_edge ? p->mbedge_limit : p->sub_bedge_limit;
    uint8_t seg[8];
    for (int n = 0; n < len; n++) {
        uint8_t *at = vertical ? plane + (y + n) * stride + x - 4 : plane + (y - 4) * stride + x + n;
        const size_t step = vertical ? 1 : stride;
        for (int k = 0; k < 8; k++) seg[k] = at[k * step];
        if (mb_edge)
            orc_lf_mb_filter(p->hev_threshold, p->interior_limit, el, seg);
        else
            orc_lf_subblock_filter(p->hev_threshold, p->interior_limit, el, seg);
        for (int k = 0; k < 8; k++) at[k * step] = seg[k];
    }
}

/* filter_v_edge_simple / filter_h_edge_simple, src/loop_filter.rs:1777-1843. */
static void edge_simple(uint8_t *plane, size_t stride, size_t x, size_t y, int vertical, uint8_t edge_limit)
{
    uint8_t seg[4];
    for (int n = 0; n < 16; n++) {
        uint8_t *at = vertical ? plane + (y + n) * stride + x - 2 : plane + (y - 2) * stride + x + n;
        const size_t step = vertical ? 1 : stride;
        for (int k = 0; k < 4; k++) seg[k] = at[k * step];
        orc_lf_simple_segment(edge_limit, seg);
        for (int k = 0; k < 4; k++) at[k * step] = seg[k];
    }
}

/* One macroblock in the section 20.6 step order: left edge, inner vertical
 * edges, top edge, inner horizontal edges.  The body of the raster loop of
 * filter_inter_frame_flags, src/loop_filter.rs:1393-1494 (and of
 * filter_frame_flags, src/loop_filter.rs:2528-2633). */
static void filter_mb(orc_lf_planes *pl, size_t mb_col, size_t mb_row, const orc_lf_params *p, int subblocks,
                      int simple)
{
    const size_t ys = pl->y_stride, cs = pl->uv_stride;
    const size_t yx = mb_col * 16, yy = mb_row * 16, cx = mb_col * 8, cy = mb_row * 8;
    if (simple) {
        if (mb_col > 0) edge_simple(pl->y, ys, yx, yy, 1, p->mbedge_limit);
        if (subblocks)
            for (int q = 1; q < 4; q++) edge_simple(pl->y, ys, yx + 4 * q, yy, 1, p->sub_bedge_limit);
        if (mb_row > 0) edge_simple(pl->y, ys, yx, yy, 0, p->mbedge_limit);
        if (subblocks)
            for (int q = 1; q < 4; q++) edge_simple(pl->y, ys, yx, yy + 4 * q, 0, p->sub_bedge_limit);
        return;
    }
    if (mb_col > 0) {
        edge_normal(pl->y, ys, yx, yy, 1, 16, p, 1);
        edge_normal(pl->u, cs, cx, cy, 1, 8, p, 1);
        edge_normal(pl->v, cs, cx, cy, 1, 8, p, 1);
    }
    if (subblocks) {
        for (int q = 1; q < 4; q++) edge_normal(pl->y, ys, yx + 4 * q, yy, 1, 16, p, 0);
        edge_normal(pl->u, cs, cx + 4, cy, 1, 8, p, 0);
        edge_normal(pl->v, cs, cx + 4, cy, 1, 8, p, 0);
    }
    if (mb_row > 0) {
        edge_normal(pl->y, ys, yx, yy, 0, 16, p, 1);
        edge_normal(pl->u, cs, cx, cy, 0, 8, p, 1);
        edge_normal(pl->v, cs, cx, cy, 0, 8, p, 1);
    }
    if (subblocks) {
        for (int q = 1; q < 4; q++) edge_normal(pl->y, ys, yx, yy + 4 * q, 0, 16, p, 0);
        edge_normal(pl->u, cs, cx, cy + 4, 0, 8, p, 0);
        edge_normal(pl->v, cs, cx, cy + 4, 0, 8, p, 0);
    }
}

/* filter_inter_frame_flags, src/loop_filter.rs:1364-1496 (with every MB
 * intra and key_frame set it is filter_frame_flags,
 * src/loop_filter.rs:2506-2636).  In place, raster order. */
void orc_lf_filter_frame(orc_lf_planes *pl, const orc_lf_mb *mbs, const orc_lf_config *cfg)
{
    for (size_t r = 0; r < pl->mb_rows; r++)
        for (size_t c = 0; c < pl->mb_cols; c++) {
            const orc_lf_mb mb = mbs[r * pl->mb_cols + c];
            const uint8_t level = orc_lf_mb_level(cfg, mb);
            if (level == 0) continue;
            const orc_lf_params p = orc_lf_derive(level, (uint8_t)cfg->sharpness_level, cfg->key_frame);
            const int subblocks = (mb.flags & ORC_LF_BPRED) || (mb.flags & ORC_LF_HAS_COEFFS);
            filter_mb(pl, c, r, &p, subblocks, cfg->simple);
        }
}

/* plane_visible_ssd, src/loop_filter.rs:1941-1960. */
static uint64_t plane_ssd(const uint8_t *a, size_t as, const uint8_t *b, size_t bs, size_t w, size_t h)
{
    uint64_t s = 0;
    for (size_t r = 0; r < h; r++)
        for (size_t x = 0; x < w; x++) {
            const int d = (int)a[r * as + x] - (int)b[r * bs + x];
            s += (uint64_t)(d * d);
        }
    return s;
}

/* reconstruction_ssd, src/loop_filter.rs:1965-1979. */
uint64_t orc_lf_ssd(const orc_lf_planes *recon, const orc_lf_source *src)
{
    const size_t cw = (src->width + 1) / 2, ch = (src->height + 1) / 2;
    return plane_ssd(recon->y, recon->y_stride, src->y, src->y_stride, src->width, src->height) +
           plane_ssd(recon->u, recon->uv_stride, src->u, src->uv_stride, cw, ch) +
           plane_ssd(recon->v, recon->uv_stride, src->v, src->uv_stride, cw, ch);
}

static orc_lf_planes clone_planes(const orc_lf_planes *pl, uint8_t **buf)
{
    const size_t ysz = pl->y_stride * pl->mb_rows * 16, csz = pl->uv_stride * pl->mb_rows * 8;
    uint8_t *m = (uint8_t *)malloc(ysz + 2 * csz);
    memcpy(m, pl->y, ysz);
    memcpy(m + ysz, pl->u, csz);
    memcpy(m + ysz + csz, pl->v, csz);
    orc_lf_planes c = *pl;
    c.y = m;
    c.u = m + ysz;
    c.v = m + ysz + csz;
    *buf = m;
    return c;
}

/* score_filter_candidate, src/loop_filter.rs:2062-2095: level 0 is the
 * unfiltered SSD (no filtering at all, whatever the segment overrides). */
uint64_t orc_lf_score(const orc_lf_planes *recon, const orc_lf_mb *mbs, const orc_lf_config *base,
                      const orc_lf_source *src, uint8_t level, uint8_t sharpness)
{
    if (level == 0) return orc_lf_ssd(recon, src);
    uint8_t *buf;
    orc_lf_planes cand = clone_planes(recon, &buf);
    orc_lf_config cfg = *base;
    cfg.loop_filter_level = level;
    cfg.sharpness_level = sharpness;
    orc_lf_filter_frame(&cand, mbs, &cfg);
    const uint64_t s = orc_lf_ssd(&cand, src);
    free(buf);
    return s;
}

/* ---- candidate scoring on a thread pool (score_candidates,
 *      src/loop_filter.rs:2097-2131): results in candidate order. ---- */
typedef struct {
    const orc_lf_planes *recon;
    const orc_lf_mb *mbs;
    const orc_lf_config *base;
    const orc_lf_source *src;
    const uint8_t *levels, *sharps;
    size_t n;
    uint64_t *out;
    atomic_size_t next;
    /* band sampling (0 bands = the whole frame) */
    const size_t *band_rows;
    size_t n_bands;
} score_job;

static uint64_t score_sampled(const score_job *j, uint8_t level, uint8_t sharp);

static void *score_worker(void *arg)
{
    score_job *j = (score_job *)arg;
    for (;;) {
        const size_t i = atomic_fetch_add(&j->next, 1);
        if (i >= j->n) return NULL;
        j->out[i] = j->n_bands ? score_sampled(j, j->levels[i], j->sharps[i])
                               : orc_lf_score(j->recon, j->mbs, j->base, j->src, j->levels[i], j->sharps[i]);
    }
}

static void run_scores(score_job *j, size_t threads)
{
    atomic_init(&j->next, 0);
    size_t t = threads < 1 ? 1 : threads;
    if (t > j->n) t = j->n;
    if (t > 64) t = 64;
    pthread_t th[64];
    size_t started = 0;
    for (size_t k = 1; k < t; k++)
        if (pthread_create(&th[started], NULL, score_worker, j) == 0) started++;
    score_worker(j);
    for (size_t k = 0; k < started; k++) pthread_join(th[k], NULL);
}

void orc_lf_score_candidates(const orc_lf_planes *recon, const orc_lf_mb *mbs, const orc_lf_config *base,
                             const orc_lf_source *src, const uint8_t *levels, const uint8_t *sharps, size_t n,
                             size_t threads, uint64_t *out)
{
    score_job j = {recon, mbs, base, src, levels, sharps, n, out, 0, NULL, 0};
    run_scores(&j, threads);
}

/* select_filter_params_threaded, src/loop_filter.rs:2133-2200: coarse grid
 * 0, 4, .., 60 plus 63 at the base sharpness, a +-3 refine around the grid
 * winner, then the other seven sharpness values at the chosen level; the
 * first strict minimum wins at every stage. */
void orc_lf_select_params(const orc_lf_planes *recon, const orc_lf_mb *mbs, const orc_lf_config *base,
                          const orc_lf_source *src, size_t threads, uint8_t *out_level, uint8_t *out_sharp)
{
    const uint8_t bs = (uint8_t)(base->sharpness_level > 7 ? 7 : base->sharpness_level);
    uint8_t lv[32], sh[32];
    uint64_t sc[32];
    size_t n = 0;
    for (int l = 0; l <= 63; l += 4) lv[n] = (uint8_t)l, sh[n++] = bs;
    lv[n] = 63, sh[n++] = bs;
    orc_lf_score_candidates(recon, mbs, base, src, lv, sh, n, threads, sc);
    uint8_t best = 0;
    uint64_t best_ssd = sc[0];
    for (size_t i = 1; i < n; i++)
        if (sc[i] < best_ssd) best_ssd = sc[i], best = lv[i];
    const int lo = best >= 3 ? best - 3 : 0, hi = best + 3 > 63 ? 63 : best + 3;
    n = 0;
    for (int l = lo; l <= hi; l++)
        if (l != best) lv[n] = (uint8_t)l, sh[n++] = bs;
    orc_lf_score_candidates(recon, mbs, base, src, lv, sh, n, threads, sc);
    for (size_t i = 0; i < n; i++)
        if (sc[i] < best_ssd) best_ssd = sc[i], best = lv[i];
    if (best == 0) {
        *out_level = 0;
        *out_sharp = bs;
        return;
    }
    n = 0;
    for (int s = 0; s <= 7; s++)
        if (s != bs) lv[n] = best, sh[n++] = (uint8_t)s;
    orc_lf_score_candidates(recon, mbs, base, src, lv, sh, n, threads, sc);
    uint8_t best_sharp = bs;
    for (size_t i = 0; i < n; i++)
        if (sc[i] < best_ssd) best_ssd = sc[i], best_sharp = sh[i];
    *out_level = best;
    *out_sharp = best_sharp;
}

/* The score of select_filter_params_sampled, src/loop_filter.rs:2288-2365:
 * bands of up to two MB rows starting every every_rows rows, each filtered
 * as a frame of its own (its first row has no top edge) and scored over its
 * visible rows. */
static uint64_t score_sampled(const score_job *j, uint8_t level, uint8_t sharp)
{
    const orc_lf_planes *pl = j->recon;
    const orc_lf_source *src = j->src;
    uint64_t total = 0;
    for (size_t b = 0; b < j->n_bands; b++) {
        const size_t first = j->band_rows[b];
        const size_t rows = pl->mb_rows - first < 2 ? pl->mb_rows - first : 2;
        const size_t y0 = first * 16;
        if (y0 >= src->height) continue;
        orc_lf_planes band = *pl;
        band.y = pl->y + y0 * pl->y_stride;
        band.u = pl->u + (y0 / 2) * pl->uv_stride;
        band.v = pl->v + (y0 / 2) * pl->uv_stride;
        band.mb_rows = rows;
        orc_lf_source bsrc = *src;
        bsrc.height = rows * 16 < src->height - y0 ? rows * 16 : src->height - y0;
        bsrc.y = src->y + y0 * src->y_stride;
        bsrc.u = src->u + (y0 / 2) * src->uv_stride;
        bsrc.v = src->v + (y0 / 2) * src->uv_stride;
        total += orc_lf_score(&band, j->mbs + first * pl->mb_cols, j->base, &bsrc, level, sharp);
    }
    return total;
}

/* select_filter_params_sampled, src/loop_filter.rs:2288-2412: every eighth
 * level plus 63, then +-4, +-2, +-1 around the running best, then (when
 * asked and the level is not 0) the other seven sharpness values. */
void orc_lf_select_params_sampled(const orc_lf_planes *recon, const orc_lf_mb *mbs, const orc_lf_config *base,
                                  const orc_lf_source *src, size_t every_rows, int sweep_sharpness, size_t threads,
                                  uint8_t *out_level, uint8_t *out_sharp)
{
    const size_t every = every_rows < 2 ? 2 : every_rows;
    size_t band_rows[4096];
    size_t nb = 0;
    for (size_t r = 0; r < recon->mb_rows && nb < 4096; r += every) band_rows[nb++] = r;
    const uint8_t bs = (uint8_t)(base->sharpness_level > 7 ? 7 : base->sharpness_level);
    uint8_t lv[16], sh[16];
    uint64_t sc[16];
    score_job j = {recon, mbs, base, src, lv, sh, 0, sc, 0, band_rows, nb};
    size_t n = 0;
    for (int l = 0; l <= 56; l += 8) lv[n] = (uint8_t)l, sh[n++] = bs;
    lv[n] = 63, sh[n++] = bs;
    j.n = n;
    run_scores(&j, threads);
    uint8_t best = lv[0];
    uint64_t best_ssd = sc[0];
    for (size_t i = 1; i < n; i++)
        if (sc[i] < best_ssd) best_ssd = sc[i], best = lv[i];
    static const uint8_t steps[3] = {4, 2, 1};
    for (int k = 0; k < 3; k++) {
        const uint8_t st = steps[k];
        n = 0;
        if (best >= st) lv[n] = best - st, sh[n++] = bs;
        if (best + st <= 63) lv[n] = best + st, sh[n++] = bs;
        j.n = n;
        run_scores(&j, threads);
        for (size_t i = 0; i < n; i++)
            if (sc[i] < best_ssd) best_ssd = sc[i], best = lv[i];
    }
    uint8_t best_sharp = bs;
    if (sweep_sharpness && best != 0) {
        n = 0;
        for (int s = 0; s <= 7; s++)
            if (s != bs) lv[n] = best, sh[n++] = (uint8_t)s;
        j.n = n;
        run_scores(&j, threads);
        for (size_t i = 0; i < n; i++)
            if (sc[i] < best_ssd) best_ssd = sc[i], best_sharp = sh[i];
    }
    *out_level = best;
    *out_sharp = best_sharp;
}
