/*
 * CPU oracle for the oxideav-vp8 ARNR + motion-compensation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference crate's integer algorithms, used as the checker for the HIP
 * path and as the timed CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never
 * links, calls or falls back to it.
 *
 * Every function cites the reference file:line it follows (paths relative
 * to the oxideav-vp8 crate root).  Pinning of each function (which
 * known-answer test, invariant or brute-force check holds it) is listed in
 * tests/test_oracle_*.py; no function is "parity unpinned".
 */
#ifndef VP8_ORACLE_H
#define VP8_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared with include/vp8g.h (same numeric values). */
enum {
    ORC_OK = 0,
    ORC_E_INVALID_DIMENSIONS = 1,
    ORC_E_REF_DIMS_MISMATCH = 2,
    ORC_E_CENTER_OUT_OF_RANGE = 3,
    ORC_E_ARG = 4
};

/* I420Frame, src/encoder.rs:5419-5434. */
typedef struct {
    uint32_t width, height;
    const uint8_t *y, *u, *v;
    size_t y_stride, uv_stride;
} orc_i420;

/* EncodeError payloads, src/encoder.rs:69-104. */
typedef struct {
    uint32_t src_w, src_h, ref_w, ref_h;
} orc_err;

/* Mv, src/motion_vector.rs:64-70 (quarter-pel, row then col). */
typedef struct {
    int16_t row, col;
} orc_mv;

/* ---------------- ARNR (src/arnr.rs) ---------------- */
uint32_t orc_pixel_weight(int32_t d, uint32_t s_scale);
void orc_weight_table(uint32_t s_scale, uint32_t table[256]);
uint32_t orc_block_sad(const orc_i420 *src, const orc_i420 *ref,
                       int bx, int by, int bw, int bh, int mvx, int mvy,
                       uint32_t early_exit_at);
uint32_t orc_block_sad_generic(const orc_i420 *src, const orc_i420 *ref,
                               int bx, int by, int bw, int bh, int mvx, int mvy,
                               uint32_t early_exit_at);
/* Exhaustive search around (sx, sy) within +-radius; out3 = {mvx, mvy,
 * sad}; minimum SAD, ties to the smallest (|mvx|+|mvy|, mvy, mvx). */
void orc_block_search(const orc_i420 *src, const orc_i420 *ref, int bx, int by, int bw, int bh,
                      int sx, int sy, int radius, int32_t out3[3]);
int orc_block_search_group(const orc_i420 *frames, size_t n, size_t center, int radius,
                           const int32_t *seeds, size_t threads, int32_t *out);
/* Exhaustive fallback of a seeded search over `first` (table layout);
 * gate = 1 queues only failed entries with a 4-neighbour passing the
 * occlusion cut-off. */
int orc_search_fallback(const orc_i420 *frames, size_t n, size_t center, const int32_t *first, int radius,
                        int gate, int32_t *out, uint32_t *queued);
int orc_build_arnr_altref_mv(const orc_i420 *frames, size_t n, size_t center, uint8_t strength,
                             const int32_t *table, uint8_t *out_y, uint8_t *out_u,
                             uint8_t *out_v, orc_err *err);
/* out3 = {mvx, mvy, sad} */
void orc_refine_search(const orc_i420 *src, const orc_i420 *ref,
                       int bx, int by, int bw, int bh, int32_t out3[3]);
void orc_accumulate_luma_block(const uint8_t *ctr_y, const uint8_t *f_y, size_t f_stride,
                               int w, int h, int bx, int by, int bw, int bh,
                               int mvx, int mvy, const uint32_t wt[256],
                               uint32_t *acc, uint32_t *wsum);
void orc_accumulate_chroma_block(const uint8_t *ctr, const uint8_t *f, size_t f_stride,
                                 int cw, int ch, int cbx, int cby, int cbw, int cbh,
                                 int cmvx, int cmvy, const uint32_t wt[256],
                                 uint32_t *acc, uint32_t *wsum);
/*
 * build_arnr_altref_threaded.  Output planes are tightly packed
 * (w*h, cw*ch, cw*ch).  `aligned_out` (optional, may be NULL) receives the
 * search table, n * brows * bcols entries of {mvx, mvy, sad} (centre rows
 * left zero).  Returns an ORC_* status.
 */
int orc_build_arnr_altref(const orc_i420 *frames, size_t n, size_t center,
                          uint8_t strength, size_t threads,
                          uint8_t *out_y, uint8_t *out_u, uint8_t *out_v,
                          orc_err *err, int32_t *aligned_out);

/* ---------------- motion compensation (src/motion_comp.rs) ---------------- */
extern const int32_t ORC_SIXTAP_FILTERS[8][6];
extern const int32_t ORC_BILINEAR_FILTERS[8][6];
orc_mv orc_stored_luma_mv(orc_mv mv);
int32_t orc_avg(int32_t c1, int32_t c2, int32_t c3, int32_t c4);
orc_mv orc_chroma_mv(orc_mv luma_mv);
orc_mv orc_apply_full_pixel(orc_mv mv);
uint8_t orc_interp(const int32_t filter[6], const uint8_t support[6]);
void orc_fetch_block_halo(const uint8_t *plane, size_t stride, size_t w, size_t h,
                          size_t blk_x, size_t blk_y, orc_mv mv, uint8_t out[81]);
void orc_sixtap_2d(const uint8_t halo[81], int mx, int my, const int32_t filters[8][6],
                   uint8_t out[16]);
void orc_fetch_luma_mb_halo(const uint8_t *plane, size_t stride, size_t w, size_t h,
                            size_t mb_x, size_t mb_y, orc_mv mv, uint8_t out[441]);
void orc_sixtap_mb_luma(const uint8_t halo[441], int mx, int my,
                        const int32_t filters[8][6], uint8_t out[256]);
void orc_fetch_chroma_mb_halo(const uint8_t *plane, size_t stride, size_t w, size_t h,
                              size_t mb_x, size_t mb_y, orc_mv mv, uint8_t out[169]);
void orc_sixtap_mb_chroma(const uint8_t halo[169], int mx, int my,
                          const int32_t filters[8][6], uint8_t out[64]);
void orc_fetch_luma_mb_whole_pixel(const uint8_t *plane, size_t stride, size_t w, size_t h,
                                   size_t mb_x, size_t mb_y, orc_mv mv, uint8_t out[256]);
void orc_fetch_chroma_mb_whole_pixel(const uint8_t *plane, size_t stride, size_t w,
                                     size_t h, size_t mb_x, size_t mb_y, orc_mv mv,
                                     uint8_t out[64]);
/* ReferencePlanes, src/motion_comp.rs:131-146. */
typedef struct {
    const uint8_t *y, *u, *v;
    size_t y_stride, uv_stride, mb_cols, mb_rows;
} orc_ref_planes;
/* predict_inter_mb; out = y[256] | u[64] | v[64] (ReconstructedMb). */
void orc_predict_inter_mb(const orc_ref_planes *ref, size_t mb_col, size_t mb_row,
                          orc_mv luma_mv, int full_pixel, const int32_t filters[8][6],
                          uint8_t out[384]);

/* ---------------- SAD scoring (src/motion_search.rs) ---------------- */
/* LumaRef, src/motion_search.rs:230-239. */
typedef struct {
    const uint8_t *plane;
    size_t stride, width, height;
} orc_luma_ref;
typedef struct {
    orc_mv mv;
    uint32_t sad;
} orc_search_result;
uint32_t orc_block_sad_16x16(const uint8_t src[256], const uint8_t pred[256]);
uint32_t orc_mb_luma_sad_at_whole_mv(const orc_luma_ref *ref, size_t mb_col, size_t mb_row,
                                     const uint8_t src[256], orc_mv mv);
uint32_t orc_mb_luma_sad_at_mv(const orc_luma_ref *ref, size_t mb_col, size_t mb_row,
                               const uint8_t src[256], orc_mv mv);
orc_search_result orc_half_pixel_refine_luma(const orc_luma_ref *ref, size_t mb_col,
                                             size_t mb_row, const uint8_t src[256],
                                             orc_mv center);
orc_search_result orc_quarter_pixel_refine_luma(const orc_luma_ref *ref, size_t mb_col,
                                                size_t mb_row, const uint8_t src[256],
                                                orc_mv center);

/* HalfPlanes, src/encoder/inter_search.rs:256-346. */
void orc_fetch_window(const orc_luma_ref *ref, int x, int y, uint8_t out[22 * 22]);
typedef struct {
    uint8_t h[22 * 17];
    uint8_t v[17 * 16];
    uint8_t d[17 * 17];
} orc_half_planes;
void orc_half_planes_new(const uint8_t window[22 * 22], const int32_t taps[6], int diagonals,
                         orc_half_planes *out);
uint32_t orc_half_planes_sad(const orc_half_planes *p, const uint8_t src[256], int dx, int dy);

/* luma_mad, src/stream.rs:2654-2666. */
double orc_luma_mad(const orc_i420 *prev, const orc_i420 *frame);

/* ---------------- SPLITMV and reconstruction (src/motion_comp.rs, src/inverse_transform.rs) ---------------- */
void orc_filter_block_4x4(const uint8_t *plane, size_t stride, size_t w, size_t h, size_t blk_x, size_t blk_y,
                          orc_mv mv, const int32_t filters[8][6], uint8_t out[16]);
void orc_split_chroma_mvs(const orc_mv luma[16], orc_mv out[4]);
void orc_predict_split_mv(const orc_ref_planes *ref, size_t mb_col, size_t mb_row, const orc_mv mvs[16],
                          int full_pixel, const int32_t filters[8][6], uint8_t out[384]);
void orc_inverse_wht_4x4(const int16_t in[16], int16_t out[16]);
void orc_inverse_dct_4x4(const int16_t in[16], int16_t out[16]);
void orc_inverse_dct_4x4_add_into(const int16_t in[16], uint8_t *plane, size_t stride, int sub_row, int sub_col);
void orc_reconstruct_inter_mb(const orc_ref_planes *ref, size_t mb_col, size_t mb_row, orc_mv mv, int full_pixel,
                              const int32_t filters[8][6], int skip, const int16_t coeffs[25][16], uint8_t out[384]);
void orc_reconstruct_split_mv_mb(const orc_ref_planes *ref, size_t mb_col, size_t mb_row, const orc_mv mvs[16],
                                 int full_pixel, const int32_t filters[8][6], int skip, const int16_t coeffs[24][16],
                                 uint8_t out[384]);
void orc_predict_split_mv_batch(const orc_ref_planes *ref, const uint32_t *mb_xy, const orc_mv *mvs, size_t n,
                                int full_pixel, const int32_t filters[8][6], uint8_t *out);
void orc_reconstruct_inter_mb_batch(const orc_ref_planes *ref, const uint32_t *mb_xy, const orc_mv *mv, size_t n,
                                    int full_pixel, const int32_t filters[8][6], const uint8_t *skip,
                                    const int16_t *coeffs, uint8_t *out);
void orc_reconstruct_split_mv_mb_batch(const orc_ref_planes *ref, const uint32_t *mb_xy, const orc_mv *mvs, size_t n,
                                       int full_pixel, const int32_t filters[8][6], const uint8_t *skip,
                                       const int16_t *coeffs, uint8_t *out);

/* ---------------- encoder inter search (src/encoder/inter_search.rs) ---------------- */
orc_search_result orc_small_diamond_search_luma(const orc_luma_ref *ref, size_t mb_col, size_t mb_row,
                                                const uint8_t src[256], orc_mv center, uint32_t max_iters);
/* MvCostTable, src/encoder/inter_search.rs:166-194: bits[component][diff + 1023], 0 = row. */
typedef struct {
    float bits[2][2047];
} orc_mv_cost_table;
/* The per-macroblock inputs of whole_pixel_search / sub_pixel_refine:
 * position, the census predictor (RefSearch::census.best), the whole-pixel
 * range, seeds, lambda_sad, good_enough and the descend flag. */
typedef struct {
    uint32_t mb_col, mb_row;
    orc_mv best;
    int32_t range_x[2], range_y[2];
    uint32_t n_seeds;
    orc_mv seeds[8];
    double lambda_sad;
    uint32_t good_enough;
    int32_t descend;
    const uint8_t *src;  /* the macroblock's 16x16 source */
} orc_mb_search;
/* SearchShape fields the two searches read (src/encoder/inter_search.rs:66-93). */
typedef struct {
    uint32_t descent_rounds;
    int32_t wide_pattern, half_diagonals, quarter;
} orc_search_shape;
/* Found, src/encoder/inter_search.rs:349-355. */
typedef struct {
    orc_mv mv;
    uint32_t sad;
    double cost;
} orc_found;
orc_found orc_whole_pixel_search(const orc_luma_ref *ref, const orc_mv_cost_table *costs, const orc_mb_search *job,
                                 const orc_search_shape *shape);
orc_found orc_sub_pixel_refine(const orc_luma_ref *ref, const orc_mv_cost_table *costs, const orc_mb_search *job,
                               orc_found start, const int32_t filters[8][6], const orc_search_shape *shape);
void orc_whole_pixel_search_batch(const orc_luma_ref *ref, const orc_mv_cost_table *costs, const orc_mb_search *jobs,
                                  size_t n, const orc_search_shape *shape, orc_found *out);
void orc_sub_pixel_refine_batch(const orc_luma_ref *ref, const orc_mv_cost_table *costs, const orc_mb_search *jobs,
                                const orc_found *start, size_t n, const int32_t filters[8][6],
                                const orc_search_shape *shape, orc_found *out);

/* ---------------- batch helpers (one call over many MBs) ---------------- */
void orc_predict_inter_mb_batch(const orc_ref_planes *ref, const uint32_t *mb_xy,
                                const orc_mv *mv, size_t n, int full_pixel,
                                const int32_t filters[8][6], uint8_t *out);
void orc_mb_luma_sad_at_mv_batch(const orc_luma_ref *ref, const uint32_t *mb_xy,
                                 const uint8_t *src, const orc_mv *mv, size_t n,
                                 uint32_t *sad);

/* ---------------- loop filter (src/loop_filter.rs) ---------------- */
/* LoopFilterParams, src/loop_filter.rs:71-84. */
typedef struct {
    uint8_t interior_limit, hev_threshold, mbedge_limit, sub_bedge_limit;
} orc_lf_params;
/* FrameFilterConfig, src/loop_filter.rs:891-944; ref_delta is
 * [CURRENT, LAST, GOLDEN, ALTREF] and mode_delta [B_PRED, ZEROMV,
 * NEAREST/NEAR/NEW, SPLITMV] (src/loop_filter.rs:1193-1209).  Same layout
 * as vp8g_lf_config. */
typedef struct {
    int32_t simple, key_frame, loop_filter_level, sharpness_level;
    int32_t segmentation_enabled, segment_abs, delta_enabled;
    int16_t segment_lf_level[4];
    int16_t ref_delta[4];
    int16_t mode_delta[4];
} orc_lf_config;
/* One macroblock's filter inputs: MacroblockModes::segment_id / y_mode,
 * the ref frame, the inter mode and mb_has_coeffs (src/loop_filter.rs:1364-1372).
 * Same layout as vp8g_lf_mb. */
typedef struct {
    uint8_t segment_id;  /* 0..3, ORC_LF_NO_SEGMENT for None */
    uint8_t ref_frame;   /* 0 intra (CURRENT), 1 LAST, 2 GOLDEN, 3 ALTREF */
    uint8_t inter_mode;  /* ORC_LF_ZEROMV, ORC_LF_SPLITMV, anything else NEAREST/NEAR/NEW */
    uint8_t flags;       /* ORC_LF_BPRED | ORC_LF_HAS_COEFFS */
} orc_lf_mb;
enum { ORC_LF_NO_SEGMENT = 255, ORC_LF_ZEROMV = 1, ORC_LF_SPLITMV = 2, ORC_LF_BPRED = 1, ORC_LF_HAS_COEFFS = 2 };
/* KeyframePlanes (MB-padded reconstruction), src/frame.rs. */
typedef struct {
    uint8_t *y, *u, *v;
    size_t y_stride, uv_stride, mb_cols, mb_rows;
} orc_lf_planes;
/* SourcePlanes, src/loop_filter.rs:1916-1932. */
typedef struct {
    size_t width, height;
    const uint8_t *y, *u, *v;
    size_t y_stride, uv_stride;
} orc_lf_source;
orc_lf_params orc_lf_derive(uint8_t level, uint8_t sharpness, int key_frame);
int orc_lf_common_adjust(int use_outer_taps, uint8_t *seg);
void orc_lf_simple_segment(uint8_t edge_limit, uint8_t *seg);
void orc_lf_subblock_filter(uint8_t hev_threshold, uint8_t interior_limit, uint8_t edge_limit, uint8_t *seg);
void orc_lf_mb_filter(uint8_t hev_threshold, uint8_t interior_limit, uint8_t edge_limit, uint8_t *seg);
uint8_t orc_lf_mb_level(const orc_lf_config *cfg, orc_lf_mb mb);
void orc_lf_filter_frame(orc_lf_planes *pl, const orc_lf_mb *mbs, const orc_lf_config *cfg);
uint64_t orc_lf_ssd(const orc_lf_planes *recon, const orc_lf_source *src);
uint64_t orc_lf_score(const orc_lf_planes *recon, const orc_lf_mb *mbs, const orc_lf_config *base,
                      const orc_lf_source *src, uint8_t level, uint8_t sharpness);
void orc_lf_score_candidates(const orc_lf_planes *recon, const orc_lf_mb *mbs, const orc_lf_config *base,
                             const orc_lf_source *src, const uint8_t *levels, const uint8_t *sharps, size_t n,
                             size_t threads, uint64_t *out);
void orc_lf_select_params(const orc_lf_planes *recon, const orc_lf_mb *mbs, const orc_lf_config *base,
                          const orc_lf_source *src, size_t threads, uint8_t *out_level, uint8_t *out_sharp);
void orc_lf_select_params_sampled(const orc_lf_planes *recon, const orc_lf_mb *mbs, const orc_lf_config *base,
                                  const orc_lf_source *src, size_t every_rows, int sweep_sharpness, size_t threads,
                                  uint8_t *out_level, uint8_t *out_sharp);

#ifdef __cplusplus
}
#endif
#endif
