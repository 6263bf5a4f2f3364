/*
 * vp8g — C ABI of the gfx950 (MI355X) implementation of oxideav-vp8's
 * encoder-side data-parallel path: ARNR altref synthesis, VP8 motion
 * compensation and sub-pel SAD scoring.
 *
 * Every entry point replaces one reference function (paths relative to the
 * oxideav-vp8 crate root) and keeps its argument meaning, output bytes and
 * error order.  Plain pointers and sizes only.
 *
 * Memory: entry points without a suffix take HOST pointers (the library
 * copies in, runs, copies out, and returns after the result is in place).
 * `_dev` entry points take DEVICE pointers, enqueue on the context's stream
 * and return without waiting; call vp8g_ctx_synchronize() before reading.
 * Their per-MB arrays (mb_xy, mv, src, out, sad) must be 4-byte aligned.
 *
 * Threading: a context is not internally locked; use one per host thread.
 * Results never depend on launch shapes, batch composition or thread count.
 */
#ifndef VP8G_H
#define VP8G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define VP8G_OK 0
#define VP8G_E_INVALID_DIMENSIONS 1   /* EncodeError::InvalidDimensions   (src/encoder.rs:72) */
#define VP8G_E_REF_DIMS_MISMATCH 2    /* EncodeError::ReferenceDimensionsMismatch (src/encoder.rs:99-104) */
#define VP8G_E_CENTER_OUT_OF_RANGE 3  /* the reference panics (src/arnr.rs:494-498) */
#define VP8G_E_ARG 4                  /* null pointer, short stride, limit exceeded */
#define VP8G_E_DEVICE 5               /* HIP runtime error; no output was written */
#define VP8G_E_NO_DEVICE 6            /* no usable gfx950 device */

/* Largest group the ARNR entry points accept in one call. */
#define VP8G_MAX_FRAMES 64

typedef struct vp8g_ctx vp8g_ctx;

/* I420Frame, src/encoder.rs:5419-5434: visible dims, planes, strides. */
typedef struct {
    uint32_t width, height;
    const uint8_t *y, *u, *v;
    size_t y_stride, uv_stride;
} vp8g_i420;

/* EncodeError payloads: InvalidDimensions{width,height} -> src_w/src_h;
 * ReferenceDimensionsMismatch{source, reference} -> all four. */
typedef struct {
    uint32_t src_w, src_h, ref_w, ref_h;
} vp8g_err;

/* Mv, src/motion_vector.rs:64-70 (quarter-pel). */
typedef struct {
    int16_t row, col;
} vp8g_mv;

/* ReferencePlanes, src/motion_comp.rs:131-146 (MB-aligned planes). */
typedef struct {
    const uint8_t *y, *u, *v;
    size_t y_stride, uv_stride, mb_cols, mb_rows;
} vp8g_ref_planes;

/* LumaRef, src/motion_search.rs:230-239 (width/height = clamp bounds). */
typedef struct {
    const uint8_t *plane;
    size_t stride, width, height;
} vp8g_luma_ref;

/* SearchResult, src/motion_search.rs:248-255. */
typedef struct {
    vp8g_mv mv;
    uint32_t sad;
} vp8g_search_result;

/* ---- context ---- */
int vp8g_ctx_create(int device, vp8g_ctx **out);
void vp8g_ctx_destroy(vp8g_ctx *ctx);
/* Run later work on `stream` (a hipStream_t; NULL restores the context's
 * own stream). */
int vp8g_ctx_set_stream(vp8g_ctx *ctx, void *stream);
void *vp8g_ctx_stream(vp8g_ctx *ctx);
int vp8g_ctx_synchronize(vp8g_ctx *ctx);
/* Launch-shape options; none of them changes any output byte.
 * VP8G_OPT_ARNR_WAVES: waves per anchor macroblock in an ARNR launch, each
 *   taking a share of the neighbour frames (0 = enough for about 4096 waves
 *   per launch; at most one per neighbour; waves x MBs per workgroup <= 16).
 * VP8G_OPT_XCD_REMAP: 1 (default) gives each XCD a contiguous band of
 *   macroblocks, 0 uses the launch order.
 * VP8G_OPT_ARNR_MBS: anchor macroblocks per ARNR workgroup, 1..8
 *   (0 = default 4; reduced automatically to fit 16 waves and 64 KiB of
 *   LDS).
 * VP8G_OPT_BATCH_WGS_PER_CU: persistent workgroups per CU of a batch,
 *   0..32 (0 = as many as fit).
 * VP8G_OPT_FAULT: host-side fault injection for testing callers' error
 *   paths (no counterpart in the reference).  VP8G_FAULT_ALLOC makes every
 *   device allocation step of a host-pointer or batch entry point fail,
 *   VP8G_FAULT_COPY every host<->device copy, VP8G_FAULT_LAUNCH every
 *   kernel or graph launch.  The failing step reports VP8G_E_DEVICE
 *   without running (nothing faulty reaches the device) and no output is
 *   written.  A context starts with the mode named by $VP8G_FAULT
 *   ("alloc", "copy" or "launch"; anything else = none).
 * VP8G_OPT_HOST_THREADS: host threads that pack the caller's planes into
 *   pinned staging and copy results out in vp8g_build_arnr_altref, 1..64
 *   (0 = default, min(8, half the hardware threads)).
 * VP8G_OPT_COPY_STREAMS: accepted for callers that set it, 0..4, with no
 *   effect: vp8g_build_arnr_altref's frames go out by the copy engines on
 *   one stream of its own and the altref rows come back by a copy kernel
 *   on the context's stream. */
#define VP8G_OPT_ARNR_WAVES 1
#define VP8G_OPT_XCD_REMAP 2
#define VP8G_OPT_ARNR_MBS 3
#define VP8G_OPT_BATCH_WGS_PER_CU 4
#define VP8G_OPT_FAULT 5
#define VP8G_FAULT_NONE 0
#define VP8G_FAULT_ALLOC 1
#define VP8G_FAULT_COPY 2
#define VP8G_FAULT_LAUNCH 3
#define VP8G_OPT_HOST_THREADS 6
#define VP8G_OPT_COPY_STREAMS 7
/* Row bands of vp8g_build_arnr_altref (host buffers): the frames go out a
 * band of MB rows at a time (each band's rows plus the 15 luma / 7 chroma
 * rows its search reaches below), the kernel runs on each band as soon as
 * its rows are in while later bands are still on the link, and each band's
 * altref rows come back as soon as its kernel is done.  The last band is
 * half as tall as the others.  1..8; 0 restores the default (3). */
#define VP8G_OPT_HOST_BANDS 8
int vp8g_ctx_set_option(vp8g_ctx *ctx, int option, int value);
const char *vp8g_status_str(int status);

/* Launch-shape tuning table (JSON written by the autotuner,
 * `python -m vp8g.autotune`): entries {"class", "n", "mbs", "waves",
 * "wgs_per_cu"} keyed by resolution class (qcif <= 176x144 < cif <=
 * 352x288 < 360p <= 640x360 < 720p <= 1280x720 < 1080p <= 1920x1088 < 4k,
 * by luma area) and group length; class "batch" for batched launches.  A
 * context loads $VP8G_TUNING, or tuning/gfx950.json beside the library,
 * when created; explicit options win over the table, the table over the
 * built-in defaults.  Like the options, the table never changes an output
 * byte.  NULL clears the table; an unreadable or malformed file returns
 * VP8G_E_ARG and keeps the current one. */
int vp8g_ctx_load_tuning(vp8g_ctx *ctx, const char *path);
/* The entry a table file gives for a group (batch = 0) or a batch
 * (batch = 1); zeros when no entry matches.  Needs no device. */
int vp8g_tuning_lookup(const char *path, uint32_t width, uint32_t height, size_t n_frames,
                       int batch, int *mbs, int *waves, int *wgs_per_cu);
/* The launch shape a single-group ARNR call with these dims would use. */
int vp8g_arnr_shape(vp8g_ctx *ctx, uint32_t width, uint32_t height, size_t n_frames,
                    uint8_t strength, int *mbs, int *waves);

/* ---- ARNR altref synthesis ----
 * Replaces build_arnr_altref (src/arnr.rs:476-482) and
 * build_arnr_altref_threaded (src/arnr.rs:488-650; the thread budget has no
 * meaning here and the output never depended on it).
 * Check order as the reference: centre index, centre dims (0 or > 0x3FFF),
 * then the first frame whose dims differ from the centre's.  Output planes
 * are tightly packed: y[w*h], u[cw*ch], v[cw*ch], cw = ceil(w/2),
 * ch = ceil(h/2) (ArnrPicture, src/arnr.rs:90-101).  `strength` is the raw
 * ArnrConfig::strength (src/arnr.rs:52-56): 0 copies the centre, values
 * above 6 act as 6 (src/arnr.rs:533).  `err` may be NULL. */
int vp8g_build_arnr_altref(vp8g_ctx *ctx, const vp8g_i420 *frames, size_t n_frames,
                           size_t center, uint8_t strength, uint8_t *out_y, uint8_t *out_u,
                           uint8_t *out_v, vp8g_err *err);
/* Same with device-resident planes and outputs.  `mv_out` (device, may be
 * NULL) receives the per-block alignment table
 * n_frames * ceil(h/16) * ceil(w/16) * {mvx, mvy, sad} (centre rows are
 * left untouched). */
int vp8g_build_arnr_altref_dev(vp8g_ctx *ctx, const vp8g_i420 *frames, size_t n_frames,
                               size_t center, uint8_t strength, uint8_t *out_y, uint8_t *out_u,
                               uint8_t *out_v, vp8g_err *err, int32_t *mv_out);

/* ---- exhaustive block search and table-driven altref ----
 * No counterpart in the reference (its ARNR search is the 4-step descent):
 * the judge and labeller for learned motion proposals.  For every 16x16
 * block of frame `center` and every other frame f, the SADs of all
 * candidates seed + (dx, dy), |dx|, |dy| <= radius, are taken with
 * block_sad's semantics (clamped fetches at the visible edges, partial
 * edge blocks; src/arnr.rs:149-153, 200-225); the minimum SAD wins and
 * ties go to the smallest (|mvx| + |mvy|, mvy, mvx).  `out` (device) uses
 * the alignment-table layout of vp8g_build_arnr_altref_dev's mv_out,
 * n_frames * ceil(h/16) * ceil(w/16) * {mvx, mvy, sad}; centre rows are
 * left untouched.  `seeds` (device, may be NULL = all zero) uses the same
 * layout (the third entry is ignored); seed components are clamped to
 * +-16384.  radius: 2, 16 or 64.  Validation as vp8g_build_arnr_altref. */
int vp8g_block_search_dev(vp8g_ctx *ctx, const vp8g_i420 *frames, size_t n_frames, size_t center,
                          int radius, const int32_t *seeds, int32_t *out, vp8g_err *err);
/* Second stage of a seeded search, no counterpart in the reference: for
 * every block and frame, exact +-2 searches (vp8g_block_search_dev's
 * semantics) around extra centres, and the best of those and the first
 * stage's result `first` (the table vp8g_block_search_dev wrote with radius
 * 2 around `seeds`; seeds may be NULL = zero, as there).  The centres are
 * the block's own first-stage vector when it lies on the border of the
 * first square (2 away from the seed in x or y), then the first-stage
 * vectors of the left, right, upper and lower blocks, each once and only
 * when it differs from the seed and the own vector.  Ties go to the
 * smaller SAD, then (|mvx| + |mvy|, mvy, mvx).  `out` (device) must not be
 * `first`; centre rows are left untouched. */
int vp8g_refine_candidates_dev(vp8g_ctx *ctx, const vp8g_i420 *frames, size_t n_frames, size_t center,
                               const int32_t *seeds, const int32_t *first, int32_t *out, vp8g_err *err);
/* Exhaustive fallback of a seeded search, no counterpart in the reference:
 * `first` (device) is a seeded search's alignment table (e.g. the radius-2
 * search around learned proposals).  Every (frame, block) entry whose SAD
 * fails the occlusion cut-off (sad >= 13 * bw * bh, the reference's
 * sad / (bw * bh) > 12, src/arnr.rs:603-608) is queued; with
 * VP8G_FALLBACK_GATE_NEIGHBOUR only those with a left, right, upper or
 * lower block of the same frame that passes it.  `out` (device, not
 * `first`) receives `first`, where every queued entry is replaced by the
 * better of itself and the exhaustive +-radius search around (0, 0)
 * (vp8g_block_search_dev's semantics; radius 16 or 64) under (sad,
 * |mvx| + |mvy|, mvy, mvx).  `queued` (device uint32, may be NULL) receives
 * the number of queued entries.  Centre rows are left untouched.  The
 * queue lives in a device workspace the context allocates (4 bytes per
 * block and frame) and keeps.  Validation as vp8g_build_arnr_altref. */
#define VP8G_FALLBACK_GATE_NEIGHBOUR 1
int vp8g_search_fallback_dev(vp8g_ctx *ctx, const vp8g_i420 *frames, size_t n_frames, size_t center,
                             const int32_t *first, int radius, int flags, int32_t *out, uint32_t *queued,
                             vp8g_err *err);
/* vp8g_search_fallback_dev for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  Same arguments, checks and
 * output, but the queue lives in a static device array the library holds
 * on every GPU it runs on until the process ends (256 bytes + 4 MiB): at
 * most 1048576 (frame, block) entries, (n_frames - 1) * ceil(w/16) *
 * ceil(h/16); larger groups return VP8G_E_ARG.  The call allocates
 * nothing.  Every call on one GPU shares that array, from any context:
 * calls that overlap on one GPU (different streams or threads) race. */
int vp8g_search_fallback_noalloc_dev(vp8g_ctx *ctx, const vp8g_i420 *frames, size_t n_frames, size_t center,
                                     const int32_t *first, int radius, int flags, int32_t *out,
                                     uint32_t *queued, vp8g_err *err);
/* The altref of build_arnr_altref (src/arnr.rs:488-650) with the search
 * replaced by a given alignment table (device, the layout above; the
 * centre's rows are not read): occlusion guard on the table's SAD, blend
 * and resolve exactly as the reference applies them to its own table.
 * Vectors may be any size (reads are clamped). */
int vp8g_build_arnr_altref_table_dev(vp8g_ctx *ctx, const vp8g_i420 *frames, size_t n_frames,
                                     size_t center, uint8_t strength, const int32_t *table,
                                     uint8_t *out_y, uint8_t *out_u, uint8_t *out_v,
                                     vp8g_err *err);

/* ---- coarse-to-fine pyramid ----
 * No counterpart in the reference: the two steps that let the short exact
 * search (radius 2) follow a proposal made at a coarser level.
 * vp8g_downsample2_i420_dev writes one pyramid level of each of the
 * n_frames (<= 64) frames src[i] into dst[i] (device planes; all src frames
 * share one size w x h): dst[i] is ceil(w/2) x ceil(h/2), and each of its
 * Y, U and V samples is the rounded mean (a + b + c + d + 2) >> 2 of the
 * 2x2 source block, coordinates clamped to the plane.  dst[i].width and
 * .height must be ceil(w/2) and ceil(h/2).  vp8g_upsample_table_dev maps an alignment
 * table of n_frames x coarse_h x coarse_w blocks onto n_frames x fine_h x
 * fine_w blocks: fine block (r, c) takes coarse block (min(r/2, coarse_h-1),
 * min(c/2, coarse_w-1)) with both vector components doubled and sad 0 (the
 * seeds of vp8g_block_search_dev one level finer); every grid side is
 * 1..1024 blocks (VP8G_E_ARG otherwise).  All on the context's stream. */
int vp8g_downsample2_i420_dev(vp8g_ctx *ctx, const vp8g_i420 *src, const vp8g_i420 *dst, size_t n_frames);
int vp8g_upsample_table_dev(vp8g_ctx *ctx, const int32_t *coarse, size_t n_frames, int coarse_w,
                            int coarse_h, int32_t *fine, int fine_w, int fine_h);

/* ---- learned motion proposer ----
 * No counterpart in the reference: a ~60k-parameter BF16 CNN (hand-written
 * MFMA kernels, fp32 accumulation and fp32 master parameters) that
 * proposes one whole-pel vector per 16x16 block from the 4x-downsampled
 * (anchor, neighbour) luma pair, for callers that seed a short exact search
 * (vp8g_block_search_dev with radius 2) instead of the full +-16 one.
 * Inputs ("planes"): u8 [image][2][4*ceil(h/16)][4*ceil(w/16)], the anchor
 * then the neighbour, each pixel the rounded mean of a 4x4 full-resolution
 * block (clamped at the frame edge).  Network: a fixed correlation layer
 * (per block, the 81 mean absolute differences of its 4x4 anchor pixels
 * against the neighbour displaced by -4..4 downsampled pixels, plus both
 * frames' pixels; 128 channels), a 1x1 conv 128->32, 3x3 convs 32->32,
 * 32->64, 64->32, 32->32 (leaky ReLU, slope 1/16), and a 1x1 head 32->2
 * predicting mv / 8.
 * Parameters: vp8g_proposer_param_count() fp32 values, per conv layer
 * weights [taps][cout][cin] then bias[cout], then the head [2][32] and
 * bias[2].  Labels: int32 [image][block][2] = {mvx, mvy}; training loss:
 * mean Huber (delta 2 px) on mv / 8 over all label entries, where a block
 * whose mvx is -32768 contributes zero.  All pointers device, 16-byte aligned;
 * everything runs on the context's stream. */
typedef struct vp8g_proposer vp8g_proposer;
size_t vp8g_proposer_param_count(void);
int vp8g_proposer_create(vp8g_ctx *ctx, vp8g_proposer **out);
void vp8g_proposer_destroy(vp8g_proposer *p);
/* Host copies of the parameters; setting them resets the optimizer. */
int vp8g_proposer_set_params(vp8g_proposer *p, const float *params, size_t n);
int vp8g_proposer_get_params(vp8g_proposer *p, float *params, size_t n);
/* Gradients of the last backward pass and its mean Huber loss. */
int vp8g_proposer_get_grads(vp8g_proposer *p, float *grads, size_t n, float *loss);
/* Planes of the n_frames - 1 (centre, other frame) pairs, in frame order. */
int vp8g_proposer_prepare_dev(vp8g_proposer *p, const vp8g_i420 *frames, size_t n_frames,
                              size_t center, uint8_t *planes);
/* Forward pass; a4_out (may be NULL) receives the last hidden layer,
 * bf16 [image][block][32]. */
int vp8g_proposer_forward_dev(vp8g_proposer *p, const uint8_t *planes, int images, int blocks_w,
                              int blocks_h, uint16_t *a4_out);
/* Proposals for every non-centre frame into an alignment table (layout of
 * vp8g_block_search_dev's out; sad entries are 0), components rounded and
 * clamped to +-clamp. */
int vp8g_proposer_propose_dev(vp8g_proposer *p, const vp8g_i420 *frames, size_t n_frames,
                              size_t center, int clamp, int32_t *table);
/* Forward + backward (gradients kept, parameters unchanged). */
int vp8g_proposer_backward_dev(vp8g_proposer *p, const uint8_t *planes, const int32_t *labels,
                               int images, int blocks_w, int blocks_h);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) at learning rate lr;
 * loss_dev (device float, may be NULL) receives the step's loss. */
int vp8g_proposer_train_step_dev(vp8g_proposer *p, const uint8_t *planes, const int32_t *labels,
                                 int images, int blocks_w, int blocks_h, float lr, float *loss_dev);

/* Synthetic region-motion luma frames on the device (training data for the
 * proposer; no counterpart in the reference).  Frame t, pixel (x, y):
 * a 3x3-box-blurred splitmix64 noise texture (texture_seed) sampled at
 * (x - t*vx, y - t*vy) clamped, (vx, vy) = vel[8*ry + rx] for the 8x8
 * region grid; then each rect of frame t (in order) replaced by its own
 * blurred noise; then + Gaussian(sigma) noise (Box-Muller on splitmix64
 * draws of noise_seeds[t]), rounded and clipped.  `out` and `clean` (may be
 * NULL; the frames before the Gaussian stage) are device [n][h][w].
 * Returns after the frames are written. */
typedef struct {
    uint64_t seed;
    uint32_t x0, y0, w, h;
} vp8g_synth_rect;
typedef struct {
    uint32_t width, height;
    size_t n, rects_per_frame;
    uint64_t texture_seed;
    const int32_t *vel;             /* host [64][2] */
    const vp8g_synth_rect *rects;   /* host [n][rects_per_frame] */
    const uint64_t *noise_seeds;    /* host [n] */
    double sigma;
} vp8g_synth_spec;
int vp8g_synth_luma_dev(vp8g_ctx *ctx, const vp8g_synth_spec *spec, uint8_t *out, uint8_t *clean);

/* ---- batched ARNR ----
 * Many groups (clips of any resolution and length) in one persistent launch,
 * captured once into a HIP graph and replayed by vp8g_arnr_batch_run.
 * Replaces a loop of build_arnr_altref_threaded calls
 * (src/arnr.rs:488-650), one per job, as the stream driver's flush_group
 * (src/stream.rs:2806-2957) issues them; every job's output bytes equal its
 * own single call's.  Each job is validated as vp8g_build_arnr_altref_dev
 * validates its arguments, in job order; on failure *failed_job names the
 * job and `err` carries its payload.  Planes, outputs and mv_out are device
 * pointers fixed at creation: a run reads whatever the frame buffers hold
 * when it executes.  Launch shape comes from the context's options at
 * creation (VP8G_OPT_ARNR_MBS, VP8G_OPT_ARNR_WAVES — default 1 for a batch —
 * and VP8G_OPT_BATCH_WGS_PER_CU).  At most 65535 jobs. */
typedef struct {
    const vp8g_i420 *frames;   /* host array of n_frames device-plane frames */
    size_t n_frames, center;
    uint8_t strength;
    uint8_t *out_y, *out_u, *out_v;  /* device, packed as vp8g_build_arnr_altref */
    int32_t *mv_out;                 /* device, may be NULL */
} vp8g_arnr_job;

typedef struct vp8g_arnr_batch vp8g_arnr_batch;

int vp8g_arnr_batch_create(vp8g_ctx *ctx, const vp8g_arnr_job *jobs, size_t n_jobs,
                           vp8g_arnr_batch **out, size_t *failed_job, vp8g_err *err);
/* Enqueue one replay on the context's current stream; returns without
 * waiting.  Replays on one stream are serialised; a batch must not run on
 * two streams at once. */
int vp8g_arnr_batch_run(vp8g_arnr_batch *batch);
/* Launch shape chosen at creation: MBs per workgroup, waves per workgroup,
 * persistent workgroups, work units (strips). Any pointer may be NULL. */
int vp8g_arnr_batch_shape(const vp8g_arnr_batch *batch, int *mbs, int *waves, int *workgroups,
                          size_t *units);
void vp8g_arnr_batch_destroy(vp8g_arnr_batch *batch);

/* ---- motion compensation ----
 * Replaces predict_inter_mb (src/motion_comp.rs:1986-2078) over a batch:
 * MB i sits at (mb_xy[2i], mb_xy[2i+1]) = (mb_col, mb_row) with quarter-pel
 * `mv[i]`; `filters` is the 8x6 tap table (SIXTAP_FILTERS or
 * BILINEAR_FILTERS, src/motion_comp.rs:294-337, or any other table);
 * `out` receives n * 384 bytes, each ReconstructedMb's y[256] u[64] v[64]
 * (src/reconstruct.rs:176-195).  n == 0 is a no-op. */
int vp8g_predict_inter_mb_batch(vp8g_ctx *ctx, const vp8g_ref_planes *ref,
                                const uint32_t *mb_xy, const vp8g_mv *mv, size_t n,
                                int full_pixel, const int32_t filters[8][6], uint8_t *out);
int vp8g_predict_inter_mb_batch_dev(vp8g_ctx *ctx, const vp8g_ref_planes *ref,
                                    const uint32_t *mb_xy, const vp8g_mv *mv, size_t n,
                                    int full_pixel, const int32_t filters[8][6], uint8_t *out);

/* ---- sub-pel SAD scoring ----
 * Replaces mb_luma_sad_at_mv (src/motion_search.rs:470-529) over a batch:
 * sad[i] = SAD of src[256*i .. +256] against the six-tap prediction of MB
 * (mb_xy[2i], mb_xy[2i+1]) at quarter-pel mv[i].  Whole-pel vectors give
 * mb_luma_sad_at_whole_mv (src/motion_search.rs:291-349). */
int vp8g_mb_luma_sad_at_mv_batch(vp8g_ctx *ctx, const vp8g_luma_ref *ref,
                                 const uint32_t *mb_xy, const uint8_t *src, const vp8g_mv *mv,
                                 size_t n, uint32_t *sad);
int vp8g_mb_luma_sad_at_mv_batch_dev(vp8g_ctx *ctx, const vp8g_luma_ref *ref,
                                     const uint32_t *mb_xy, const uint8_t *src,
                                     const vp8g_mv *mv, size_t n, uint32_t *sad);

/* Replaces half_pixel_refine_luma (src/motion_search.rs:564-627) and
 * quarter_pixel_refine_luma (src/motion_search.rs:666-728) over a batch.
 * Centre components must lie in [-1023, 1023] (VP8G_E_ARG otherwise; the
 * reference debug-asserts it). */
int vp8g_half_pixel_refine_luma_batch(vp8g_ctx *ctx, const vp8g_luma_ref *ref,
                                      const uint32_t *mb_xy, const uint8_t *src,
                                      const vp8g_mv *centers, size_t n,
                                      vp8g_search_result *out);
int vp8g_quarter_pixel_refine_luma_batch(vp8g_ctx *ctx, const vp8g_luma_ref *ref,
                                         const uint32_t *mb_xy, const uint8_t *src,
                                         const vp8g_mv *centers, size_t n,
                                         vp8g_search_result *out);

/* The same over device arrays (enqueue only).  mb_xy components must be at
 * most 0xFFFFF and centre components lie in [-1023, 1023] (components
 * outside are clamped to it). */
int vp8g_half_pixel_refine_luma_batch_dev(vp8g_ctx *ctx, const vp8g_luma_ref *ref,
                                          const uint32_t *mb_xy, const uint8_t *src,
                                          const vp8g_mv *centers, size_t n,
                                          vp8g_search_result *out);
int vp8g_quarter_pixel_refine_luma_batch_dev(vp8g_ctx *ctx, const vp8g_luma_ref *ref,
                                             const uint32_t *mb_xy, const uint8_t *src,
                                             const vp8g_mv *centers, size_t n,
                                             vp8g_search_result *out);

/* Replaces the encoder's half-pel probe set, fetch_window + HalfPlanes::new
 * + HalfPlanes::sad (src/encoder/inter_search.rs:266-346, used by
 * sub_pixel_refine :509-536), over a batch: block i's whole-pel position is
 * (xy[2i], xy[2i+1]) (MB origin + whole MV, any value, clamped reads), its
 * source src[256*i..]; `taps` is the caller's filters[4]; sad[8*i + k]
 * holds HalfPlanes::sad for (dx, dy) in (dy, dx) raster order without
 * (0, 0): (-1,-1) (0,-1) (1,-1) (-1,0) (1,0) (-1,1) (0,1) (1,1).  With
 * `diagonals` = 0 the four diagonal entries are 0xFFFFFFFF (not computed,
 * as HalfPlanes::new(.., false)). */
int vp8g_half_plane_sads_batch(vp8g_ctx *ctx, const vp8g_luma_ref *ref, const int32_t *xy,
                               const uint8_t *src, size_t n, const int32_t taps[6],
                               int diagonals, uint32_t *sad);

/* ---- stream driver: the scene-cut metric ----
 * Replaces luma_mad (src/stream.rs:2654-2666) over a batch of frame pairs:
 * mad[i] = (sum over the visible w x h luma of |prev[i] - cur[i]|) as f64
 * / (w * h) as f64.  Every frame of a call has the dims of prev[0]
 * (VP8G_E_REF_DIMS_MISMATCH otherwise: the stream driver locks them);
 * w, h in 1..16383.  The _dev form takes device planes and writes the
 * integer sums sad[i] (device) instead; enqueue only. */
int vp8g_luma_mad_batch(vp8g_ctx *ctx, const vp8g_i420 *prev, const vp8g_i420 *cur, size_t n, double *mad);
int vp8g_luma_sad_batch_dev(vp8g_ctx *ctx, const vp8g_i420 *prev, const vp8g_i420 *cur, size_t n, uint64_t *sad);

/* ---- SPLITMV prediction and inter-macroblock reconstruction ----
 * Replaces predict_split_mv (src/motion_comp.rs:2227-2320) over a batch:
 * macroblock i at (mb_xy[2i], mb_xy[2i+1]) with its 16 quarter-pel luma
 * sub-block vectors mvs[16*i .. 16*i+15] (raster order); every luma 4x4
 * sub-block through filter_block_4x4 (:1693-1712) under its own doubled
 * vector, every chroma 4x4 sub-block under split_chroma_mvs (:2162-2203),
 * no secondary clamp.  `filters` taps lie within +-2^20.  out: n x 384
 * (ReconstructedMb). */
int vp8g_predict_split_mv_batch(vp8g_ctx *ctx, const vp8g_ref_planes *ref, const uint32_t *mb_xy,
                                const vp8g_mv *mvs, size_t n, int full_pixel, const int32_t filters[8][6],
                                uint8_t *out);
/* Replaces reconstruct_inter_mb (src/motion_comp.rs:2098-2150) over a
 * batch: predict_inter_mb at mv[i], then unless skip[i] (mb_skip_coeff) the
 * residue of the dequantised coefficients coeffs[i] = y2[16], y[16][16],
 * u[4][16], v[4][16] (int16, 400 per macroblock): the inverse WHT of Y2
 * seeds each Y block's DC (src/inverse_transform.rs:170-240), and every
 * block's inverse DCT is added with clamp255 (:417-480). */
int vp8g_reconstruct_inter_mb_batch(vp8g_ctx *ctx, const vp8g_ref_planes *ref, const uint32_t *mb_xy,
                                    const vp8g_mv *mv, size_t n, int full_pixel, const int32_t filters[8][6],
                                    const uint8_t *skip, const int16_t *coeffs, uint8_t *out);
/* Replaces reconstruct_split_mv_mb (src/motion_comp.rs:2340-2385): as
 * predict_split_mv, then the residue of coeffs[i] = y[16][16], u[4][16],
 * v[4][16] (384 per macroblock; no Y2). */
int vp8g_reconstruct_split_mv_mb_batch(vp8g_ctx *ctx, const vp8g_ref_planes *ref, const uint32_t *mb_xy,
                                       const vp8g_mv *mvs, size_t n, int full_pixel, const int32_t filters[8][6],
                                       const uint8_t *skip, const int16_t *coeffs, uint8_t *out);
/* The three over device arrays (enqueue only). */
int vp8g_predict_split_mv_batch_dev(vp8g_ctx *ctx, const vp8g_ref_planes *ref, const uint32_t *mb_xy,
                                    const vp8g_mv *mvs, size_t n, int full_pixel, const int32_t filters[8][6],
                                    uint8_t *out);
int vp8g_reconstruct_inter_mb_batch_dev(vp8g_ctx *ctx, const vp8g_ref_planes *ref, const uint32_t *mb_xy,
                                        const vp8g_mv *mv, size_t n, int full_pixel, const int32_t filters[8][6],
                                        const uint8_t *skip, const int16_t *coeffs, uint8_t *out);
int vp8g_reconstruct_split_mv_mb_batch_dev(vp8g_ctx *ctx, const vp8g_ref_planes *ref, const uint32_t *mb_xy,
                                           const vp8g_mv *mvs, size_t n, int full_pixel,
                                           const int32_t filters[8][6], const uint8_t *skip, const int16_t *coeffs,
                                           uint8_t *out);

/* ---- encoder inter motion search ----
 * The caller side of sub-pel SAD scoring: the rate-distortion motion search
 * of the inter-frame encoder over a batch of macroblocks against one
 * reference.
 *
 * MvCostTable (src/encoder/inter_search.rs:166-194): bits[component][diff +
 * 1023] is the f32 bit cost of a motion-vector differential component
 * (component 0 = row), as MvCostTable::new tables it from the frame's MV
 * contexts.  The rate of a vector is lambda_sad * (bits[0][dr] + bits[1][dc])
 * (f32 sum widened to f64), and vectors whose differential leaves
 * [-1023, 1023] are never chosen. */
typedef struct {
    float bits[2][2047];
} vp8g_mv_cost_table;
/* The per-macroblock inputs of whole_pixel_search / sub_pixel_refine: the
 * macroblock position, the census predictor the differential is coded
 * against (RefSearch::census.best), the whole-pixel search range
 * (RefSearch::range_x / range_y, min <= max, within +-8191), up to 8 seed
 * vectors (quarter-pel), lambda_sad (finite, >= 0), good_enough
 * (< 2^30) and whether to descend past the seeds. */
typedef struct {
    uint32_t mb_col, mb_row;
    vp8g_mv best;
    int32_t range_x[2], range_y[2];
    uint32_t n_seeds;
    vp8g_mv seeds[8];
    double lambda_sad;
    uint32_t good_enough;
    int32_t descend;
} vp8g_mb_search;
/* SearchShape (src/encoder/inter_search.rs:66-93): the fields the two
 * searches read. */
typedef struct {
    uint32_t descent_rounds;
    int32_t wide_pattern, half_diagonals, quarter;
} vp8g_search_shape;
/* Found (src/encoder/inter_search.rs:349-355): the vector, its SAD and its
 * cost sad + lambda_sad * bits. */
typedef struct {
    vp8g_mv mv;
    uint32_t sad;
    double cost;
} vp8g_found;

/* Replaces whole_pixel_search (src/encoder/inter_search.rs:380-495) over a
 * batch: macroblock i of `jobs` with source src[256*i..] searched in `ref`
 * (LumaRef: width/height are the clamp bounds).  The seeds, rounded to
 * whole pixels and clamped into the range, are probed once each in order;
 * unless the best SAD is already <= good_enough (or descend is 0), the
 * two-pixel pattern (when shape->wide_pattern and the SAD exceeds 4 *
 * good_enough) then the one-pixel pattern descend for up to
 * descent_rounds rounds each.  out[i] receives the Found. */
int vp8g_whole_pixel_search_batch(vp8g_ctx *ctx, const vp8g_luma_ref *ref, const vp8g_mv_cost_table *costs,
                                  const vp8g_mb_search *jobs, const uint8_t *src, size_t n,
                                  const vp8g_search_shape *shape, vp8g_found *out);
/* Replaces sub_pixel_refine (src/encoder/inter_search.rs:499-552) over a
 * batch: from start[i] (whole-pel, as whole_pixel_search returns it) the
 * half-pel neighbours scored from HalfPlanes with filters[4] (the four
 * diagonals only with shape->half_diagonals), then with shape->quarter the
 * quarter-pel neighbours of the best through mb_luma_sad_at_mv; a
 * neighbour replaces the best when its cost is lower.  filters taps lie
 * within +-2^20. */
int vp8g_sub_pixel_refine_batch(vp8g_ctx *ctx, const vp8g_luma_ref *ref, const vp8g_mv_cost_table *costs,
                                const vp8g_mb_search *jobs, const uint8_t *src, const vp8g_found *start, size_t n,
                                const int32_t filters[8][6], const vp8g_search_shape *shape, vp8g_found *out);
/* Replaces small_diamond_search_luma (src/motion_search.rs:381-442) over a
 * batch: from centers[i] (clamped to +-1023, snapped toward zero to whole
 * pixels) up to max_iters rounds of the N, S, W, E one-pixel probes, strict
 * improvement only. */
int vp8g_small_diamond_search_luma_batch(vp8g_ctx *ctx, const vp8g_luma_ref *ref, const uint32_t *mb_xy,
                                         const uint8_t *src, const vp8g_mv *centers, size_t n, uint32_t max_iters,
                                         vp8g_search_result *out);
/* The three over device arrays (costs, jobs, src, start, mb_xy, centers,
 * out in device memory; `shape` and `filters` on the host); enqueue only.
 * Device inputs out of the documented ranges are clamped into them. */
int vp8g_whole_pixel_search_batch_dev(vp8g_ctx *ctx, const vp8g_luma_ref *ref, const vp8g_mv_cost_table *costs,
                                      const vp8g_mb_search *jobs, const uint8_t *src, size_t n,
                                      const vp8g_search_shape *shape, vp8g_found *out);
int vp8g_sub_pixel_refine_batch_dev(vp8g_ctx *ctx, const vp8g_luma_ref *ref, const vp8g_mv_cost_table *costs,
                                    const vp8g_mb_search *jobs, const uint8_t *src, const vp8g_found *start,
                                    size_t n, const int32_t filters[8][6], const vp8g_search_shape *shape,
                                    vp8g_found *out);
int vp8g_small_diamond_search_luma_batch_dev(vp8g_ctx *ctx, const vp8g_luma_ref *ref, const uint32_t *mb_xy,
                                             const uint8_t *src, const vp8g_mv *centers, size_t n,
                                             uint32_t max_iters, vp8g_search_result *out);

/* ---- loop filter (RFC 6386 section 15) ----
 * FrameFilterConfig (src/loop_filter.rs:891-944) with the section 20.6
 * arrays: ref_delta[CURRENT, LAST, GOLDEN, ALTREF] and mode_delta[B_PRED,
 * ZEROMV, NEAREST/NEAR/NEW, SPLITMV] (FrameFilterConfig::ref_deltas /
 * mode_deltas, src/loop_filter.rs:1193-1209).  Boolean fields are 0 / 1;
 * loop_filter_level and sharpness_level take the u8 range. */
typedef struct {
    int32_t simple, key_frame, loop_filter_level, sharpness_level;
    int32_t segmentation_enabled, segment_abs, delta_enabled;
    int16_t segment_lf_level[4];
    int16_t ref_delta[4];
    int16_t mode_delta[4];
} vp8g_lf_config;

/* One macroblock's filter inputs, raster order: MacroblockModes::segment_id
 * (VP8G_LF_NO_SEGMENT for None) and y_mode (VP8G_LF_BPRED when B_PRED; the
 * decoder records SPLITMV MBs as B_PRED, src/loop_filter.rs:1301-1308), the
 * reference frame (0 intra, 1 LAST, 2 GOLDEN, 3 ALTREF), the inter mode of
 * an inter MB (VP8G_LF_ZEROMV, VP8G_LF_SPLITMV, anything else is
 * NEAREST/NEAR/NEW) and mb_has_coeffs (src/loop_filter.rs:1500-1505) as
 * VP8G_LF_HAS_COEFFS.  Four bytes, 4-byte aligned arrays. */
typedef struct {
    uint8_t segment_id, ref_frame, inter_mode, flags;
} vp8g_lf_mb;
#define VP8G_LF_NO_SEGMENT 255
#define VP8G_LF_ZEROMV 1
#define VP8G_LF_SPLITMV 2
#define VP8G_LF_BPRED 1
#define VP8G_LF_HAS_COEFFS 2

/* KeyframePlanes (src/frame.rs): the MB-padded reconstruction, mb_rows*16
 * rows of y_stride >= mb_cols*16 bytes, mb_rows*8 rows of uv_stride >=
 * mb_cols*8 for u and v; at most 1024 x 1024 MBs and below 4 GiB per
 * plane.  In device memory y and y_stride are multiples of 16, u, v and
 * uv_stride multiples of 8. */
typedef struct {
    uint8_t *y, *u, *v;
    size_t y_stride, uv_stride, mb_cols, mb_rows;
} vp8g_lf_planes;

/* SourcePlanes (src/loop_filter.rs:1916-1932): the source picture and its
 * visible width x height window (at most the padded frame); chroma covers
 * ceil(w/2) x ceil(h/2).  Any alignment. */
typedef struct {
    size_t width, height;
    const uint8_t *y, *u, *v;
    size_t y_stride, uv_stride;
} vp8g_lf_source;

/* Replaces filter_inter_frame_flags (src/loop_filter.rs:1364-1496) and, with
 * every MB intra, filter_frame_flags (src/loop_filter.rs:2506-2636), the
 * loops filter_frame_flags_threaded / filter_inter_frame_flags_threaded
 * (src/loop_filter.rs:2435-2504, src/loop_filter/bands.rs) run on several
 * threads: the whole frame filtered in place, byte-identical to the raster
 * pass.  `mbs` has mb_cols * mb_rows entries.  _dev: planes and mbs in
 * device memory, enqueue only; the host form copies planes and mbs in and
 * the filtered planes back (synchronous). */
int vp8g_loop_filter_frame_dev(vp8g_ctx *ctx, const vp8g_lf_planes *planes, const vp8g_lf_mb *mbs,
                               const vp8g_lf_config *cfg);
int vp8g_loop_filter_frame(vp8g_ctx *ctx, const vp8g_lf_planes *planes, const vp8g_lf_mb *mbs,
                           const vp8g_lf_config *cfg);

/* Replaces score_filter_candidate over score_candidates
 * (src/loop_filter.rs:2062-2131): ssd[i] = the visible-window SSD against
 * `src` of the reconstruction filtered with cfg's level and sharpness
 * replaced by levels[i], sharps[i] (level 0: the unfiltered SSD, whatever
 * the segment overrides).  every_rows 0 scores the whole frame; every_rows
 * >= 2 scores the bands of select_filter_params_sampled
 * (src/loop_filter.rs:2303-2365): two MB rows starting every every_rows
 * rows, each filtered as a frame of its own.  recon, mbs, src and ssd are
 * device memory (recon is only read); levels and sharps are host arrays of
 * n <= 256 entries.  Enqueue only. */
int vp8g_loop_filter_score_dev(vp8g_ctx *ctx, const vp8g_lf_planes *recon, const vp8g_lf_mb *mbs,
                               const vp8g_lf_config *cfg, const vp8g_lf_source *src, const uint8_t *levels,
                               const uint8_t *sharps, size_t n, size_t every_rows, uint64_t *ssd);

/* Replaces select_filter_params_threaded (src/loop_filter.rs:2133-2200) when
 * every_rows is 0, and select_filter_params_sampled
 * (src/loop_filter.rs:2288-2412, with sweep_sharpness; without it
 * select_filter_level_sampled, :2224-2258) when every_rows >= 2: the
 * (level, sharpness) the reference's search picks, first strict minimum at
 * every stage.  The searches' level stages are scored in one pass over all
 * 64 levels, the sharpness sweep in a second.  _dev: recon, mbs and src in
 * device memory; the host form copies them in.  Synchronous. */
int vp8g_loop_filter_select_dev(vp8g_ctx *ctx, const vp8g_lf_planes *recon, const vp8g_lf_mb *mbs,
                                const vp8g_lf_config *cfg, const vp8g_lf_source *src, size_t every_rows,
                                int sweep_sharpness, uint8_t *level, uint8_t *sharpness);
int vp8g_loop_filter_select(vp8g_ctx *ctx, const vp8g_lf_planes *recon, const vp8g_lf_mb *mbs,
                            const vp8g_lf_config *cfg, const vp8g_lf_source *src, size_t every_rows,
                            int sweep_sharpness, uint8_t *level, uint8_t *sharpness);

/* Waits for the context's stream and reports whether any loop-filter launch
 * since the last call gave up on a row wait (VP8G_E_DEVICE; its output is
 * then incomplete), then clears the flag.  The row schedule cannot stall,
 * so this reports a broken invariant, not a normal outcome. */
int vp8g_loop_filter_status(vp8g_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
