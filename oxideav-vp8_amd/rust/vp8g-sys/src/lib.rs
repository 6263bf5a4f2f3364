//! Raw bindings to `libvp8g`, the gfx950 (MI355X) implementation of
//! oxideav-vp8's encoder-side data-parallel path: ARNR altref synthesis,
//! VP8 motion compensation, sub-pel SAD scoring and the encoder's inter
//! motion search.
//!
//! One item per declaration of `include/vp8g.h`, in the header's order,
//! with the same names, field order and argument order.  C array
//! parameters appear as the pointers they decay to (`const int32_t
//! filters[8][6]` is `*const [i32; 6]`, pointing at the first of 8 rows).
//!
//! Memory: entry points without a suffix take host pointers and return
//! after the result is in place; `_dev` entry points take device pointers,
//! enqueue on the context's stream and return without waiting
//! (`vp8g_ctx_synchronize` before reading).  A context is not internally
//! locked: use one per host thread.
#![no_std]
#![allow(non_camel_case_types)]

use core::ffi::{c_char, c_int, c_void};
use core::marker::{PhantomData, PhantomPinned};

// ---------------------------------------------------------------- status

pub const VP8G_OK: c_int = 0;
/// `EncodeError::InvalidDimensions` (src/encoder.rs:72).
pub const VP8G_E_INVALID_DIMENSIONS: c_int = 1;
/// `EncodeError::ReferenceDimensionsMismatch` (src/encoder.rs:99-104).
pub const VP8G_E_REF_DIMS_MISMATCH: c_int = 2;
/// The reference panics (src/arnr.rs:494-498).
pub const VP8G_E_CENTER_OUT_OF_RANGE: c_int = 3;
/// Null pointer, short stride, limit exceeded.
pub const VP8G_E_ARG: c_int = 4;
/// HIP runtime error; no output was written.
pub const VP8G_E_DEVICE: c_int = 5;
/// No usable gfx950 device.
pub const VP8G_E_NO_DEVICE: c_int = 6;

/// Frames per ARNR call (the descriptors travel in kernel arguments).
pub const VP8G_MAX_FRAMES: c_int = 64;

// ---------------------------------------------------------------- types

/// Opaque per-thread context (device, stream, arenas).
#[repr(C)]
pub struct vp8g_ctx {
    _private: [u8; 0],
    _marker: PhantomData<(*mut u8, PhantomPinned)>,
}

/// `I420Frame` (src/encoder.rs:5419-5434): visible dims, planes, strides.
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct vp8g_i420 {
    pub width: u32,
    pub height: u32,
    pub y: *const u8,
    pub u: *const u8,
    pub v: *const u8,
    pub y_stride: usize,
    pub uv_stride: usize,
}

/// `EncodeError` payloads: `InvalidDimensions { width, height }` in
/// `src_w` / `src_h`; `ReferenceDimensionsMismatch { source, reference }`
/// in all four.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq)]
pub struct vp8g_err {
    pub src_w: u32,
    pub src_h: u32,
    pub ref_w: u32,
    pub ref_h: u32,
}

/// `Mv` (src/motion_vector.rs:64-70), quarter-pel.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq)]
pub struct vp8g_mv {
    pub row: i16,
    pub col: i16,
}

/// `ReferencePlanes` (src/motion_comp.rs:131-146), macroblock-aligned.
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct vp8g_ref_planes {
    pub y: *const u8,
    pub u: *const u8,
    pub v: *const u8,
    pub y_stride: usize,
    pub uv_stride: usize,
    pub mb_cols: usize,
    pub mb_rows: usize,
}

/// `LumaRef` (src/motion_search.rs:230-239); width / height are the clamp
/// bounds.
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct vp8g_luma_ref {
    pub plane: *const u8,
    pub stride: usize,
    pub width: usize,
    pub height: usize,
}

/// `SearchResult` (src/motion_search.rs:248-255).
#[repr(C)]
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq)]
pub struct vp8g_search_result {
    pub mv: vp8g_mv,
    pub sad: u32,
}

// ---------------------------------------------------------------- context

/// Launch-shape options (`vp8g_ctx_set_option`); none changes an output byte.
pub const VP8G_OPT_ARNR_WAVES: c_int = 1;
pub const VP8G_OPT_XCD_REMAP: c_int = 2;
pub const VP8G_OPT_ARNR_MBS: c_int = 3;
pub const VP8G_OPT_BATCH_WGS_PER_CU: c_int = 4;
/// Host-side fault injection for callers' error paths.
pub const VP8G_OPT_FAULT: c_int = 5;
pub const VP8G_FAULT_NONE: c_int = 0;
pub const VP8G_FAULT_ALLOC: c_int = 1;
pub const VP8G_FAULT_COPY: c_int = 2;
pub const VP8G_FAULT_LAUNCH: c_int = 3;
/// Host threads of `vp8g_build_arnr_altref`'s packing and copy-out.
pub const VP8G_OPT_HOST_THREADS: c_int = 6;
/// Streams carrying `vp8g_build_arnr_altref`'s host<->device copies.
pub const VP8G_OPT_COPY_STREAMS: c_int = 7;
/// Row bands of `vp8g_build_arnr_altref` (1..8; 0 = default).
pub const VP8G_OPT_HOST_BANDS: c_int = 8;

/// `vp8g_search_fallback[_noalloc]_dev` flag: queue only failed entries
/// with a passing left, right, upper or lower neighbour.
pub const VP8G_FALLBACK_GATE_NEIGHBOUR: c_int = 1;

/// Opaque learned motion proposer (no counterpart in the reference).
#[repr(C)]
pub struct vp8g_proposer {
    _private: [u8; 0],
    _marker: PhantomData<(*mut u8, PhantomPinned)>,
}

/// A rectangle of per-frame noise in the synthetic training frames.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct vp8g_synth_rect {
    pub seed: u64,
    pub x0: u32,
    pub y0: u32,
    pub w: u32,
    pub h: u32,
}

/// Synthetic region-motion luma (training data for the proposer).
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct vp8g_synth_spec {
    pub width: u32,
    pub height: u32,
    pub n: usize,
    pub rects_per_frame: usize,
    pub texture_seed: u64,
    /// host `[64][2]`
    pub vel: *const i32,
    /// host `[n][rects_per_frame]`
    pub rects: *const vp8g_synth_rect,
    /// host `[n]`
    pub noise_seeds: *const u64,
    pub sigma: f64,
}

/// One group of a batched ARNR launch.
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct vp8g_arnr_job {
    /// host array of `n_frames` device-plane frames
    pub frames: *const vp8g_i420,
    pub n_frames: usize,
    pub center: usize,
    pub strength: u8,
    pub out_y: *mut u8,
    pub out_u: *mut u8,
    pub out_v: *mut u8,
    /// device alignment table, may be null
    pub mv_out: *mut i32,
}

/// Opaque graph-captured batch of ARNR jobs.
#[repr(C)]
pub struct vp8g_arnr_batch {
    _private: [u8; 0],
    _marker: PhantomData<(*mut u8, PhantomPinned)>,
}

/// `MvCostTable` (src/encoder/inter_search.rs:166-194):
/// `bits[component][diff + 1023]`, component 0 = row.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vp8g_mv_cost_table {
    pub bits: [[f32; 2047]; 2],
}

/// The per-macroblock inputs of `whole_pixel_search` / `sub_pixel_refine`
/// (src/encoder/inter_search.rs:380-552).
#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct vp8g_mb_search {
    pub mb_col: u32,
    pub mb_row: u32,
    /// `RefSearch::census.best`, the predictor differentials are coded against
    pub best: vp8g_mv,
    pub range_x: [i32; 2],
    pub range_y: [i32; 2],
    pub n_seeds: u32,
    pub seeds: [vp8g_mv; 8],
    pub lambda_sad: f64,
    pub good_enough: u32,
    pub descend: i32,
}

/// `SearchShape` (src/encoder/inter_search.rs:66-93): the fields the two
/// searches read.
#[repr(C)]
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq)]
pub struct vp8g_search_shape {
    pub descent_rounds: u32,
    pub wide_pattern: i32,
    pub half_diagonals: i32,
    pub quarter: i32,
}

/// `Found` (src/encoder/inter_search.rs:349-355).
#[repr(C)]
#[derive(Debug, Clone, Copy, Default, PartialEq)]
pub struct vp8g_found {
    pub mv: vp8g_mv,
    pub sad: u32,
    pub cost: f64,
}

/// `FrameFilterConfig` (src/loop_filter.rs:891-944) with `ref_delta` =
/// [CURRENT, LAST, GOLDEN, ALTREF] and `mode_delta` = [B_PRED, ZEROMV,
/// NEAREST/NEAR/NEW, SPLITMV] (src/loop_filter.rs:1193-1209).
#[repr(C)]
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq)]
pub struct vp8g_lf_config {
    pub simple: i32,
    pub key_frame: i32,
    pub loop_filter_level: i32,
    pub sharpness_level: i32,
    pub segmentation_enabled: i32,
    pub segment_abs: i32,
    pub delta_enabled: i32,
    pub segment_lf_level: [i16; 4],
    pub ref_delta: [i16; 4],
    pub mode_delta: [i16; 4],
}

/// One macroblock's loop-filter inputs (segment id, ref frame, inter mode,
/// `VP8G_LF_BPRED | VP8G_LF_HAS_COEFFS`).
#[repr(C)]
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq)]
pub struct vp8g_lf_mb {
    pub segment_id: u8,
    pub ref_frame: u8,
    pub inter_mode: u8,
    pub flags: u8,
}

/// `KeyframePlanes` (src/frame.rs): the MB-padded reconstruction.
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct vp8g_lf_planes {
    pub y: *mut u8,
    pub u: *mut u8,
    pub v: *mut u8,
    pub y_stride: usize,
    pub uv_stride: usize,
    pub mb_cols: usize,
    pub mb_rows: usize,
}

/// `SourcePlanes` (src/loop_filter.rs:1916-1932).
#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct vp8g_lf_source {
    pub width: usize,
    pub height: usize,
    pub y: *const u8,
    pub u: *const u8,
    pub v: *const u8,
    pub y_stride: usize,
    pub uv_stride: usize,
}

pub const VP8G_LF_NO_SEGMENT: c_int = 255;
pub const VP8G_LF_ZEROMV: c_int = 1;
pub const VP8G_LF_SPLITMV: c_int = 2;
pub const VP8G_LF_BPRED: c_int = 1;
pub const VP8G_LF_HAS_COEFFS: c_int = 2;

extern "C" {
    // ------------------------------------------------------------ context
    pub fn vp8g_ctx_create(device: c_int, out: *mut *mut vp8g_ctx) -> c_int;
    pub fn vp8g_ctx_destroy(ctx: *mut vp8g_ctx);
    /// `stream` is a `hipStream_t`; null restores the context's own stream.
    pub fn vp8g_ctx_set_stream(ctx: *mut vp8g_ctx, stream: *mut c_void) -> c_int;
    pub fn vp8g_ctx_stream(ctx: *mut vp8g_ctx) -> *mut c_void;
    pub fn vp8g_ctx_synchronize(ctx: *mut vp8g_ctx) -> c_int;
    pub fn vp8g_ctx_set_option(ctx: *mut vp8g_ctx, option: c_int, value: c_int) -> c_int;
    pub fn vp8g_status_str(status: c_int) -> *const c_char;

    // ------------------------------------------------------------ tuning
    pub fn vp8g_ctx_load_tuning(ctx: *mut vp8g_ctx, path: *const c_char) -> c_int;
    pub fn vp8g_tuning_lookup(path: *const c_char, width: u32, height: u32, n_frames: usize, batch: c_int,
                              mbs: *mut c_int, waves: *mut c_int, wgs_per_cu: *mut c_int) -> c_int;
    pub fn vp8g_arnr_shape(ctx: *mut vp8g_ctx, width: u32, height: u32, n_frames: usize, strength: u8,
                           mbs: *mut c_int, waves: *mut c_int) -> c_int;

    // ------------------------------------------------------------ ARNR altref synthesis
    /// `build_arnr_altref` (src/arnr.rs:476-482) and
    /// `build_arnr_altref_threaded` (src/arnr.rs:488-650).
    pub fn vp8g_build_arnr_altref(ctx: *mut vp8g_ctx, frames: *const vp8g_i420, n_frames: usize, center: usize,
                                  strength: u8, out_y: *mut u8, out_u: *mut u8, out_v: *mut u8,
                                  err: *mut vp8g_err) -> c_int;
    pub fn vp8g_build_arnr_altref_dev(ctx: *mut vp8g_ctx, frames: *const vp8g_i420, n_frames: usize, center: usize,
                                      strength: u8, out_y: *mut u8, out_u: *mut u8, out_v: *mut u8,
                                      err: *mut vp8g_err, mv_out: *mut i32) -> c_int;

    // ------------------------------------------------------------ block search, table-driven altref
    pub fn vp8g_block_search_dev(ctx: *mut vp8g_ctx, frames: *const vp8g_i420, n_frames: usize, center: usize,
                                 radius: c_int, seeds: *const i32, out: *mut i32, err: *mut vp8g_err) -> c_int;
    pub fn vp8g_refine_candidates_dev(ctx: *mut vp8g_ctx, frames: *const vp8g_i420, n_frames: usize,
                                      center: usize, seeds: *const i32, first: *const i32, out: *mut i32,
                                      err: *mut vp8g_err) -> c_int;
    pub fn vp8g_search_fallback_dev(ctx: *mut vp8g_ctx, frames: *const vp8g_i420, n_frames: usize, center: usize,
                                    first: *const i32, radius: c_int, flags: c_int, out: *mut i32,
                                    queued: *mut u32, err: *mut vp8g_err) -> c_int;
    /// Allocation-free form for callers that cannot allocate (graph
    /// capture, serving loops): static 4 MiB queue per GPU, at most 2^20
    /// entries; overlapping calls on one GPU race.
    pub fn vp8g_search_fallback_noalloc_dev(ctx: *mut vp8g_ctx, frames: *const vp8g_i420, n_frames: usize,
                                            center: usize, first: *const i32, radius: c_int, flags: c_int,
                                            out: *mut i32, queued: *mut u32, err: *mut vp8g_err) -> c_int;
    pub fn vp8g_build_arnr_altref_table_dev(ctx: *mut vp8g_ctx, frames: *const vp8g_i420, n_frames: usize,
                                            center: usize, strength: u8, table: *const i32, out_y: *mut u8,
                                            out_u: *mut u8, out_v: *mut u8, err: *mut vp8g_err) -> c_int;

    // ------------------------------------------------------------ coarse-to-fine pyramid
    pub fn vp8g_downsample2_i420_dev(ctx: *mut vp8g_ctx, src: *const vp8g_i420, dst: *const vp8g_i420,
                                     n_frames: usize) -> c_int;
    pub fn vp8g_upsample_table_dev(ctx: *mut vp8g_ctx, coarse: *const i32, n_frames: usize, coarse_w: c_int,
                                   coarse_h: c_int, fine: *mut i32, fine_w: c_int, fine_h: c_int) -> c_int;

    // ------------------------------------------------------------ learned motion proposer
    pub fn vp8g_proposer_param_count() -> usize;
    pub fn vp8g_proposer_create(ctx: *mut vp8g_ctx, out: *mut *mut vp8g_proposer) -> c_int;
    pub fn vp8g_proposer_destroy(p: *mut vp8g_proposer);
    pub fn vp8g_proposer_set_params(p: *mut vp8g_proposer, params: *const f32, n: usize) -> c_int;
    pub fn vp8g_proposer_get_params(p: *mut vp8g_proposer, params: *mut f32, n: usize) -> c_int;
    pub fn vp8g_proposer_get_grads(p: *mut vp8g_proposer, grads: *mut f32, n: usize, loss: *mut f32) -> c_int;
    pub fn vp8g_proposer_prepare_dev(p: *mut vp8g_proposer, frames: *const vp8g_i420, n_frames: usize,
                                     center: usize, planes: *mut u8) -> c_int;
    pub fn vp8g_proposer_forward_dev(p: *mut vp8g_proposer, planes: *const u8, images: c_int, blocks_w: c_int,
                                     blocks_h: c_int, a4_out: *mut u16) -> c_int;
    pub fn vp8g_proposer_propose_dev(p: *mut vp8g_proposer, frames: *const vp8g_i420, n_frames: usize,
                                     center: usize, clamp: c_int, table: *mut i32) -> c_int;
    pub fn vp8g_proposer_backward_dev(p: *mut vp8g_proposer, planes: *const u8, labels: *const i32, images: c_int,
                                      blocks_w: c_int, blocks_h: c_int) -> c_int;
    pub fn vp8g_proposer_train_step_dev(p: *mut vp8g_proposer, planes: *const u8, labels: *const i32,
                                        images: c_int, blocks_w: c_int, blocks_h: c_int, lr: f32,
                                        loss_dev: *mut f32) -> c_int;
    pub fn vp8g_synth_luma_dev(ctx: *mut vp8g_ctx, spec: *const vp8g_synth_spec, out: *mut u8,
                               clean: *mut u8) -> c_int;

    // ------------------------------------------------------------ batched ARNR
    /// Many groups in one persistent, graph-captured launch: a loop of
    /// `build_arnr_altref_threaded` calls (src/arnr.rs:488-650) as
    /// `flush_group` (src/stream.rs:2806-2957) issues them.
    pub fn vp8g_arnr_batch_create(ctx: *mut vp8g_ctx, jobs: *const vp8g_arnr_job, n_jobs: usize,
                                  out: *mut *mut vp8g_arnr_batch, failed_job: *mut usize,
                                  err: *mut vp8g_err) -> c_int;
    pub fn vp8g_arnr_batch_run(batch: *mut vp8g_arnr_batch) -> c_int;
    pub fn vp8g_arnr_batch_shape(batch: *const vp8g_arnr_batch, mbs: *mut c_int, waves: *mut c_int,
                                 workgroups: *mut c_int, units: *mut usize) -> c_int;
    pub fn vp8g_arnr_batch_destroy(batch: *mut vp8g_arnr_batch);

    // ------------------------------------------------------------ motion compensation
    /// `predict_inter_mb` (src/motion_comp.rs:1986-2078), n macroblocks;
    /// `out` receives n x 384 bytes (`ReconstructedMb` y, u, v).
    pub fn vp8g_predict_inter_mb_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_ref_planes, mb_xy: *const u32,
                                       mv: *const vp8g_mv, n: usize, full_pixel: c_int, filters: *const [i32; 6],
                                       out: *mut u8) -> c_int;
    pub fn vp8g_predict_inter_mb_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_ref_planes, mb_xy: *const u32,
                                           mv: *const vp8g_mv, n: usize, full_pixel: c_int,
                                           filters: *const [i32; 6], out: *mut u8) -> c_int;

    // ------------------------------------------------------------ sub-pel SAD scoring
    /// `mb_luma_sad_at_mv` (src/motion_search.rs:470-529), n macroblocks.
    pub fn vp8g_mb_luma_sad_at_mv_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref, mb_xy: *const u32,
                                        src: *const u8, mv: *const vp8g_mv, n: usize, sad: *mut u32) -> c_int;
    pub fn vp8g_mb_luma_sad_at_mv_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref, mb_xy: *const u32,
                                            src: *const u8, mv: *const vp8g_mv, n: usize, sad: *mut u32) -> c_int;
    /// `half_pixel_refine_luma` (src/motion_search.rs:564-627).
    pub fn vp8g_half_pixel_refine_luma_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref, mb_xy: *const u32,
                                             src: *const u8, centers: *const vp8g_mv, n: usize,
                                             out: *mut vp8g_search_result) -> c_int;
    /// `quarter_pixel_refine_luma` (src/motion_search.rs:666-728).
    pub fn vp8g_quarter_pixel_refine_luma_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                                mb_xy: *const u32, src: *const u8, centers: *const vp8g_mv,
                                                n: usize, out: *mut vp8g_search_result) -> c_int;
    pub fn vp8g_half_pixel_refine_luma_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                                 mb_xy: *const u32, src: *const u8, centers: *const vp8g_mv,
                                                 n: usize, out: *mut vp8g_search_result) -> c_int;
    pub fn vp8g_quarter_pixel_refine_luma_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                                    mb_xy: *const u32, src: *const u8, centers: *const vp8g_mv,
                                                    n: usize, out: *mut vp8g_search_result) -> c_int;
    /// `HalfPlanes::new` + `HalfPlanes::sad` (src/encoder/inter_search.rs:266-346);
    /// `taps` points at the caller's `filters[4]` (6 taps).
    pub fn vp8g_half_plane_sads_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref, xy: *const i32,
                                      src: *const u8, n: usize, taps: *const i32, diagonals: c_int,
                                      sad: *mut u32) -> c_int;

    // ------------------------------------------------------------ stream driver
    /// `luma_mad` (src/stream.rs:2654-2666) over frame pairs.
    pub fn vp8g_luma_mad_batch(ctx: *mut vp8g_ctx, prev: *const vp8g_i420, cur: *const vp8g_i420, n: usize,
                               mad: *mut f64) -> c_int;
    pub fn vp8g_luma_sad_batch_dev(ctx: *mut vp8g_ctx, prev: *const vp8g_i420, cur: *const vp8g_i420, n: usize,
                                   sad: *mut u64) -> c_int;

    // ------------------------------------------------------------ SPLITMV and reconstruction
    /// `predict_split_mv` (src/motion_comp.rs:2227-2320); `mvs` holds 16
    /// vectors per macroblock.
    pub fn vp8g_predict_split_mv_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_ref_planes, mb_xy: *const u32,
                                       mvs: *const vp8g_mv, n: usize, full_pixel: c_int,
                                       filters: *const [i32; 6], out: *mut u8) -> c_int;
    /// `reconstruct_inter_mb` (src/motion_comp.rs:2098-2150); 400
    /// coefficients per macroblock (y2, y, u, v).
    pub fn vp8g_reconstruct_inter_mb_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_ref_planes, mb_xy: *const u32,
                                           mv: *const vp8g_mv, n: usize, full_pixel: c_int,
                                           filters: *const [i32; 6], skip: *const u8, coeffs: *const i16,
                                           out: *mut u8) -> c_int;
    /// `reconstruct_split_mv_mb` (src/motion_comp.rs:2340-2385); 384
    /// coefficients per macroblock (y, u, v).
    pub fn vp8g_reconstruct_split_mv_mb_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_ref_planes,
                                              mb_xy: *const u32, mvs: *const vp8g_mv, n: usize, full_pixel: c_int,
                                              filters: *const [i32; 6], skip: *const u8, coeffs: *const i16,
                                              out: *mut u8) -> c_int;
    pub fn vp8g_predict_split_mv_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_ref_planes, mb_xy: *const u32,
                                           mvs: *const vp8g_mv, n: usize, full_pixel: c_int,
                                           filters: *const [i32; 6], out: *mut u8) -> c_int;
    pub fn vp8g_reconstruct_inter_mb_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_ref_planes,
                                               mb_xy: *const u32, mv: *const vp8g_mv, n: usize, full_pixel: c_int,
                                               filters: *const [i32; 6], skip: *const u8, coeffs: *const i16,
                                               out: *mut u8) -> c_int;
    pub fn vp8g_reconstruct_split_mv_mb_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_ref_planes,
                                                  mb_xy: *const u32, mvs: *const vp8g_mv, n: usize,
                                                  full_pixel: c_int, filters: *const [i32; 6], skip: *const u8,
                                                  coeffs: *const i16, out: *mut u8) -> c_int;

    // ------------------------------------------------------------ encoder inter motion search
    /// `whole_pixel_search` (src/encoder/inter_search.rs:380-495).
    pub fn vp8g_whole_pixel_search_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                         costs: *const vp8g_mv_cost_table, jobs: *const vp8g_mb_search,
                                         src: *const u8, n: usize, shape: *const vp8g_search_shape,
                                         out: *mut vp8g_found) -> c_int;
    /// `sub_pixel_refine` (src/encoder/inter_search.rs:499-552).
    pub fn vp8g_sub_pixel_refine_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                       costs: *const vp8g_mv_cost_table, jobs: *const vp8g_mb_search,
                                       src: *const u8, start: *const vp8g_found, n: usize,
                                       filters: *const [i32; 6], shape: *const vp8g_search_shape,
                                       out: *mut vp8g_found) -> c_int;
    /// `small_diamond_search_luma` (src/motion_search.rs:381-442).
    pub fn vp8g_small_diamond_search_luma_batch(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                                mb_xy: *const u32, src: *const u8, centers: *const vp8g_mv,
                                                n: usize, max_iters: u32, out: *mut vp8g_search_result) -> c_int;
    pub fn vp8g_whole_pixel_search_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                             costs: *const vp8g_mv_cost_table, jobs: *const vp8g_mb_search,
                                             src: *const u8, n: usize, shape: *const vp8g_search_shape,
                                             out: *mut vp8g_found) -> c_int;
    pub fn vp8g_sub_pixel_refine_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                           costs: *const vp8g_mv_cost_table, jobs: *const vp8g_mb_search,
                                           src: *const u8, start: *const vp8g_found, n: usize,
                                           filters: *const [i32; 6], shape: *const vp8g_search_shape,
                                           out: *mut vp8g_found) -> c_int;
    pub fn vp8g_small_diamond_search_luma_batch_dev(ctx: *mut vp8g_ctx, r#ref: *const vp8g_luma_ref,
                                                    mb_xy: *const u32, src: *const u8, centers: *const vp8g_mv,
                                                    n: usize, max_iters: u32,
                                                    out: *mut vp8g_search_result) -> c_int;
    // ------------------------------------------------------------ loop filter
    pub fn vp8g_loop_filter_frame_dev(ctx: *mut vp8g_ctx, planes: *const vp8g_lf_planes, mbs: *const vp8g_lf_mb,
                                      cfg: *const vp8g_lf_config) -> c_int;
    pub fn vp8g_loop_filter_frame(ctx: *mut vp8g_ctx, planes: *const vp8g_lf_planes, mbs: *const vp8g_lf_mb,
                                  cfg: *const vp8g_lf_config) -> c_int;
    pub fn vp8g_loop_filter_score_dev(ctx: *mut vp8g_ctx, recon: *const vp8g_lf_planes, mbs: *const vp8g_lf_mb,
                                      cfg: *const vp8g_lf_config, src: *const vp8g_lf_source, levels: *const u8,
                                      sharps: *const u8, n: usize, every_rows: usize, ssd: *mut u64) -> c_int;
    pub fn vp8g_loop_filter_select_dev(ctx: *mut vp8g_ctx, recon: *const vp8g_lf_planes, mbs: *const vp8g_lf_mb,
                                       cfg: *const vp8g_lf_config, src: *const vp8g_lf_source, every_rows: usize,
                                       sweep_sharpness: c_int, level: *mut u8, sharpness: *mut u8) -> c_int;
    pub fn vp8g_loop_filter_select(ctx: *mut vp8g_ctx, recon: *const vp8g_lf_planes, mbs: *const vp8g_lf_mb,
                                   cfg: *const vp8g_lf_config, src: *const vp8g_lf_source, every_rows: usize,
                                   sweep_sharpness: c_int, level: *mut u8, sharpness: *mut u8) -> c_int;
    pub fn vp8g_loop_filter_status(ctx: *mut vp8g_ctx) -> c_int;
}
