//! Safe drop-ins for oxideav-vp8's data-parallel encoder functions, run on
//! a gfx950 (MI355X) GPU through `vp8g-sys`.
//!
//! Each function keeps the signature, argument meaning, output bytes and
//! error behaviour of the oxideav-vp8 function it replaces (cited on each
//! item).  The types mirror the reference's field for field, so a caller
//! converts with a struct literal.  Functions with the reference's
//! signature panic where the reference panics and, since the reference
//! cannot fail past its own checks, also on a device failure (message:
//! the library's status string); the `*_batch` forms return
//! `Result<_, DeviceError>` instead and take a whole frame of macroblocks
//! per call, which is where the GPU pays.
//!
//! Every call runs on a per-thread context (device `$VP8G_DEVICE`, default
//! 0) created on first use.
use std::cell::RefCell;
use std::ffi::CStr;
use std::fmt;
use std::ptr::{self, NonNull};

use vp8g_sys as sys;

// ---------------------------------------------------------------- types

/// `I420Frame` (src/encoder.rs:5419-5434).
#[derive(Debug, Clone, Copy)]
pub struct I420Frame<'a> {
    pub width: u32,
    pub height: u32,
    pub y: &'a [u8],
    pub u: &'a [u8],
    pub v: &'a [u8],
    pub y_stride: usize,
    pub uv_stride: usize,
}

/// `ArnrConfig` (src/arnr.rs:52-56).
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub struct ArnrConfig {
    pub strength: u8,
}

impl Default for ArnrConfig {
    fn default() -> Self {
        ArnrConfig { strength: 3 }
    }
}

/// `ArnrPicture` (src/arnr.rs:90-101): tightly packed planes.
#[derive(Debug, Clone, PartialEq, Eq)]
pub struct ArnrPicture {
    pub width: u32,
    pub height: u32,
    pub y: Vec<u8>,
    pub u: Vec<u8>,
    pub v: Vec<u8>,
}

/// The two `EncodeError` variants (src/encoder.rs:69-104) ARNR returns.
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum EncodeError {
    InvalidDimensions { width: u32, height: u32 },
    ReferenceDimensionsMismatch { source: (u32, u32), reference: (u32, u32) },
}

/// `Mv` (src/motion_vector.rs:64-70), quarter-pel.
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq, Hash)]
pub struct Mv {
    pub row: i16,
    pub col: i16,
}

/// `ReferencePlanes` (src/motion_comp.rs:131-146).
#[derive(Debug, Clone, Copy)]
pub struct ReferencePlanes<'a> {
    pub y: &'a [u8],
    pub u: &'a [u8],
    pub v: &'a [u8],
    pub y_stride: usize,
    pub uv_stride: usize,
    pub mb_cols: usize,
    pub mb_rows: usize,
}

/// `LumaRef` (src/motion_search.rs:230-239).
#[derive(Debug, Clone, Copy)]
pub struct LumaRef<'a> {
    pub plane: &'a [u8],
    pub stride: usize,
    pub width: usize,
    pub height: usize,
}

/// `SearchResult` (src/motion_search.rs:248-255).
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub struct SearchResult {
    pub mv: Mv,
    pub sad: u32,
}

/// `ReconstructedMb` (src/reconstruct.rs:176-195).
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub struct ReconstructedMb {
    pub y: [u8; 256],
    pub u: [u8; 64],
    pub v: [u8; 64],
}

impl Default for ReconstructedMb {
    fn default() -> Self {
        ReconstructedMb { y: [0; 256], u: [0; 64], v: [0; 64] }
    }
}

/// A status from the library that the reference has no counterpart for:
/// an argument beyond the library's limits (`VP8G_E_ARG`), a HIP runtime
/// error (`VP8G_E_DEVICE`) or no usable gfx950 device (`VP8G_E_NO_DEVICE`).
#[derive(Debug, Clone, PartialEq, Eq)]
pub struct DeviceError {
    pub status: i32,
    pub what: &'static str,
}

impl fmt::Display for DeviceError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        // SAFETY: vp8g_status_str returns a static NUL-terminated string for any status.
        let s = unsafe { CStr::from_ptr(sys::vp8g_status_str(self.status)) };
        write!(f, "vp8g: {} failed: {}", self.what, s.to_string_lossy())
    }
}

impl std::error::Error for DeviceError {}

fn check(status: i32, what: &'static str) -> Result<(), DeviceError> {
    if status == sys::VP8G_OK {
        Ok(())
    } else {
        Err(DeviceError { status, what })
    }
}

// ---------------------------------------------------------------- context

/// One library context: a device, a stream and the context's arenas.
/// Not internally locked, so it is `Send` but not `Sync`.
pub struct Context {
    raw: NonNull<sys::vp8g_ctx>,
}

// SAFETY: a context may move between threads; it is never shared (no Sync).
unsafe impl Send for Context {}

impl Context {
    pub fn new(device: i32) -> Result<Context, DeviceError> {
        let mut raw = ptr::null_mut();
        // SAFETY: `raw` is a valid out-pointer.
        check(unsafe { sys::vp8g_ctx_create(device, &mut raw) }, "vp8g_ctx_create")?;
        Ok(Context { raw: NonNull::new(raw).expect("vp8g_ctx_create returned OK and a null context") })
    }

    pub fn as_ptr(&self) -> *mut sys::vp8g_ctx {
        self.raw.as_ptr()
    }

    /// Waits for the context's stream.
    pub fn synchronize(&self) -> Result<(), DeviceError> {
        // SAFETY: the context is live.
        check(unsafe { sys::vp8g_ctx_synchronize(self.as_ptr()) }, "vp8g_ctx_synchronize")
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        // SAFETY: created by vp8g_ctx_create, destroyed once.
        unsafe { sys::vp8g_ctx_destroy(self.raw.as_ptr()) }
    }
}

thread_local! {
    static CTX: RefCell<Option<Context>> = const { RefCell::new(None) };
}

/// Runs `f` with this thread's context, creating it on first use.
pub fn with_context<R>(f: impl FnOnce(&Context) -> Result<R, DeviceError>) -> Result<R, DeviceError> {
    CTX.with(|cell| {
        let mut slot = cell.borrow_mut();
        if slot.is_none() {
            let device = std::env::var("VP8G_DEVICE").ok().and_then(|s| s.parse().ok()).unwrap_or(0);
            *slot = Some(Context::new(device)?);
        }
        f(slot.as_ref().unwrap())
    })
}

fn or_panic<T>(r: Result<T, DeviceError>) -> T {
    r.unwrap_or_else(|e| panic!("{e}"))
}

// ---------------------------------------------------------------- bounds

/// Bytes a plane of `rows` rows of `cols` bytes at `stride` spans.
fn plane_span(stride: usize, cols: usize, rows: usize) -> usize {
    if rows == 0 || cols == 0 {
        0
    } else {
        stride * (rows - 1) + cols
    }
}

fn check_frame(f: &I420Frame<'_>) {
    let (w, h) = (f.width as usize, f.height as usize);
    let (cw, ch) = (w.div_ceil(2), h.div_ceil(2));
    assert!(f.y_stride >= w && f.uv_stride >= cw, "I420Frame strides below the plane widths");
    assert!(f.y.len() >= plane_span(f.y_stride, w, h), "I420Frame::y shorter than y_stride * height");
    assert!(f.u.len() >= plane_span(f.uv_stride, cw, ch), "I420Frame::u shorter than uv_stride * uv_height");
    assert!(f.v.len() >= plane_span(f.uv_stride, cw, ch), "I420Frame::v shorter than uv_stride * uv_height");
}

fn desc(f: &I420Frame<'_>) -> sys::vp8g_i420 {
    sys::vp8g_i420 {
        width: f.width,
        height: f.height,
        y: f.y.as_ptr(),
        u: f.u.as_ptr(),
        v: f.v.as_ptr(),
        y_stride: f.y_stride,
        uv_stride: f.uv_stride,
    }
}

fn ref_planes(r: &ReferencePlanes<'_>) -> sys::vp8g_ref_planes {
    let (lw, lh, cw, ch) = (16 * r.mb_cols, 16 * r.mb_rows, 8 * r.mb_cols, 8 * r.mb_rows);
    assert!(r.y_stride >= lw && r.uv_stride >= cw, "ReferencePlanes strides below the plane widths");
    assert!(r.y.len() >= plane_span(r.y_stride, lw, lh), "ReferencePlanes::y shorter than its planes");
    assert!(r.u.len() >= plane_span(r.uv_stride, cw, ch), "ReferencePlanes::u shorter than its planes");
    assert!(r.v.len() >= plane_span(r.uv_stride, cw, ch), "ReferencePlanes::v shorter than its planes");
    sys::vp8g_ref_planes {
        y: r.y.as_ptr(),
        u: r.u.as_ptr(),
        v: r.v.as_ptr(),
        y_stride: r.y_stride,
        uv_stride: r.uv_stride,
        mb_cols: r.mb_cols,
        mb_rows: r.mb_rows,
    }
}

fn luma_ref(r: &LumaRef<'_>) -> sys::vp8g_luma_ref {
    assert!(r.stride >= r.width, "LumaRef stride below its width");
    assert!(r.plane.len() >= plane_span(r.stride, r.width, r.height), "LumaRef::plane shorter than stride * height");
    sys::vp8g_luma_ref { plane: r.plane.as_ptr(), stride: r.stride, width: r.width, height: r.height }
}

fn mv_c(m: Mv) -> sys::vp8g_mv {
    sys::vp8g_mv { row: m.row, col: m.col }
}

fn mb_xy(mbs: &[(usize, usize)]) -> Vec<u32> {
    mbs.iter()
        .flat_map(|&(c, r)| {
            // the library takes coordinates up to 2^20 per axis
            [u32::try_from(c).unwrap_or(u32::MAX), u32::try_from(r).unwrap_or(u32::MAX)]
        })
        .collect()
}

fn unpack_mbs(out: &[u8]) -> Vec<ReconstructedMb> {
    out.chunks_exact(384)
        .map(|c| {
            let mut m = ReconstructedMb::default();
            m.y.copy_from_slice(&c[..256]);
            m.u.copy_from_slice(&c[256..320]);
            m.v.copy_from_slice(&c[320..]);
            m
        })
        .collect()
}

// ---------------------------------------------------------------- ARNR

/// `build_arnr_altref` (src/arnr.rs:476-482).  Groups of up to
/// `VP8G_MAX_FRAMES` (64) frames; longer ones panic.
pub fn build_arnr_altref(frames: &[I420Frame<'_>], center: usize, cfg: &ArnrConfig)
    -> Result<ArnrPicture, EncodeError> {
    build_arnr_altref_threaded(frames, center, cfg, 1)
}

/// `build_arnr_altref_threaded` (src/arnr.rs:488-650).  The thread budget
/// has no meaning on the GPU; the picture never depended on it.
pub fn build_arnr_altref_threaded(frames: &[I420Frame<'_>], center: usize, cfg: &ArnrConfig, _threads: usize)
    -> Result<ArnrPicture, EncodeError> {
    assert!(center < frames.len(), "build_arnr_altref: center index {center} out of range ({} frames)",
            frames.len());
    let ctr = &frames[center];
    // The library validates before it reads a pixel, so only frames that
    // pass its checks (the centre's dims) are ever read: check their slices.
    for f in frames.iter().filter(|f| f.width == ctr.width && f.height == ctr.height) {
        if (1..=0x3FFF).contains(&f.width) && (1..=0x3FFF).contains(&f.height) {
            check_frame(f);
        }
    }
    let descs: Vec<sys::vp8g_i420> = frames.iter().map(desc).collect();
    let (w, h) = (ctr.width as usize, ctr.height as usize);
    let (cw, ch) = (w.div_ceil(2), h.div_ceil(2));
    let valid = (1..=0x3FFF).contains(&ctr.width) && (1..=0x3FFF).contains(&ctr.height);
    let (yl, cl) = if valid { (w * h, cw * ch) } else { (0, 0) };
    let mut out = ArnrPicture { width: ctr.width, height: ctr.height, y: vec![0; yl], u: vec![0; cl], v: vec![0; cl] };
    let mut err = sys::vp8g_err::default();
    let st = or_panic(with_context(|c| {
        // SAFETY: descriptors point into slices checked above; outputs hold w*h and cw*ch bytes.
        Ok(unsafe {
            sys::vp8g_build_arnr_altref(c.as_ptr(), descs.as_ptr(), descs.len(), center, cfg.strength,
                                        out.y.as_mut_ptr(), out.u.as_mut_ptr(), out.v.as_mut_ptr(), &mut err)
        })
    }));
    match st {
        sys::VP8G_OK => Ok(out),
        sys::VP8G_E_INVALID_DIMENSIONS => Err(EncodeError::InvalidDimensions { width: err.src_w, height: err.src_h }),
        sys::VP8G_E_REF_DIMS_MISMATCH => Err(EncodeError::ReferenceDimensionsMismatch {
            source: (err.src_w, err.src_h),
            reference: (err.ref_w, err.ref_h),
        }),
        s => panic!("{}", DeviceError { status: s, what: "vp8g_build_arnr_altref" }),
    }
}

// ---------------------------------------------------------------- motion compensation

/// `predict_inter_mb` over many macroblocks of one reference:
/// macroblock i at `mbs[i] = (mb_col, mb_row)` with `mvs[i]`.
pub fn predict_inter_mb_batch(reference: &ReferencePlanes<'_>, mbs: &[(usize, usize)], mvs: &[Mv], full_pixel: bool,
                              filters: &[[i32; 6]; 8]) -> Result<Vec<ReconstructedMb>, DeviceError> {
    assert_eq!(mbs.len(), mvs.len(), "one vector per macroblock");
    let rp = ref_planes(reference);
    let xy = mb_xy(mbs);
    let mv: Vec<sys::vp8g_mv> = mvs.iter().copied().map(mv_c).collect();
    let mut out = vec![0u8; 384 * mbs.len()];
    with_context(|c| {
        // SAFETY: planes checked by ref_planes; arrays hold n entries; out holds n*384 bytes.
        check(unsafe {
            sys::vp8g_predict_inter_mb_batch(c.as_ptr(), &rp, xy.as_ptr(), mv.as_ptr(), mbs.len(), full_pixel as i32,
                                             filters.as_ptr(), out.as_mut_ptr())
        }, "vp8g_predict_inter_mb_batch")
    })?;
    Ok(unpack_mbs(&out))
}

/// `predict_inter_mb` (src/motion_comp.rs:1986-2078).
pub fn predict_inter_mb(reference: &ReferencePlanes<'_>, mb_col: usize, mb_row: usize, luma_mv: Mv, full_pixel: bool,
                        filters: &[[i32; 6]; 8]) -> ReconstructedMb {
    or_panic(predict_inter_mb_batch(reference, &[(mb_col, mb_row)], &[luma_mv], full_pixel, filters))[0]
}

/// `predict_split_mv` over many macroblocks: 16 raster-order luma
/// sub-block vectors per macroblock.
pub fn predict_split_mv_batch(reference: &ReferencePlanes<'_>, mbs: &[(usize, usize)], split_luma_mvs: &[[Mv; 16]],
                              full_pixel: bool, filters: &[[i32; 6]; 8]) -> Result<Vec<ReconstructedMb>, DeviceError> {
    assert_eq!(mbs.len(), split_luma_mvs.len(), "16 vectors per macroblock");
    let rp = ref_planes(reference);
    let xy = mb_xy(mbs);
    let mv: Vec<sys::vp8g_mv> = split_luma_mvs.iter().flatten().copied().map(mv_c).collect();
    let mut out = vec![0u8; 384 * mbs.len()];
    with_context(|c| {
        // SAFETY: as predict_inter_mb_batch, 16 vectors per macroblock.
        check(unsafe {
            sys::vp8g_predict_split_mv_batch(c.as_ptr(), &rp, xy.as_ptr(), mv.as_ptr(), mbs.len(), full_pixel as i32,
                                             filters.as_ptr(), out.as_mut_ptr())
        }, "vp8g_predict_split_mv_batch")
    })?;
    Ok(unpack_mbs(&out))
}

/// `predict_split_mv` (src/motion_comp.rs:2227-2320).
pub fn predict_split_mv(reference: &ReferencePlanes<'_>, mb_col: usize, mb_row: usize, split_luma_mvs: &[Mv; 16],
                        full_pixel: bool, filters: &[[i32; 6]; 8]) -> ReconstructedMb {
    or_panic(predict_split_mv_batch(reference, &[(mb_col, mb_row)], std::slice::from_ref(split_luma_mvs), full_pixel,
                                    filters))[0]
}

/// The dequantised coefficients of one macroblock: the Y2 block (unused by
/// SPLITMV), 16 Y blocks, 4 U and 4 V blocks.
#[derive(Debug, Clone, Copy)]
pub struct MbCoeffs<'a> {
    pub y2: &'a [i16; 16],
    pub y: &'a [[i16; 16]; 16],
    pub u: &'a [[i16; 16]; 4],
    pub v: &'a [[i16; 16]; 4],
}

fn pack_coeffs(c: &MbCoeffs<'_>, with_y2: bool, dst: &mut Vec<i16>) {
    if with_y2 {
        dst.extend_from_slice(c.y2);
    }
    for b in c.y.iter().chain(c.u.iter()).chain(c.v.iter()) {
        dst.extend_from_slice(b);
    }
}

/// `reconstruct_inter_mb` over many macroblocks; `skip[i]` is
/// `mb_skip_coeff`.
pub fn reconstruct_inter_mb_batch(reference: &ReferencePlanes<'_>, mbs: &[(usize, usize)], mvs: &[Mv],
                                  full_pixel: bool, filters: &[[i32; 6]; 8], skip: &[bool], coeffs: &[MbCoeffs<'_>])
    -> Result<Vec<ReconstructedMb>, DeviceError> {
    let n = mbs.len();
    assert!(mvs.len() == n && skip.len() == n && coeffs.len() == n, "one entry per macroblock");
    let rp = ref_planes(reference);
    let xy = mb_xy(mbs);
    let mv: Vec<sys::vp8g_mv> = mvs.iter().copied().map(mv_c).collect();
    let sk: Vec<u8> = skip.iter().map(|&s| s as u8).collect();
    let mut cf = Vec::with_capacity(400 * n);
    coeffs.iter().for_each(|c| pack_coeffs(c, true, &mut cf));
    let mut out = vec![0u8; 384 * n];
    with_context(|c| {
        // SAFETY: 400 coefficients and one skip flag per macroblock; out holds n*384 bytes.
        check(unsafe {
            sys::vp8g_reconstruct_inter_mb_batch(c.as_ptr(), &rp, xy.as_ptr(), mv.as_ptr(), n, full_pixel as i32,
                                                 filters.as_ptr(), sk.as_ptr(), cf.as_ptr(), out.as_mut_ptr())
        }, "vp8g_reconstruct_inter_mb_batch")
    })?;
    Ok(unpack_mbs(&out))
}

/// `reconstruct_inter_mb` (src/motion_comp.rs:2098-2150).
#[allow(clippy::too_many_arguments)]
pub fn reconstruct_inter_mb(reference: &ReferencePlanes<'_>, mb_col: usize, mb_row: usize, luma_mv: Mv,
                            full_pixel: bool, filters: &[[i32; 6]; 8], mb_skip_coeff: bool,
                            y2_coeffs_dequant: &[i16; 16], y_coeffs_dequant: &[[i16; 16]; 16],
                            u_coeffs_dequant: &[[i16; 16]; 4], v_coeffs_dequant: &[[i16; 16]; 4]) -> ReconstructedMb {
    let c = MbCoeffs { y2: y2_coeffs_dequant, y: y_coeffs_dequant, u: u_coeffs_dequant, v: v_coeffs_dequant };
    or_panic(reconstruct_inter_mb_batch(reference, &[(mb_col, mb_row)], &[luma_mv], full_pixel, filters,
                                        &[mb_skip_coeff], &[c]))[0]
}

/// `reconstruct_split_mv_mb` over many macroblocks (the Y2 block of each
/// `MbCoeffs` is not read).
pub fn reconstruct_split_mv_mb_batch(reference: &ReferencePlanes<'_>, mbs: &[(usize, usize)],
                                     split_luma_mvs: &[[Mv; 16]], full_pixel: bool, filters: &[[i32; 6]; 8],
                                     skip: &[bool], coeffs: &[MbCoeffs<'_>])
    -> Result<Vec<ReconstructedMb>, DeviceError> {
    let n = mbs.len();
    assert!(split_luma_mvs.len() == n && skip.len() == n && coeffs.len() == n, "one entry per macroblock");
    let rp = ref_planes(reference);
    let xy = mb_xy(mbs);
    let mv: Vec<sys::vp8g_mv> = split_luma_mvs.iter().flatten().copied().map(mv_c).collect();
    let sk: Vec<u8> = skip.iter().map(|&s| s as u8).collect();
    let mut cf = Vec::with_capacity(384 * n);
    coeffs.iter().for_each(|c| pack_coeffs(c, false, &mut cf));
    let mut out = vec![0u8; 384 * n];
    with_context(|c| {
        // SAFETY: 16 vectors, 384 coefficients and one skip flag per macroblock.
        check(unsafe {
            sys::vp8g_reconstruct_split_mv_mb_batch(c.as_ptr(), &rp, xy.as_ptr(), mv.as_ptr(), n, full_pixel as i32,
                                                    filters.as_ptr(), sk.as_ptr(), cf.as_ptr(), out.as_mut_ptr())
        }, "vp8g_reconstruct_split_mv_mb_batch")
    })?;
    Ok(unpack_mbs(&out))
}

/// `reconstruct_split_mv_mb` (src/motion_comp.rs:2340-2385).
#[allow(clippy::too_many_arguments)]
pub fn reconstruct_split_mv_mb(reference: &ReferencePlanes<'_>, mb_col: usize, mb_row: usize,
                               split_luma_mvs: &[Mv; 16], full_pixel: bool, filters: &[[i32; 6]; 8],
                               mb_skip_coeff: bool, y_coeffs_dequant: &[[i16; 16]; 16],
                               u_coeffs_dequant: &[[i16; 16]; 4], v_coeffs_dequant: &[[i16; 16]; 4])
    -> ReconstructedMb {
    let c = MbCoeffs { y2: &[0; 16], y: y_coeffs_dequant, u: u_coeffs_dequant, v: v_coeffs_dequant };
    or_panic(reconstruct_split_mv_mb_batch(reference, &[(mb_col, mb_row)], std::slice::from_ref(split_luma_mvs),
                                           full_pixel, filters, &[mb_skip_coeff], &[c]))[0]
}

// ---------------------------------------------------------------- sub-pel SAD scoring and searches

/// `mb_luma_sad_at_mv` over many macroblocks: `sad[i]` of `src_y[i]`
/// against macroblock `mbs[i]` at `mvs[i]`.
pub fn mb_luma_sad_at_mv_batch(reference: LumaRef<'_>, mbs: &[(usize, usize)], src_y: &[[u8; 256]], mvs: &[Mv])
    -> Result<Vec<u32>, DeviceError> {
    let n = mbs.len();
    assert!(src_y.len() == n && mvs.len() == n, "one source block and vector per macroblock");
    let lr = luma_ref(&reference);
    let xy = mb_xy(mbs);
    let mv: Vec<sys::vp8g_mv> = mvs.iter().copied().map(mv_c).collect();
    let mut sad = vec![0u32; n];
    with_context(|c| {
        // SAFETY: plane checked by luma_ref; n blocks of 256 bytes; n outputs.
        check(unsafe {
            sys::vp8g_mb_luma_sad_at_mv_batch(c.as_ptr(), &lr, xy.as_ptr(), src_y.as_ptr().cast(), mv.as_ptr(), n,
                                              sad.as_mut_ptr())
        }, "vp8g_mb_luma_sad_at_mv_batch")
    })?;
    Ok(sad)
}

/// `mb_luma_sad_at_mv` (src/motion_search.rs:470-529).
pub fn mb_luma_sad_at_mv(reference: LumaRef<'_>, mb_col: usize, mb_row: usize, src_y: &[u8; 256], mv: Mv) -> u32 {
    or_panic(mb_luma_sad_at_mv_batch(reference, &[(mb_col, mb_row)], std::slice::from_ref(src_y), &[mv]))[0]
}

type RefineFn = unsafe extern "C" fn(*mut sys::vp8g_ctx, *const sys::vp8g_luma_ref, *const u32, *const u8,
                                     *const sys::vp8g_mv, usize, *mut sys::vp8g_search_result) -> i32;

fn refine_batch(f: RefineFn, what: &'static str, reference: LumaRef<'_>, mbs: &[(usize, usize)],
                src_y: &[[u8; 256]], centers: &[Mv]) -> Result<Vec<SearchResult>, DeviceError> {
    let n = mbs.len();
    assert!(src_y.len() == n && centers.len() == n, "one source block and centre per macroblock");
    let lr = luma_ref(&reference);
    let xy = mb_xy(mbs);
    let ct: Vec<sys::vp8g_mv> = centers.iter().copied().map(mv_c).collect();
    let mut out = vec![sys::vp8g_search_result::default(); n];
    with_context(|c| {
        // SAFETY: as mb_luma_sad_at_mv_batch; n results.
        check(unsafe { f(c.as_ptr(), &lr, xy.as_ptr(), src_y.as_ptr().cast(), ct.as_ptr(), n, out.as_mut_ptr()) },
              what)
    })?;
    Ok(out.iter().map(|r| SearchResult { mv: Mv { row: r.mv.row, col: r.mv.col }, sad: r.sad }).collect())
}

/// `half_pixel_refine_luma` over many macroblocks (centres within
/// [-1023, 1023], whole-pel).
pub fn half_pixel_refine_luma_batch(reference: LumaRef<'_>, mbs: &[(usize, usize)], src_y: &[[u8; 256]],
                                    centers: &[Mv]) -> Result<Vec<SearchResult>, DeviceError> {
    refine_batch(sys::vp8g_half_pixel_refine_luma_batch, "vp8g_half_pixel_refine_luma_batch", reference, mbs, src_y,
                 centers)
}

/// `half_pixel_refine_luma` (src/motion_search.rs:564-627).
pub fn half_pixel_refine_luma(reference: LumaRef<'_>, mb_col: usize, mb_row: usize, src_y: &[u8; 256],
                              whole_pixel_center: Mv) -> SearchResult {
    or_panic(half_pixel_refine_luma_batch(reference, &[(mb_col, mb_row)], std::slice::from_ref(src_y),
                                          &[whole_pixel_center]))[0]
}

/// `quarter_pixel_refine_luma` over many macroblocks.
pub fn quarter_pixel_refine_luma_batch(reference: LumaRef<'_>, mbs: &[(usize, usize)], src_y: &[[u8; 256]],
                                       centers: &[Mv]) -> Result<Vec<SearchResult>, DeviceError> {
    refine_batch(sys::vp8g_quarter_pixel_refine_luma_batch, "vp8g_quarter_pixel_refine_luma_batch", reference, mbs,
                 src_y, centers)
}

/// `quarter_pixel_refine_luma` (src/motion_search.rs:666-728).
pub fn quarter_pixel_refine_luma(reference: LumaRef<'_>, mb_col: usize, mb_row: usize, src_y: &[u8; 256],
                                 half_pixel_center: Mv) -> SearchResult {
    or_panic(quarter_pixel_refine_luma_batch(reference, &[(mb_col, mb_row)], std::slice::from_ref(src_y),
                                             &[half_pixel_center]))[0]
}

/// `small_diamond_search_luma` over many macroblocks.
pub fn small_diamond_search_luma_batch(reference: LumaRef<'_>, mbs: &[(usize, usize)], src_y: &[[u8; 256]],
                                       centers: &[Mv], max_iters: u32) -> Result<Vec<SearchResult>, DeviceError> {
    let n = mbs.len();
    assert!(src_y.len() == n && centers.len() == n, "one source block and centre per macroblock");
    let lr = luma_ref(&reference);
    let xy = mb_xy(mbs);
    let ct: Vec<sys::vp8g_mv> = centers.iter().copied().map(mv_c).collect();
    let mut out = vec![sys::vp8g_search_result::default(); n];
    with_context(|c| {
        // SAFETY: as mb_luma_sad_at_mv_batch; n results.
        check(unsafe {
            sys::vp8g_small_diamond_search_luma_batch(c.as_ptr(), &lr, xy.as_ptr(), src_y.as_ptr().cast(),
                                                      ct.as_ptr(), n, max_iters, out.as_mut_ptr())
        }, "vp8g_small_diamond_search_luma_batch")
    })?;
    Ok(out.iter().map(|r| SearchResult { mv: Mv { row: r.mv.row, col: r.mv.col }, sad: r.sad }).collect())
}

/// `small_diamond_search_luma` (src/motion_search.rs:381-442).
pub fn small_diamond_search_luma(reference: LumaRef<'_>, mb_col: usize, mb_row: usize, src_y: &[u8; 256],
                                 center: Mv, max_iters: u32) -> SearchResult {
    or_panic(small_diamond_search_luma_batch(reference, &[(mb_col, mb_row)], std::slice::from_ref(src_y), &[center],
                                             max_iters))[0]
}

// ---------------------------------------------------------------- stream driver

/// `luma_mad` (src/stream.rs:2654-2666) of each (previous, current) pair;
/// every frame has the dims of the first previous frame.
pub fn luma_mad_batch(prev: &[I420Frame<'_>], cur: &[I420Frame<'_>]) -> Result<Vec<f64>, DeviceError> {
    assert_eq!(prev.len(), cur.len(), "one previous frame per current frame");
    for f in prev.iter().chain(cur) {
        let (w, h) = (f.width as usize, f.height as usize);
        assert!(f.y_stride >= w && f.y.len() >= plane_span(f.y_stride, w, h), "luma plane shorter than its frame");
    }
    let p: Vec<sys::vp8g_i420> = prev.iter().map(desc).collect();
    let c: Vec<sys::vp8g_i420> = cur.iter().map(desc).collect();
    let mut mad = vec![0f64; prev.len()];
    with_context(|ctx| {
        // SAFETY: luma planes checked above (only luma is read); n outputs.
        check(unsafe { sys::vp8g_luma_mad_batch(ctx.as_ptr(), p.as_ptr(), c.as_ptr(), p.len(), mad.as_mut_ptr()) },
              "vp8g_luma_mad_batch")
    })?;
    Ok(mad)
}

// ---------------------------------------------------------------- loop filter

/// `FrameFilterConfig` (src/loop_filter.rs:891-944).
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq)]
pub struct FrameFilterConfig {
    pub simple: bool,
    pub key_frame: bool,
    pub loop_filter_level: u8,
    pub sharpness_level: u8,
    pub segmentation_enabled: bool,
    pub segment_abs: bool,
    pub segment_lf_level: [i16; 4],
    pub delta_enabled: bool,
    pub ref_delta_current: i16,
    pub bpred_mode_delta: i16,
    pub ref_delta_last: i16,
    pub ref_delta_golden: i16,
    pub ref_delta_altref: i16,
    pub zero_mv_mode_delta: i16,
    pub other_mv_mode_delta: i16,
    pub split_mv_mode_delta: i16,
}

impl FrameFilterConfig {
    fn c(&self) -> sys::vp8g_lf_config {
        sys::vp8g_lf_config {
            simple: self.simple as i32,
            key_frame: self.key_frame as i32,
            loop_filter_level: self.loop_filter_level as i32,
            sharpness_level: self.sharpness_level as i32,
            segmentation_enabled: self.segmentation_enabled as i32,
            segment_abs: self.segment_abs as i32,
            delta_enabled: self.delta_enabled as i32,
            segment_lf_level: self.segment_lf_level,
            // FrameFilterConfig::ref_deltas / mode_deltas (src/loop_filter.rs:1193-1209)
            ref_delta: [self.ref_delta_current, self.ref_delta_last, self.ref_delta_golden, self.ref_delta_altref],
            mode_delta: [self.bpred_mode_delta, self.zero_mv_mode_delta, self.other_mv_mode_delta,
                         self.split_mv_mode_delta],
        }
    }
}

/// `KeyframePlanes` (src/frame.rs): the MB-padded reconstruction.
#[derive(Debug, Clone, PartialEq, Eq)]
pub struct KeyframePlanes {
    pub y: Vec<u8>,
    pub u: Vec<u8>,
    pub v: Vec<u8>,
    pub y_stride: usize,
    pub uv_stride: usize,
    pub mb_cols: usize,
    pub mb_rows: usize,
}

/// `SourcePlanes` (src/loop_filter.rs:1916-1932).
#[derive(Debug, Clone, Copy)]
pub struct SourcePlanes<'a> {
    pub width: usize,
    pub height: usize,
    pub y: &'a [u8],
    pub u: &'a [u8],
    pub v: &'a [u8],
    pub y_stride: usize,
    pub uv_stride: usize,
}

/// `motion_comp::RefFrame` of an inter macroblock.
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum RefFrame {
    Last,
    Golden,
    AltRef,
}

/// `near_mv::InterMode` of an inter macroblock.
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum InterMode {
    Nearest,
    Near,
    Zero,
    New,
    Split,
}

/// What the section 15 pass reads of one macroblock: `MacroblockModes::segment_id`,
/// `y_mode == B_PRED` (SPLITMV macroblocks are recorded as B_PRED,
/// src/loop_filter.rs:1301-1308), `mb_has_coeffs`, and for an inter
/// macroblock its reference frame and inter mode.
#[derive(Debug, Clone, Copy, Default, PartialEq, Eq)]
pub struct MbFilterInfo {
    pub segment_id: Option<u8>,
    pub bpred: bool,
    pub has_coeffs: bool,
    pub inter: Option<(RefFrame, InterMode)>,
}

fn lf_records(mbs: &[MbFilterInfo]) -> Vec<sys::vp8g_lf_mb> {
    mbs.iter()
        .map(|m| {
            let (ref_frame, inter_mode) = match m.inter {
                None => (0u8, 0u8),
                Some((r, i)) => (
                    match r {
                        RefFrame::Last => 1,
                        RefFrame::Golden => 2,
                        RefFrame::AltRef => 3,
                    },
                    match i {
                        InterMode::Zero => sys::VP8G_LF_ZEROMV as u8,
                        InterMode::Split => sys::VP8G_LF_SPLITMV as u8,
                        _ => 3,
                    },
                ),
            };
            sys::vp8g_lf_mb {
                segment_id: m.segment_id.unwrap_or(sys::VP8G_LF_NO_SEGMENT as u8),
                ref_frame,
                inter_mode,
                flags: (if m.bpred { sys::VP8G_LF_BPRED as u8 } else { 0 })
                    | (if m.has_coeffs { sys::VP8G_LF_HAS_COEFFS as u8 } else { 0 }),
            }
        })
        .collect()
}

fn check_lf_planes(p: &KeyframePlanes) {
    let (c, r) = (p.mb_cols, p.mb_rows);
    assert!(p.y_stride >= c * 16 && p.uv_stride >= c * 8, "plane stride shorter than the MB columns");
    assert!(p.y.len() >= plane_span(p.y_stride, c * 16, r * 16), "luma plane shorter than its MB rows");
    assert!(p.u.len() >= plane_span(p.uv_stride, c * 8, r * 8) && p.v.len() >= plane_span(p.uv_stride, c * 8, r * 8),
            "chroma plane shorter than its MB rows");
}

fn lf_planes(p: &mut KeyframePlanes) -> sys::vp8g_lf_planes {
    sys::vp8g_lf_planes {
        y: p.y.as_mut_ptr(),
        u: p.u.as_mut_ptr(),
        v: p.v.as_mut_ptr(),
        y_stride: p.y_stride,
        uv_stride: p.uv_stride,
        mb_cols: p.mb_cols,
        mb_rows: p.mb_rows,
    }
}

fn lf_source(s: &SourcePlanes<'_>) -> sys::vp8g_lf_source {
    let (cw, ch) = (s.width.div_ceil(2), s.height.div_ceil(2));
    assert!(s.y.len() >= plane_span(s.y_stride, s.width, s.height), "source luma shorter than its window");
    assert!(s.u.len() >= plane_span(s.uv_stride, cw, ch) && s.v.len() >= plane_span(s.uv_stride, cw, ch),
            "source chroma shorter than its window");
    sys::vp8g_lf_source {
        width: s.width,
        height: s.height,
        y: s.y.as_ptr(),
        u: s.u.as_ptr(),
        v: s.v.as_ptr(),
        y_stride: s.y_stride,
        uv_stride: s.uv_stride,
    }
}

/// `filter_inter_frame_flags` (src/loop_filter.rs:1364-1496): the frame
/// filtered in place, byte for byte the raster pass.  Like the reference, a
/// record count that is not mb_cols * mb_rows leaves the planes alone.
pub fn filter_inter_frame_flags(planes: &mut KeyframePlanes, mbs: &[MbFilterInfo], config: &FrameFilterConfig)
                                -> Result<(), DeviceError> {
    if planes.mb_cols == 0 || planes.mb_rows == 0 || mbs.len() != planes.mb_cols * planes.mb_rows {
        return Ok(());
    }
    check_lf_planes(planes);
    let rec = lf_records(mbs);
    let cfg = config.c();
    let p = lf_planes(planes);
    with_context(|ctx| {
        // SAFETY: planes checked above; one record per MB.
        check(unsafe { sys::vp8g_loop_filter_frame(ctx.as_ptr(), &p, rec.as_ptr(), &cfg) }, "vp8g_loop_filter_frame")
    })
}

/// `filter_frame_flags` (src/loop_filter.rs:2506-2636): a key frame, every
/// macroblock intra.
pub fn filter_frame_flags(planes: &mut KeyframePlanes, segment_ids: &[Option<u8>], bpred: &[bool],
                          has_coeffs: &[bool], config: &FrameFilterConfig) -> Result<(), DeviceError> {
    if bpred.len() != has_coeffs.len() || segment_ids.len() != bpred.len() {
        return Ok(());
    }
    let mbs: Vec<MbFilterInfo> = (0..bpred.len())
        .map(|i| MbFilterInfo { segment_id: segment_ids[i], bpred: bpred[i], has_coeffs: has_coeffs[i], inter: None })
        .collect();
    filter_inter_frame_flags(planes, &mbs, config)
}

fn select(planes: &KeyframePlanes, mbs: &[MbFilterInfo], base: &FrameFilterConfig, src: &SourcePlanes<'_>,
          every_rows: usize, sweep: bool) -> Result<(u8, u8), DeviceError> {
    assert_eq!(mbs.len(), planes.mb_cols * planes.mb_rows, "one record per macroblock");
    check_lf_planes(planes);
    let rec = lf_records(mbs);
    let cfg = base.c();
    let s = lf_source(src);
    // the call only reads the planes
    let p = sys::vp8g_lf_planes {
        y: planes.y.as_ptr() as *mut u8,
        u: planes.u.as_ptr() as *mut u8,
        v: planes.v.as_ptr() as *mut u8,
        y_stride: planes.y_stride,
        uv_stride: planes.uv_stride,
        mb_cols: planes.mb_cols,
        mb_rows: planes.mb_rows,
    };
    let (mut level, mut sharp) = (0u8, 0u8);
    with_context(|ctx| {
        // SAFETY: planes and source checked above; one record per MB.
        check(unsafe {
            sys::vp8g_loop_filter_select(ctx.as_ptr(), &p, rec.as_ptr(), &cfg, &s, every_rows, sweep as i32,
                                         &mut level, &mut sharp)
        }, "vp8g_loop_filter_select")
    })?;
    Ok((level, sharp))
}

/// `select_filter_params_threaded` (src/loop_filter.rs:2133-2200): the
/// (level, sharpness) of the reference's search; the candidates are
/// scored on the device, whatever `threads` says.
pub fn select_filter_params_threaded(planes: &KeyframePlanes, mbs: &[MbFilterInfo], base: &FrameFilterConfig,
                                     src: &SourcePlanes<'_>, _threads: usize) -> (u8, u8) {
    or_panic(select(planes, mbs, base, src, 0, true))
}

/// `select_filter_params_sampled` (src/loop_filter.rs:2288-2412) with
/// `FilterSampling { every_rows, sweep_sharpness }`.
pub fn select_filter_params_sampled(planes: &KeyframePlanes, mbs: &[MbFilterInfo], base: &FrameFilterConfig,
                                    src: &SourcePlanes<'_>, every_rows: usize, sweep_sharpness: bool,
                                    _threads: usize) -> (u8, u8) {
    or_panic(select(planes, mbs, base, src, every_rows.max(2), sweep_sharpness))
}
