//! Properties of the safe wrappers that fix their outputs without a second
//! implementation: the strength-0 altref is the centre frame; a whole-pel
//! prediction inside the frame is a copy of the displaced block; the SAD at
//! a whole-pel vector is the plain SAD of that block; batches equal single
//! calls; the reference's errors come back as its `EncodeError` values.
//! Each test returns early (and says so) on a machine without a gfx950
//! device.
use vp8g::*;

fn have_device() -> bool {
    match Context::new(0) {
        Ok(_) => true,
        Err(e) => {
            eprintln!("skipping: {e}");
            false
        }
    }
}

/// The six-tap table of RFC 6386 section 18.3 (the reference's
/// SIXTAP_FILTERS, src/motion_comp.rs:294-337).
const SIXTAP: [[i32; 6]; 8] = [
    [0, 0, 128, 0, 0, 0],
    [0, -6, 123, 12, -1, 0],
    [2, -11, 108, 36, -8, 1],
    [0, -9, 93, 50, -6, 0],
    [3, -16, 77, 77, -16, 3],
    [0, -6, 50, 93, -9, 0],
    [1, -8, 36, 108, -11, 2],
    [0, -1, 12, 123, -6, 0],
];

fn noise(n: usize, seed: u64) -> Vec<u8> {
    let mut s = seed.wrapping_mul(0x9E37_79B9_7F4A_7C15) | 1;
    (0..n)
        .map(|_| {
            s ^= s << 13;
            s ^= s >> 7;
            s ^= s << 17;
            (s >> 24) as u8
        })
        .collect()
}

#[test]
fn strength_zero_copies_the_centre_and_errors_match_the_reference() {
    if !have_device() {
        return;
    }
    let (w, h) = (35usize, 19usize);
    let (cw, ch) = (w.div_ceil(2), h.div_ceil(2));
    let planes: Vec<(Vec<u8>, Vec<u8>, Vec<u8>)> =
        (0..3).map(|i| (noise(w * h, i), noise(cw * ch, 10 + i), noise(cw * ch, 20 + i))).collect();
    let frames: Vec<I420Frame> = planes
        .iter()
        .map(|(y, u, v)| I420Frame { width: w as u32, height: h as u32, y, u, v, y_stride: w, uv_stride: cw })
        .collect();
    let pic = build_arnr_altref(&frames, 1, &ArnrConfig { strength: 0 }).unwrap();
    assert_eq!((pic.y.as_slice(), pic.u.as_slice(), pic.v.as_slice()),
               (planes[1].0.as_slice(), planes[1].1.as_slice(), planes[1].2.as_slice()));
    // strength 3 is deterministic and the thread budget changes nothing
    let a = build_arnr_altref(&frames, 1, &ArnrConfig::default()).unwrap();
    let b = build_arnr_altref_threaded(&frames, 1, &ArnrConfig::default(), 8).unwrap();
    assert_eq!(a, b);

    let mut odd = frames.clone();
    odd[2].width = 33;
    assert_eq!(build_arnr_altref(&odd, 1, &ArnrConfig::default()),
               Err(EncodeError::ReferenceDimensionsMismatch { source: (35, 19), reference: (33, 19) }));
    let mut zero = frames.clone();
    zero[1].width = 0;
    assert_eq!(build_arnr_altref(&zero, 1, &ArnrConfig::default()),
               Err(EncodeError::InvalidDimensions { width: 0, height: 19 }));
}

#[test]
#[should_panic(expected = "out of range")]
fn centre_out_of_range_panics() {
    let y = [0u8; 4];
    let f = I420Frame { width: 2, height: 2, y: &y, u: &y[..1], v: &y[..1], y_stride: 2, uv_stride: 1 };
    let _ = build_arnr_altref(&[f], 1, &ArnrConfig::default());
}

#[test]
fn whole_pel_prediction_and_sad_are_copies_and_plain_sads() {
    if !have_device() {
        return;
    }
    let (mbc, mbr) = (6usize, 4usize);
    let (lw, lh, cw, ch) = (16 * mbc, 16 * mbr, 8 * mbc, 8 * mbr);
    let (y, u, v) = (noise(lw * lh, 1), noise(cw * ch, 2), noise(cw * ch, 3));
    let rp = ReferencePlanes { y: &y, u: &u, v: &v, y_stride: lw, uv_stride: cw, mb_cols: mbc, mb_rows: mbr };
    // macroblock (2, 1) displaced by (+8, -8) pixels: quarter-pel (row -32, col 32)
    let mv = Mv { row: -32, col: 32 };
    let mb = predict_inter_mb(&rp, 2, 1, mv, false, &SIXTAP);
    for r in 0..16 {
        for c in 0..16 {
            assert_eq!(mb.y[16 * r + c], y[(16 + r - 8) * lw + 32 + c + 8]);
        }
    }
    // chroma moves by half: (+4, -4)
    for r in 0..8 {
        for c in 0..8 {
            assert_eq!(mb.u[8 * r + c], u[(8 + r - 4) * cw + 16 + c + 4]);
        }
    }
    let lr = LumaRef { plane: &y, stride: lw, width: lw, height: lh };
    let src: [u8; 256] = noise(256, 9).try_into().unwrap();
    let want: u32 = (0..256).map(|i| (src[i] as i32 - mb.y[i] as i32).unsigned_abs()).sum();
    assert_eq!(mb_luma_sad_at_mv(lr, 2, 1, &src, mv), want);
    assert_eq!(mb_luma_sad_at_mv(lr, 2, 1, &mb.y, mv), 0);

    // a batch equals its single calls
    let mbs: Vec<(usize, usize)> = (0..mbr).flat_map(|r| (0..mbc).map(move |c| (c, r))).collect();
    let mvs: Vec<Mv> = (0..mbs.len()).map(|i| Mv { row: (i as i16 % 9) - 4, col: 3 - (i as i16 % 7) }).collect();
    let batch = predict_inter_mb_batch(&rp, &mbs, &mvs, false, &SIXTAP).unwrap();
    for (i, &(c, r)) in mbs.iter().enumerate() {
        assert_eq!(batch[i], predict_inter_mb(&rp, c, r, mvs[i], false, &SIXTAP));
    }
    let srcs = vec![src; mbs.len()];
    let sads = mb_luma_sad_at_mv_batch(lr, &mbs, &srcs, &mvs).unwrap();
    for (i, &(c, r)) in mbs.iter().enumerate() {
        assert_eq!(sads[i], mb_luma_sad_at_mv(lr, c, r, &src, mvs[i]));
    }
    // sixteen equal SPLITMV vectors predict as the whole macroblock
    assert_eq!(predict_split_mv(&rp, 2, 1, &[mv; 16], false, &SIXTAP), mb);
    // a skipped macroblock reconstructs to its prediction
    let z16 = [0i16; 16];
    let zy = [[0i16; 16]; 16];
    let zc = [[0i16; 16]; 4];
    assert_eq!(reconstruct_inter_mb(&rp, 2, 1, mv, false, &SIXTAP, true, &z16, &zy, &zc, &zc), mb);
}
