//! Builds libvp8g.so for gfx950 from the HIP sources in `../../csrc` with
//! hipcc, into OUT_DIR, and links it (with an rpath to OUT_DIR) plus the
//! ROCm runtime.
//!
//! Every `.hip` / `.cpp` file of the source directory is one translation
//! unit, compiled in parallel with the flags below, then linked with
//! `hipcc -shared`.  Environment:
//!   HIPCC        compiler (default $ROCM_PATH/bin/hipcc)
//!   ROCM_PATH    ROCm root (default /opt/rocm)
//!   VP8G_LIB_DIR directory holding a prebuilt libvp8g.so; with the
//!                `prebuilt` feature (or this variable set) no hipcc runs.
use std::env;
use std::path::{Path, PathBuf};
use std::process::Command;
use std::thread;

/// hipcc flags of every translation unit.  Packed-FP32 instruction
/// selection stays off.
const HIP_FLAGS: &[&str] = &[
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
];

fn link_runtime(rocm: &str) {
    println!("cargo:rustc-link-search=native={rocm}/lib");
    println!("cargo:rustc-link-lib=dylib=amdhip64");
}

fn main() {
    let rocm = env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string());
    println!("cargo:rerun-if-env-changed=VP8G_LIB_DIR");
    println!("cargo:rerun-if-env-changed=HIPCC");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");

    if let Ok(dir) = env::var("VP8G_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
        println!("cargo:rustc-link-lib=dylib=vp8g");
        link_runtime(&rocm);
        return;
    }
    if env::var_os("CARGO_FEATURE_PREBUILT").is_some() {
        panic!("feature `prebuilt` needs VP8G_LIB_DIR, the directory holding libvp8g.so");
    }

    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let csrc = manifest.join("../../csrc").canonicalize().expect("HIP sources at ../../csrc");
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let hipcc = env::var("HIPCC").unwrap_or_else(|_| format!("{rocm}/bin/hipcc"));
    println!("cargo:rerun-if-changed={}", csrc.display());

    let mut sources: Vec<PathBuf> = std::fs::read_dir(&csrc)
        .expect("read csrc")
        .map(|e| e.unwrap().path())
        .filter(|p| matches!(p.extension().and_then(|x| x.to_str()), Some("hip") | Some("cpp")))
        .collect();
    sources.sort();
    for entry in std::fs::read_dir(&csrc).unwrap() {
        println!("cargo:rerun-if-changed={}", entry.unwrap().path().display());
    }

    let compile = |src: &Path| -> PathBuf {
        let obj = out.join(src.file_stem().unwrap()).with_extension("o");
        let mut cmd = Command::new(&hipcc);
        cmd.args(HIP_FLAGS);
        if src.extension().and_then(|x| x.to_str()) == Some("cpp") {
            cmd.args(["-x", "hip"]);
        }
        cmd.arg("-c").arg(src).arg("-o").arg(&obj);
        let st = cmd.status().unwrap_or_else(|e| panic!("running {hipcc}: {e}"));
        assert!(st.success(), "hipcc failed on {}", src.display());
        obj
    };
    let objs: Vec<PathBuf> = thread::scope(|s| {
        let handles: Vec<_> = sources.iter().map(|src| s.spawn(move || compile(src))).collect();
        handles.into_iter().map(|h| h.join().expect("compile thread")).collect()
    });

    let lib = out.join("libvp8g.so");
    let st = Command::new(&hipcc)
        .args(["--offload-arch=gfx950", "-fPIC", "-shared", "-o"])
        .arg(&lib)
        .args(&objs)
        .arg("-ldl")
        .status()
        .expect("running hipcc to link");
    assert!(st.success(), "linking libvp8g.so failed");

    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
    println!("cargo:rustc-link-lib=dylib=vp8g");
    link_runtime(&rocm);
}
