// build.rs for the `hip` feature: compiles the HIP kernels and the C ABI
// with hipcc for gfx950 into $OUT_DIR/libfastatari_hip.so and links it.
// Replaces the nvcc/PTX step of the reference's build.rs:12-36.  Sources and
// options are those of fastatari/build.py (one object per translation unit).
use std::path::Path;
use std::process::Command;

fn main() {
    #[cfg(feature = "hip")]
    build_hip();
}

#[cfg(feature = "hip")]
fn build_hip() {
    let manifest = std::env::var("CARGO_MANIFEST_DIR").unwrap();
    let out_dir = std::env::var("OUT_DIR").unwrap();
    let csrc = Path::new(&manifest).join("fast-atari-rs_amd/csrc");
    let hipcc = std::env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".into());
    let flags = [
        "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC",
        // packed FP32 instructions are disabled for every HIP translation unit
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        // keeps the frame loop's wave-uniform control state in scalar registers
        "-fno-slp-vectorize",
        "-pthread",
    ];
    // (source, extra options): both frame kernels keep branches whose
    // conditions are wave-uniform as scalar branches; the rest is host code
    // with its own small kernels (fa_render.hip: the rendering console).
    let uniform_skip: &[&str] = &["-mllvm", "-structurizecfg-skip-uniform-regions=true"];
    let sources: [(&str, &[&str]); 9] = [
        ("fa_kernels.hip", uniform_skip),
        ("fa_kernels_sorted.hip", uniform_skip),
        ("fa_lib.hip", &[]),
        ("fa_batch.hip", &[]),
        ("fa_serve.hip", &[]),
        ("fa_adapter.hip", &[]),
        ("fa_noalloc.hip", &[]),
        ("fa_probe.hip", &[]),
        ("fa_render.hip", &[]),
    ];
    let mut objs = Vec::new();
    for (src, extra) in sources.iter() {
        let obj = Path::new(&out_dir).join(src.replace(".hip", ".o"));
        let status = Command::new(&hipcc)
            .args(flags)
            .args(extra.iter())
            .arg("-c")
            .arg("-o")
            .arg(&obj)
            .arg(csrc.join(src))
            .status()
            .expect("failed to run hipcc (is ROCm installed?)");
        assert!(status.success(), "hipcc failed on {}", src);
        objs.push(obj);
    }
    let lib = Path::new(&out_dir).join("libfastatari_hip.so");
    let status = Command::new(&hipcc)
        .args(["--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o"])
        .arg(&lib)
        .args(objs.iter())
        .status()
        .expect("failed to run hipcc (is ROCm installed?)");
    assert!(status.success(), "hipcc link failed");
    println!("cargo:rustc-link-search=native={}", out_dir);
    println!("cargo:rustc-link-lib=dylib=fastatari_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out_dir);
    let headers = ["fa_emu.hpp", "fa_decode.hpp", "fa_frame.hpp", "fa_uniform.hpp", "fa_internal.hpp"];
    for f in sources.iter().map(|(s, _)| *s).chain(headers.iter().copied()) {
        println!("cargo:rerun-if-changed=fast-atari-rs_amd/csrc/{}", f);
    }
    println!("cargo:rerun-if-changed=include/fastatari.h");
}
