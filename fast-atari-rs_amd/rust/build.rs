// build.rs for the `hip` feature: compiles the HIP kernels and the C ABI
// with hipcc for gfx950 into $OUT_DIR/libfastatari_hip.so and links it.
// Replaces the nvcc/PTX step of the reference's build.rs:12-36.
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
    // fa_kernels.hip holds every kernel; the others are host code (C ABI,
    // multi-ROM batches, server)
    let sources = ["fa_kernels.hip", "fa_lib.hip", "fa_batch.hip", "fa_serve.hip"];
    let lib = Path::new(&out_dir).join("libfastatari_hip.so");
    let hipcc = std::env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".into());
    let status = Command::new(hipcc)
        .args([
            "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-pthread", "-o",
        ])
        .arg(&lib)
        .args(sources.iter().map(|f| csrc.join(f)))
        .status()
        .expect("failed to run hipcc (is ROCm installed?)");
    assert!(status.success(), "hipcc failed");
    println!("cargo:rustc-link-search=native={}", out_dir);
    println!("cargo:rustc-link-lib=dylib=fastatari_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out_dir);
    for f in sources.iter().chain(["fa_emu.hpp", "fa_decode.hpp", "fa_internal.hpp"].iter()) {
        println!("cargo:rerun-if-changed=fast-atari-rs_amd/csrc/{}", f);
    }
    println!("cargo:rerun-if-changed=include/fastatari.h");
}
