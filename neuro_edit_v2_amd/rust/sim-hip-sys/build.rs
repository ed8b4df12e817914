//! Link step: libneurohip.so is built offline by `make -C
//! neuro_edit_v2_amd/csrc` (hipcc, gfx950) and found through
//! NEUROHIP_LIB_DIR, defaulting to the in-tree build directory.
use std::path::PathBuf;

fn main() {
    let dir = std::env::var("NEUROHIP_LIB_DIR").map(PathBuf::from).unwrap_or_else(|_| {
        PathBuf::from(std::env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../_build")
    });
    println!("cargo:rerun-if-env-changed=NEUROHIP_LIB_DIR");
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=neurohip");
    // the library loads libamdhip64 from ROCm
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
