// Links libferray_gemm.so: from FERRAY_GEMM_LIB_DIR when set, else the
// in-tree build next to this crate (../ferray_amd).  The library itself
// depends on the ROCm runtime (libamdhip64), found through its own runpath
// or LD_LIBRARY_PATH at run time.
fn main() {
    let dir = std::env::var("FERRAY_GEMM_LIB_DIR")
        .unwrap_or_else(|_| format!("{}/../ferray_amd", env!("CARGO_MANIFEST_DIR")));
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=ferray_gemm");
    println!("cargo:rerun-if-env-changed=FERRAY_GEMM_LIB_DIR");
}
