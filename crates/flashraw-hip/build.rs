//! Links libflashraw_hip.so: from FLASHRAW_HIP_LIB_DIR when set (the directory the MI355X
//! build wrote it to, `flashraw_amd/lib`), else from the linker's default search path.
fn main() {
    println!("cargo:rerun-if-env-changed=FLASHRAW_HIP_LIB_DIR");
    if let Ok(dir) = std::env::var("FLASHRAW_HIP_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    }
    println!("cargo:rustc-link-lib=dylib=flashraw_hip");
}
