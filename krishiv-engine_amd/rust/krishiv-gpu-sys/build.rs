// Links libkgpu.so: KGPU_LIB_DIR names the directory that holds it (the
// in-tree build puts it in krishiv-engine_amd/kgpu/).
fn main() {
    let dir = std::env::var("KGPU_LIB_DIR").unwrap_or_else(|_| {
        concat!(env!("CARGO_MANIFEST_DIR"), "/../../kgpu").to_string()
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=kgpu");
    println!("cargo:rerun-if-env-changed=KGPU_LIB_DIR");
}
