// Links libvt3d_dense.so. The library is built by video-to-3d_amd/csrc/Makefile into
// video-to-3d_amd/vt3d_dense/; VT3D_DENSE_LIB_DIR overrides where it is found. At run time
// the loader finds it through LD_LIBRARY_PATH or the final binary's rpath.
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=VT3D_DENSE_LIB_DIR");
    let dir = match env::var_os("VT3D_DENSE_LIB_DIR") {
        Some(d) => PathBuf::from(d),
        None => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap())
            .join("..")
            .join("..")
            .join("vt3d_dense"),
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=vt3d_dense");
}
