//! Link against libsdr_amd.so: the directory comes from SDR_AMD_LIB_DIR, or defaults to
//! the package directory that `python -m sdr_amd.build` writes the library into.
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=SDR_AMD_LIB_DIR");
    let dir = match env::var("SDR_AMD_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => {
            let here = PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR"));
            here.join("..").join("..")
        }
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=sdr_amd");
    // the library is found at run time next to where it was linked
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
