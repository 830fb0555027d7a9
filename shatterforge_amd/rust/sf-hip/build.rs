// Links libsf_hip.so: set SF_HIP_DIR to the folder holding it (shatterforge_amd/ after `make -C shatterforge_amd/csrc`).
fn main() {
    if let Ok(dir) = std::env::var("SF_HIP_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
    }
    println!("cargo:rustc-link-lib=dylib=sf_hip");
    println!("cargo:rerun-if-env-changed=SF_HIP_DIR");
}
