// Link libvba.so: VBA_LIB_DIR names the directory that holds it (the
// library's own build directory, visloc-rs_amd/); without it the linker's
// default search path is used.
fn main() {
    println!("cargo:rerun-if-env-changed=VBA_LIB_DIR");
    match std::env::var("VBA_LIB_DIR") {
        Ok(dir) => println!("cargo:rustc-link-search=native={dir}"),
        Err(_) => println!("cargo:warning=VBA_LIB_DIR is not set: libvba.so must be on the linker's search path"),
    }
}
