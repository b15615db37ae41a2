// Builds libbuffgpu.so with hipcc (gfx950) through the repository's own build
// step and links it: the repository root is BUFF_AMD_ROOT, else two levels
// above this crate (buff_amd/rust).
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    let root = match env::var("BUFF_AMD_ROOT") {
        Ok(r) => PathBuf::from(r),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("..").join(".."),
    };
    let status = Command::new("python3")
        .args(&["-m", "buff_amd.build"])
        .current_dir(&root)
        .status()
        .expect("running python3 -m buff_amd.build (hipcc --offload-arch=gfx950)");
    assert!(status.success(), "building libbuffgpu.so failed");
    let lib = root.join("buff_amd");
    println!("cargo:rustc-link-search=native={}", lib.display());
    println!("cargo:rustc-link-lib=dylib=buffgpu");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib.display());
    for f in ["include/buffgpu.h", "buff_amd/csrc"].iter() {
        println!("cargo:rerun-if-changed={}", root.join(f).display());
    }
}
