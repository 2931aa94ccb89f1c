// Links the prebuilt shared library. GRAPH_RENDER_FA2_LIB_DIR names the directory holding
// libgraph_render_fa2.so (graph_render_amd/_lib after `python graph_render_amd/build.py`).
fn main() {
    println!("cargo:rerun-if-env-changed=GRAPH_RENDER_FA2_LIB_DIR");
    if let Ok(dir) = std::env::var("GRAPH_RENDER_FA2_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
    }
    println!("cargo:rustc-link-lib=dylib=graph_render_fa2");
}
