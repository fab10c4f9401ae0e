//! hipcc step for zml's build.zig: compiles libzmlhip.so (the MI355X / gfx950
//! CBLAS + LAPACKE drop-in) from this repository's sources and links it into
//! the artifacts that consume `-Dlink_cblas` / `-Dlink_lapacke`
//! (reference build.zig:21-24, :45-53, :96-102).
//!
//! UNVERIFIED: written against Zig 0.15's std.Build API without a Zig
//! toolchain to compile it (none is installed where this library is built).
//!
//! Usage in zml's build.zig, after the link_cblas / link_lapacke options:
//!
//!     const zmlhip = @import("third_party/zml_amd/zml_amd/zig/build_zmlhip.zig");
//!     if (b.option(bool, "hip", "Use libzmlhip (MI355X) for CBLAS/LAPACKE") orelse false) {
//!         const lib = zmlhip.addLibrary(b, "third_party/zml_amd");
//!         zmlhip.linkInto(b, lib, exe, "third_party/zml_amd");
//!         zmlhip.linkInto(b, lib, lib_unit_tests, "third_party/zml_amd");
//!     }
//!
//! then `zig build -Dhip=true -Dlink_cblas=zmlhip -Dlink_lapacke=zmlhip test`.
const std = @import("std");

/// Flags of every translation unit (the same as zml_amd/build.py): gfx950
/// code, packed FP32 instructions disabled, C++17.
const common_flags = [_][]const u8{
    "-O3",                 "-std=c++17",        "-fPIC",       "--offload-arch=gfx950",
    "-Xclang",             "-target-feature",   "-Xclang",     "-packed-fp32-ops",
};

/// The HIP / C++ sources of the library (zml_amd/csrc), compiled as HIP.
const sources = [_][]const u8{
    "batched.hip", "blas_f64.hip", "capi.cpp", "gemm_f64.hip", "getrf_f64.hip",
    "mixed.hip",   "mp_gemm.hip",  "potrf_f64.hip", "runtime.cpp",
};

pub const Library = struct {
    /// the linked shared library (zml_amd/libzmlhip.so equivalent, in the cache)
    file: std.Build.LazyPath,
    /// the step that produces it
    step: *std.Build.Step,
};

/// One hipcc compile per source, then one hipcc link; `root` is where this
/// repository is vendored, relative to zml's build root.
pub fn addLibrary(b: *std.Build, root: []const u8) Library {
    const hipcc = b.option([]const u8, "hipcc", "hipcc executable") orelse "/opt/rocm/bin/hipcc";
    const link = b.addSystemCommand(&.{ hipcc, "-shared", "-fPIC", "--offload-arch=gfx950" });
    for (sources) |src| {
        const cc = b.addSystemCommand(&.{hipcc});
        if (std.mem.endsWith(u8, src, ".cpp")) cc.addArgs(&.{ "-x", "hip" });
        cc.addArgs(&common_flags);
        cc.addArg("-c");
        cc.addFileArg(b.path(b.fmt("{s}/zml_amd/csrc/{s}", .{ root, src })));
        cc.addArg("-o");
        const obj = cc.addOutputFileArg(b.fmt("{s}.o", .{src}));
        // headers the sources include: rebuild when one of them changes
        cc.addFileInput(b.path(b.fmt("{s}/zml_amd/csrc/zmlhip_internal.hpp", .{root})));
        cc.addFileInput(b.path(b.fmt("{s}/include/zmlhip.h", .{root})));
        link.addFileArg(obj);
    }
    link.addArg("-o");
    const so = link.addOutputFileArg("libzmlhip.so");
    return .{ .file = so, .step = &link.step };
}

/// Makes `art` include this repository's cblas.h / lapacke.h (the names
/// src/c.zig @cImports) and link libzmlhip, found at run time through an rpath.
pub fn linkInto(b: *std.Build, lib: Library, art: *std.Build.Step.Compile, root: []const u8) void {
    art.step.dependOn(lib.step);
    art.root_module.addIncludePath(b.path(b.fmt("{s}/include", .{root})));
    art.root_module.addLibraryPath(lib.file.dirname());
    art.root_module.addRPath(lib.file.dirname());
    art.root_module.linkSystemLibrary("zmlhip", .{});
    art.linkLibC();
}
