//! Zig declarations of libzmlhip's device-resident API (include/zmlhip.h):
//! device pointers, a hipStream_t passed as an opaque pointer, asynchronous
//! except for the mixed solvers. The CBLAS / LAPACKE drop-ins need no
//! declaration here: zml already @cImports cblas.h and lapacke.h (src/c.zig).
//!
//! UNVERIFIED: no Zig toolchain was available to compile this file.
pub const Stream = ?*anyopaque;

pub const row_major: c_int = 101;
pub const col_major: c_int = 102;

/// zmlhip_ir_opts: trailing-update precision (0 FP32, 1 FP16, 2 BF16) and the
/// refinement / GMRES-IR / pivoting controls of the mixed solvers.
pub const IrOpts = extern struct {
    precision: c_int = 1,
    max_iter: c_int = 0,
    gmres_restart: c_int = 0,
    gmres_max_iter: c_int = 0,
    no_gmres: c_int = 0,
    tol: f64 = 0,
    no_optimistic: c_int = 0,
    pivot_threshold: f64 = 0,
};

/// zmlhip_ir_result
pub const IrResult = extern struct {
    info: c_int = 0,
    iterations: c_int = 0,
    gmres_iterations: c_int = 0,
    fallback: c_int = 0,
    backward_error: f64 = 0,
    panels: c_int = 0,
    panels_optimistic: c_int = 0,
};

pub const ESYNC: c_int = -1002;

pub extern fn zmlhip_last_error() [*:0]const u8;
pub extern fn zmlhip_stream_synchronize(stream: Stream) c_int;

pub extern fn zmlhip_dgemm(layout: c_int, transa: c_int, transb: c_int, m: c_int, n: c_int, k: c_int,
    alpha: f64, a: [*]const f64, lda: c_int, b: [*]const f64, ldb: c_int, beta: f64, c: [*]f64, ldc: c_int,
    stream: Stream) c_int;
pub extern fn zmlhip_dgemv(layout: c_int, trans: c_int, m: c_int, n: c_int, alpha: f64, a: [*]const f64,
    lda: c_int, x: [*]const f64, incx: c_int, beta: f64, y: [*]f64, incy: c_int, stream: Stream) c_int;
pub extern fn zmlhip_dgemv_noalloc(layout: c_int, trans: c_int, m: c_int, n: c_int, alpha: f64,
    a: [*]const f64, lda: c_int, x: [*]const f64, incx: c_int, beta: f64, y: [*]f64, incy: c_int,
    stream: Stream) c_int;
pub extern fn zmlhip_dlange(layout: c_int, norm: u8, m: c_int, n: c_int, a: [*]const f64, lda: c_int,
    result: *f64, stream: Stream) c_int;

/// *info: 0, the first zero pivot (1-based), or ESYNC when a panel's bounded
/// wait gave up.
pub extern fn zmlhip_dgetrf(layout: c_int, m: c_int, n: c_int, a: [*]f64, lda: c_int, ipiv: [*]i32,
    info: *i32, stream: Stream) c_int;
pub extern fn zmlhip_dgetrs(layout: c_int, trans: u8, n: c_int, nrhs: c_int, a: [*]const f64, lda: c_int,
    ipiv: [*]const i32, b: [*]f64, ldb: c_int, stream: Stream) c_int;
pub extern fn zmlhip_dpotrf(layout: c_int, uplo: u8, n: c_int, a: [*]f64, lda: c_int, info: *i32,
    stream: Stream) c_int;
pub extern fn zmlhip_dpotrs(layout: c_int, uplo: u8, n: c_int, nrhs: c_int, a: [*]const f64, lda: c_int,
    b: [*]f64, ldb: c_int, stream: Stream) c_int;

pub extern fn zmlhip_dgesv_mixed(layout: c_int, n: c_int, nrhs: c_int, a: [*]const f64, lda: c_int,
    b: [*]const f64, ldb: c_int, x: [*]f64, ldx: c_int, opts: ?*const IrOpts, result: *IrResult,
    stream: Stream) c_int;
pub extern fn zmlhip_dposv_mixed(layout: c_int, uplo: u8, n: c_int, nrhs: c_int, a: [*]const f64,
    lda: c_int, b: [*]const f64, ldb: c_int, x: [*]f64, ldx: c_int, opts: ?*const IrOpts,
    result: *IrResult, stream: Stream) c_int;

pub const BatchPlan = opaque {};
pub extern fn zmlhip_batch_plan_create(plan: *?*BatchPlan, count: c_int, n: [*]const c_int,
    nrhs: [*]const c_int, spd: ?[*]const c_int, a: [*]const [*]const f64, b: [*]const [*]const f64,
    x: [*]const [*]f64, opts: ?*const IrOpts, autotune: c_int, stream: Stream) c_int;
pub extern fn zmlhip_batch_plan_execute(plan: *BatchPlan, stream: Stream) c_int;
pub extern fn zmlhip_batch_plan_status(plan: *BatchPlan, status: [*]c_int, stream: Stream) c_int;
pub extern fn zmlhip_batch_plan_destroy(plan: *BatchPlan) c_int;
