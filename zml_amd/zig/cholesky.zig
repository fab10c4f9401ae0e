//! linalg.llt / utu / cholesky on `Layout` (reference
//! src/linalg/cholesky.zig:17-54,95-188), for zml's tree with this library
//! behind `-Dlink_lapacke`: the factorization is lapack.potrf, i.e.
//! LAPACKE_dpotrf from libzmlhip under -Dlink_lapacke=zmlhip. Fixes D1
//! (Order -> Layout), D2 (orderOf -> layoutOf) and D3 (matrixType's
//! .symmetric_dense / .hermitian_dense) on this path.
//!
//! UNVERIFIED: no Zig toolchain was available to compile this file.
const std = @import("std");

const types = @import("../types.zig");
const Layout = types.Layout;
const Uplo = types.Uplo;
const layoutOf = types.layoutOf;
const Numeric = types.Numeric;

const matrix = @import("../matrix.zig");
const linalg = @import("../linalg.zig");

/// A = F^T F (uplo .upper, F = U) or F F^T (.lower, F = L), F in an owned
/// n x n buffer with ld = n in the input's layout; the other triangle of the
/// buffer is whatever the copy held.
fn Factor(T: type, uplo: Uplo, layout: Layout) type {
    return struct {
        const Self = @This();

        _size: u32,
        _f: [*]T,
        _hermitian: bool,

        pub fn deinit(self: *Self, allocator: std.mem.Allocator) void {
            allocator.free(self._f[0 .. self._size * self._size]);
            self.* = undefined;
        }

        fn view(self: *const Self) matrix.triangular.Dense(T, uplo, .non_unit, layout) {
            return .{ .data = self._f, .rows = self._size, .cols = self._size, .ld = self._size,
                      .flags = .{ .owns_data = false } };
        }

        pub const l = if (uplo == .lower) view else @compileError("l() is the factor of an LLT");
        pub const u = if (uplo == .upper) view else @compileError("u() is the factor of a UTU");
    };
}

pub fn LLT(T: type, layout: Layout) type {
    return Factor(T, .lower, layout);
}
pub fn UTU(T: type, layout: Layout) type {
    return Factor(T, .upper, layout);
}

fn isSymOrHerm(comptime A: type) bool {
    if (!types.isMatrix(A)) return false;
    const mt = types.matrixType(A);
    return mt == .symmetric_dense or mt == .hermitian_dense;
}

/// Copies A into the requested triangle (mirrored when A stores the other
/// one), factors it with potrf and returns the owning factor. A non-positive
/// pivot gives error.SingularMatrix (cholesky.zig:124-127).
fn factor(comptime want: Uplo, allocator: std.mem.Allocator, a: anytype, ctx: anytype)
    !Factor(Numeric(@TypeOf(a)), want, layoutOf(@TypeOf(a))) {
    const A = @TypeOf(a);
    comptime if (!isSymOrHerm(A))
        @compileError("cholesky: argument must be a symmetric or hermitian dense matrix, got " ++ @typeName(A));
    comptime if (types.isArbitraryPrecision(Numeric(A)))
        @compileError("zml.linalg cholesky is not implemented for arbitrary precision types");
    types.validateContext(@TypeOf(ctx), .{});

    const n: u32 = a.size;
    var f = if (comptime types.uploOf(A) == want) try a.copy(allocator, ctx) else try a.copyInverseUplo(allocator, ctx);
    errdefer f.deinit(allocator);
    const info: i32 = try linalg.lapack.potrf(layoutOf(A), want, @intCast(n), f.data, @intCast(f.ld), ctx);
    if (info != 0) return linalg.Error.SingularMatrix;
    return .{ ._size = n, ._f = f.data, ._hermitian = comptime types.matrixType(A) == .hermitian_dense };
}

pub fn llt(allocator: std.mem.Allocator, a: anytype, ctx: anytype) !LLT(Numeric(@TypeOf(a)), layoutOf(@TypeOf(a))) {
    return factor(.lower, allocator, a, ctx);
}

pub fn utu(allocator: std.mem.Allocator, a: anytype, ctx: anytype) !UTU(Numeric(@TypeOf(a)), layoutOf(@TypeOf(a))) {
    return factor(.upper, allocator, a, ctx);
}

/// The factor in the triangle A stores.
pub fn cholesky(allocator: std.mem.Allocator, a: anytype, ctx: anytype)
    !Factor(Numeric(@TypeOf(a)), types.uploOf(@TypeOf(a)), layoutOf(@TypeOf(a))) {
    return factor(comptime types.uploOf(@TypeOf(a)), allocator, a, ctx);
}
