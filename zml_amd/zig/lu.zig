//! linalg.plu restored on `Layout` (reference src/linalg/lu.zig:59-128,178-221,
//! export commented out at src/linalg.zig:51-57), for zml's tree with this
//! library behind `-Dlink_lapacke`. Drop-in for src/linalg/lu.zig's PLU/plu;
//! re-export with `pub const plu = @import("linalg/lu.zig").plu;` in
//! src/linalg.zig. Fixes the snapshot's defects on this path:
//!   D1  types.Order -> types.Layout (values 101 / 102, types.zig:59-104)
//!   D2  types.orderOf -> types.layoutOf (types.zig:746)
//!   D3  matrix.General -> matrix.general.Dense; matrixType's .general_dense
//!   D4  PLU/plu exported again
//!   D5  lapack.getrf with an empty context (the typed dgetrf passes null)
//! The factorization itself is lapack.getrf, i.e. LAPACKE_dgetrf from
//! libzmlhip when zml is built with -Dlink_lapacke=zmlhip.
//!
//! UNVERIFIED: no Zig toolchain was available to compile this file.
const std = @import("std");

const types = @import("../types.zig");
const Layout = types.Layout;
const layoutOf = types.layoutOf;
const Numeric = types.Numeric;
const int = @import("../int.zig");

const matrix = @import("../matrix.zig");
const linalg = @import("../linalg.zig");

/// P A = L U of an m x n matrix: the factors share one owned buffer (unit
/// lower L below the diagonal, U on and above it) in the input's layout, and
/// ipiv holds LAPACK's 1-based interchanges (row i was swapped with row
/// ipiv[i] - 1, for i = 0, 1, ... in order).
pub fn PLU(T: type, layout: Layout) type {
    return struct {
        const Self = @This();

        _rows: u32,
        _cols: u32,
        _ipiv: []i32,
        _lu: [*]T,

        fn ld(self: *const Self) u32 {
            return if (comptime layout == .col_major) self._rows else self._cols;
        }

        pub fn deinit(self: *Self, allocator: std.mem.Allocator) void {
            allocator.free(self._ipiv);
            allocator.free(self._lu[0 .. self._rows * self._cols]);
            self.* = undefined;
        }

        /// The row permutation as a matrix: the interchanges replayed in
        /// reverse order on the identity (lu.zig:82-99).
        pub fn p(self: *const Self, allocator: std.mem.Allocator) !matrix.Permutation(T) {
            var perm: matrix.Permutation(T) = try .init(allocator, self._rows);
            errdefer perm.deinit(allocator);
            for (0..self._rows) |i| perm.data[i] = @intCast(i);
            var k: usize = self._ipiv.len;
            while (k > 0) {
                k -= 1;
                const q: usize = @intCast(self._ipiv[k] - 1);
                std.mem.swap(u32, &perm.data[k], &perm.data[q]);
            }
            return perm;
        }

        /// Unit lower m x min(m, n) view of the shared buffer.
        pub fn l(self: *const Self) matrix.triangular.Dense(T, .lower, .unit, layout) {
            return .{ .data = self._lu, .rows = self._rows, .cols = int.min(self._rows, self._cols),
                      .ld = self.ld(), .flags = .{ .owns_data = false } };
        }

        /// Upper min(m, n) x n view of the shared buffer.
        pub fn u(self: *const Self) matrix.triangular.Dense(T, .upper, .non_unit, layout) {
            return .{ .data = self._lu, .rows = int.min(self._rows, self._cols), .cols = self._cols,
                      .ld = self.ld(), .flags = .{ .owns_data = false } };
        }
    };
}

/// Copies the general dense matrix `a`, factors the copy with partial
/// pivoting and returns the owning factorization. An exactly zero pivot gives
/// error.SingularMatrix (lu.zig:212-213); the copy and ipiv are freed then.
pub fn plu(allocator: std.mem.Allocator, a: anytype, ctx: anytype) !PLU(Numeric(@TypeOf(a)), layoutOf(@TypeOf(a))) {
    const A = @TypeOf(a);
    comptime if (!types.isMatrix(A) or types.matrixType(A) != .general_dense)
        @compileError("plu: argument must be a general dense matrix, got " ++ @typeName(A));
    comptime if (types.isArbitraryPrecision(Numeric(A)))
        @compileError("zml.linalg.plu is not implemented for arbitrary precision types");
    types.validateContext(@TypeOf(ctx), .{});

    const m: u32 = a.rows;
    const n: u32 = a.cols;
    var lu = try a.copy(allocator, ctx);
    errdefer lu.deinit(allocator);
    const ipiv = try allocator.alloc(i32, int.min(m, n));
    errdefer allocator.free(ipiv);

    const info: i32 = try linalg.lapack.getrf(layoutOf(A), @intCast(m), @intCast(n), lu.data,
        @intCast(lu.ld), ipiv.ptr, .{});
    if (info != 0) return linalg.Error.SingularMatrix;

    return .{ ._rows = m, ._cols = n, ._ipiv = ipiv, ._lu = lu.data };
}
