//! linalg.matmul of two general dense matrices through blas.gemm (reference
//! src/linalg/matmul/dgedge.zig:11-41), on `Layout`: with -Dlink_cblas=zmlhip
//! the product is cblas_dgemm / cblas_sgemm of libzmlhip. Fixes D1 / D2 / D3
//! (Order, orderOf, matrix.dense.General) on this path. The result takes A's
//! layout; a B stored in the other layout is read transposed.
//!
//! UNVERIFIED: no Zig toolchain was available to compile this file.
const std = @import("std");

const types = @import("../../types.zig");
const MulCoerce = types.MulCoerce;
const layoutOf = types.layoutOf;

const matrix = @import("../../matrix.zig");
const blas = @import("../blas.zig");

pub inline fn mm(allocator: std.mem.Allocator, a: anytype, b: anytype, ctx: anytype) !MulCoerce(@TypeOf(a), @TypeOf(b)) {
    const A = @TypeOf(a);
    const B = @TypeOf(b);
    const C = types.Coerce(types.Numeric(A), types.Numeric(B));
    const layout = comptime layoutOf(A);

    var c: matrix.general.Dense(C, layout) = try .init(allocator, a.rows, b.cols);
    errdefer c.deinit(allocator);

    try blas.gemm(layout, .no_trans, if (comptime layoutOf(B) == layout) .no_trans else .trans,
        @intCast(a.rows), @intCast(b.cols), @intCast(a.cols), 1, a.data, @intCast(a.ld), b.data,
        @intCast(b.ld), 0, c.data, @intCast(c.ld), ctx);
    return c;
}
