//! The GPU row of the reference's benchmark harness: the same workload as
//! libs/pairing/src/bench.zig:87-111 (nttWithTwiddles over 2^20 BN254 Fr,
//! input fromInt(i % 1000), root rootOfUnity(2^20), precomputed table), run
//! through GpuNtt (gpu.zig).  A maintainer adds these declarations to
//! bench.zig and one line to its main, next to the CPU row at :223:
//!
//!     _ = bench("NTT 2^20 Fr (twiddles, GPU)", 5, nttFrGpu);
//!
//! Each call copies the 32 MiB buffer to the device and back (the host-buffer
//! entry point keeps the reference's blocking, in-place semantics), so the row
//! reports the PCIe-inclusive time.  Not compiled here: the build image has no
//! Zig toolchain.
const std = @import("std");
const zf = @import("zig-field");
const ntt_mod = @import("zig-ntt");
const gpu = @import("gpu.zig");

const FrScalar = zf.Field(0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001);

var gpu_ctx: ?gpu.GpuNtt = null;
var gpu_buf: []FrScalar = &.{};
var gpu_tw: []const []const FrScalar = &.{};
var gpu_sink: usize = 0;

fn nttGpuSetup() void {
    if (gpu_buf.len != 0) return;
    const F = FrScalar;
    gpu_ctx = gpu.GpuNtt.init(0, 20) catch |e| std.debug.panic("GpuNtt.init failed: {s}", .{@errorName(e)});
    gpu_buf = std.heap.page_allocator.alloc(F, 1 << 20) catch @panic("oom");
    for (gpu_buf, 0..) |*x, i| x.* = F.fromInt(@as(u64, i % 1000));
    const root = F.rootOfUnity(gpu_buf.len) catch |e| std.debug.panic("rootOfUnity(2^20) failed: {s}", .{@errorName(e)});
    // Registered with the context: later nttWithTwiddles calls read the
    // device-resident copy instead of uploading the table.
    gpu_tw = gpu_ctx.?.precomputeTwiddles(F, 20, root, std.heap.page_allocator) catch @panic("oom");
}

pub fn nttFrGpu() void {
    nttGpuSetup();
    gpu_ctx.?.nttWithTwiddles(FrScalar, gpu_buf, 20, gpu_tw) catch |e|
        std.debug.panic("GpuNtt.nttWithTwiddles failed: {s}", .{@errorName(e)});
    gpu_sink ^= @intFromBool(gpu_buf[0].isZero());
}

// The CPU row's own function, for a side-by-side check of the two backends:
// the same input through zig-ntt and through GpuNtt gives the same bytes.
pub fn checkAgainstCpu() !void {
    const F = FrScalar;
    const a = try std.heap.page_allocator.alloc(F, 1 << 12);
    defer std.heap.page_allocator.free(a);
    const b = try std.heap.page_allocator.alloc(F, 1 << 12);
    defer std.heap.page_allocator.free(b);
    for (a, b, 0..) |*x, *y, i| {
        x.* = F.fromInt(@as(u64, i * 7 + 3));
        y.* = x.*;
    }
    const root = try F.rootOfUnity(a.len);
    try ntt_mod.ntt(F, a, 12, root);
    var g = try gpu.GpuNtt.init(0, 12);
    defer g.deinit();
    try g.ntt(F, b, 12, root);
    if (!std.mem.eql(u8, std.mem.sliceAsBytes(a), std.mem.sliceAsBytes(b))) return error.Mismatch;
}
