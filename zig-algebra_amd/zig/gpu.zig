//! GpuNtt: the zig-ntt call surface (libs/ntt/src/root.zig) on the MI355X
//! library libzntt_gpu.so, through its C ABI (include/zntt_gpu.h).
//!
//! A maintainer adds this file as libs/ntt/src/gpu.zig and wires the module
//! to the library (build.zig, next to the zig-ntt module at :519-527):
//!
//!     ntt_mod.addIncludePath(b.path("include"));
//!     ntt_mod.addLibraryPath(.{ .cwd_relative = "<path>/zig-algebra_amd/zntt_amd" });
//!     ntt_mod.linkSystemLibrary("zntt_gpu", .{});
//!     ntt_mod.link_libc = true;
//!
//! Every function mirrors root.zig signature for signature, with the same
//! argument meaning, element layout and error sets.  Written against Zig
//! 0.16 (build.zig.zon:5); the build image of this library has no Zig
//! toolchain, so it has not been compiled here.

const std = @import("std");

pub const Ctx = opaque {};
extern fn zntt_ctx_create(out: *?*Ctx, device: c_int, hbm_budget: u64) c_int;
extern fn zntt_ctx_destroy(ctx: *Ctx) void;
extern fn zntt_ctx_reserve(ctx: *Ctx, max_log_n: usize) c_int;
extern fn zntt_bit_reverse(ctx: *Ctx, field: c_int, data: *anyopaque, len: usize) c_int;
extern fn zntt_ntt(ctx: *Ctx, field: c_int, data: *anyopaque, len: usize, log_n: usize, root: *const anyopaque) c_int;
extern fn zntt_intt(ctx: *Ctx, field: c_int, data: *anyopaque, len: usize, log_n: usize, root: *const anyopaque) c_int;
extern fn zntt_precompute_twiddles(ctx: *Ctx, field: c_int, log_n: usize, root: *const anyopaque,
                                   stages: [*]const *anyopaque, num_stages: usize) c_int;
extern fn zntt_free_twiddles(ctx: *Ctx, stages: [*]const *const anyopaque, num_stages: usize) c_int;
extern fn zntt_ntt_with_twiddles(ctx: *Ctx, field: c_int, data: *anyopaque, len: usize, log_n: usize,
                                 stages: [*]const *const anyopaque, lens: [*]const usize, n: usize) c_int;
extern fn zntt_intt_with_twiddles(ctx: *Ctx, field: c_int, data: *anyopaque, len: usize, log_n: usize,
                                  stages: [*]const *const anyopaque, lens: [*]const usize, n: usize) c_int;

fn fieldId(comptime F: type) c_int {
    if (F.MODULUS == 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001) {
        comptime std.debug.assert(@sizeOf(F) == 32);
        return 1;
    }
    if (F.MODULUS == 0xFFFFFFFF00000001) {
        comptime std.debug.assert(@sizeOf(F) == 8);
        return 2;
    }
    if (F.MODULUS == 0x78000001) {
        comptime std.debug.assert(@sizeOf(F) == 8);
        return 3;
    }
    if (F.MODULUS == 0x7F000001) {
        comptime std.debug.assert(@sizeOf(F) == 8);
        return 4;
    }
    @compileError("GpuNtt supports BN254 Fr, Goldilocks, BabyBear and KoalaBear");
}

pub const GpuNtt = struct {
    ctx: *Ctx,

    // max_log_n: the largest transform the caller will run.  Everything the
    // seven functions need up to 2^max_log_n is allocated here
    // (zntt_ctx_reserve), so a full device surfaces as error.OutOfMemory from
    // init and never from a transform.
    pub fn init(device: c_int, max_log_n: usize) !GpuNtt {
        var c: ?*Ctx = null;
        if (zntt_ctx_create(&c, device, 0) != 0) return error.NoDevice;
        switch (zntt_ctx_reserve(c.?, max_log_n)) {
            0 => {},
            5 => {
                zntt_ctx_destroy(c.?);
                return error.OutOfMemory;
            },
            9 => {
                zntt_ctx_destroy(c.?);
                return error.Unsupported;
            },
            else => {
                zntt_ctx_destroy(c.?);
                return error.DeviceError;
            },
        }
        return .{ .ctx = c.? };
    }
    pub fn deinit(self: GpuNtt) void {
        zntt_ctx_destroy(self.ctx);
    }

    // Device-only statuses (OutOfMemory = 5 from a transform, DeviceError =
    // 6, Unsupported = 9) are not in zig-ntt's error sets.  After init's
    // reservation a transform at log_n <= max_log_n allocates nothing, so
    // they come only from a HIP failure or a size past the reservation; the
    // shim panics on them so every signature below keeps the reference's
    // exact error set.
    fn device(st: c_int) noreturn {
        std.debug.panic("zig-ntt GPU backend: status {d}", .{st});
    }

    pub fn bitReverse(self: GpuNtt, comptime F: type, data: []F) error{InvalidLength}!void {
        switch (zntt_bit_reverse(self.ctx, fieldId(F), data.ptr, data.len)) {
            0 => {},
            3 => return error.InvalidLength,
            else => |st| device(st),
        }
    }

    pub fn ntt(self: GpuNtt, comptime F: type, data: []F, log_n: usize, root: F)
        error{ LogTooLarge, LengthMismatch, InvalidLength }!void {
        switch (zntt_ntt(self.ctx, fieldId(F), data.ptr, data.len, log_n, &root)) {
            0 => {},
            1 => return error.LogTooLarge,
            2 => return error.LengthMismatch,
            else => |st| device(st),
        }
    }

    pub fn intt(self: GpuNtt, comptime F: type, data: []F, log_n: usize, root: F)
        error{ LogTooLarge, LengthMismatch, InvalidLength }!void {
        switch (zntt_intt(self.ctx, fieldId(F), data.ptr, data.len, log_n, &root)) {
            0 => {},
            1 => return error.LogTooLarge,
            2 => return error.LengthMismatch,
            else => |st| device(st),
        }
    }

    // The caller's allocator owns the table, as in root.zig:112-134; the
    // library fills it and keeps a device copy keyed by the stage pointers.
    pub fn precomputeTwiddles(self: GpuNtt, comptime F: type, log_n: usize, root: F,
                              allocator: std.mem.Allocator) error{ LogTooLarge, OutOfMemory }![]const []const F {
        if (log_n >= @bitSizeOf(usize)) return error.LogTooLarge;
        const tw = try allocator.alloc([]F, log_n);
        var done: usize = 0;
        errdefer {
            for (tw[0..done]) |t| allocator.free(t);
            allocator.free(tw);
        }
        const ptrs = try allocator.alloc(*anyopaque, log_n);
        defer allocator.free(ptrs);
        for (tw, 0..) |*t, s| {
            t.* = try allocator.alloc(F, @as(usize, 1) << @intCast(s));
            ptrs[s] = t.*.ptr;
            done += 1;
        }
        switch (zntt_precompute_twiddles(self.ctx, fieldId(F), log_n, &root, ptrs.ptr, log_n)) {
            0 => return tw,
            1 => return error.LogTooLarge,
            5 => return error.OutOfMemory,
            else => |st| device(st),
        }
    }

    pub fn freeTwiddles(self: GpuNtt, comptime F: type, twiddles: []const []const F,
                        allocator: std.mem.Allocator) void {
        var buf: [64]*const anyopaque = undefined;
        for (twiddles, 0..) |t, s| buf[s] = t.ptr;
        _ = zntt_free_twiddles(self.ctx, &buf, twiddles.len);
        for (twiddles) |t| allocator.free(t);
        allocator.free(twiddles);
    }

    pub fn nttWithTwiddles(self: GpuNtt, comptime F: type, data: []F, log_n: usize, twiddles: []const []const F)
        error{ LogTooLarge, LengthMismatch, InvalidTwiddles, InvalidLength }!void {
        return self.twiddled(zntt_ntt_with_twiddles, F, data, log_n, twiddles);
    }

    pub fn inttWithTwiddles(self: GpuNtt, comptime F: type, data: []F, log_n: usize, twiddles: []const []const F)
        error{ LogTooLarge, LengthMismatch, InvalidTwiddles, InvalidLength }!void {
        return self.twiddled(zntt_intt_with_twiddles, F, data, log_n, twiddles);
    }

    fn twiddled(self: GpuNtt, comptime f: anytype, comptime F: type, data: []F, log_n: usize,
                twiddles: []const []const F) error{ LogTooLarge, LengthMismatch, InvalidTwiddles, InvalidLength }!void {
        if (twiddles.len > 64) return error.InvalidTwiddles;
        var ptrs: [64]*const anyopaque = undefined;
        var lens: [64]usize = undefined;
        for (twiddles, 0..) |t, s| {
            ptrs[s] = t.ptr;
            lens[s] = t.len;
        }
        switch (f(self.ctx, fieldId(F), data.ptr, data.len, log_n, &ptrs, &lens, twiddles.len)) {
            0 => {},
            1 => return error.LogTooLarge,
            2 => return error.LengthMismatch,
            4 => return error.InvalidTwiddles,
            else => |st| device(st),
        }
    }
};
