// 3-D Life-like engines on a torus, with the reference's Moore list.
//
// In 3-D the reference's MooreNeighborhood keeps the centre and drops
// the offset (1, 1, 1): it filters itertools.product([-1, 0, 1], repeat=3)
// by `!= (0, 0)`, which never matches a 3-tuple, and keeps the first 26
// entries (xentica/core/topology/neighborhood.py:159-167).  So for a
// Life-like rule (examples/game_of_life.py:190-197) the count of a cell is
// the sum of its full 3x3x3 cube minus the cell at (x+1, y+1, z+1), and
// next = ((B >> count) & 1) | ((S >> count) & 1) & state.
//
// life3d_bits_kernel: 1 bit per cell, 32 cells per uint32 word, x fastest.
//   A wavefront is split into G = 64 / Wd lane groups; a group's Wd lanes
//   hold one x-row (lane = 32-bit word), so x-neighbours are a rotation
//   inside the group (ds_bpermute) and the torus wrap in x is exact.  Each
//   group owns a strip of YS rows (registers) and the wave streams a range
//   of z planes: for every plane, X = 3-cell x-sums (2 bits) of the YS+2
//   rows, XY = 3-row y-sums (4 bits); the count of plane z is the
//   bit-sliced sum XY(z-1) + XY(z) + XY(z+1) (5 bits) minus the (1,1,1)
//   bit.  B5/S45 collapses to three 3-input logic ops.
// life3d_u8_kernel: one uint8 per cell, one thread per cell reading its 26
//   neighbours directly -- a plain restatement kept as an independent check
//   of the bit-sliced kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "xentica_hip.h"
#include "xentica_kernels.h"
#include "common.h"

namespace {

constexpr int kWave = 64;
constexpr int kRsrcFlags = 0x00020000;  // dword 3 of a raw buffer resource (CDNA)

template <uint8_t kTable>
__device__ __forceinline__ uint32_t op3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, kTable);
}
#include "bitop3.h"
constexpr uint8_t kXor3 = bitop3_table([](bool a, bool b, bool c) { return a ^ b ^ c; });
constexpr uint8_t kMaj3 = bitop3_table([](bool a, bool b, bool c) { return (a + b + c) >= 2; });
constexpr uint8_t kMux = bitop3_table([](bool a, bool b, bool c) { return a ? b : c; });
constexpr uint8_t kNotANotBC = bitop3_table([](bool a, bool b, bool c) { return !a && !b && c; });
constexpr uint8_t kNotAAndBOrC = bitop3_table([](bool a, bool b, bool c) { return !a && (b || c); });
constexpr uint8_t kTwoOrOneAlive =
    bitop3_table([](bool b1, bool b0, bool s) { return (b1 && !b0) || (!b1 && b0 && s); });

__device__ __forceinline__ int wrap(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// value of lane `src` (ds_bpermute)
__device__ __forceinline__ uint32_t lane_read(uint32_t v, int src) {
    return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(v)));
}

struct Count5 {
    uint32_t b0, b1, b2, b3, b4;
};

// sum of three 2-bit numbers -> 4 bits (0..9)
__device__ __forceinline__ void add3_2(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1,
                                       uint32_t c0, uint32_t c1, uint32_t (&s)[4]) {
    s[0] = op3<kXor3>(a0, b0, c0);
    const uint32_t k0 = op3<kMaj3>(a0, b0, c0);
    const uint32_t x = op3<kXor3>(a1, b1, c1);
    const uint32_t y = op3<kMaj3>(a1, b1, c1);
    s[1] = x ^ k0;
    const uint32_t k1 = x & k0;
    s[2] = y ^ k1;
    s[3] = y & k1;
}

// sum of three 4-bit numbers (each 0..9) -> 5 bits (0..27)
__device__ __forceinline__ Count5 add3_4(const uint32_t (&a)[4], const uint32_t (&b)[4],
                                         const uint32_t (&c)[4]) {
    Count5 r;
    r.b0 = op3<kXor3>(a[0], b[0], c[0]);
    const uint32_t k = op3<kMaj3>(a[0], b[0], c[0]);
    const uint32_t t1 = op3<kXor3>(a[1], b[1], c[1]);
    const uint32_t u1 = op3<kMaj3>(a[1], b[1], c[1]);
    r.b1 = t1 ^ k;
    const uint32_t v1 = t1 & k;
    const uint32_t t2 = op3<kXor3>(a[2], b[2], c[2]);
    const uint32_t u2 = op3<kMaj3>(a[2], b[2], c[2]);
    r.b2 = op3<kXor3>(t2, u1, v1);
    const uint32_t u3 = op3<kMaj3>(t2, u1, v1);
    const uint32_t t4 = op3<kXor3>(a[3], b[3], c[3]);
    const uint32_t u4 = op3<kMaj3>(a[3], b[3], c[3]);
    r.b3 = op3<kXor3>(t4, u2, u3);
    const uint32_t u5 = op3<kMaj3>(t4, u2, u3);
    r.b4 = u4 | u5;  // the sum never reaches 32
    return r;
}

// count - bit (bit is set only where the 27-sum includes it)
__device__ __forceinline__ Count5 sub_bit(Count5 c, uint32_t bit) {
    Count5 r;
    r.b0 = c.b0 ^ bit;
    uint32_t br = ~c.b0 & bit;
    r.b1 = c.b1 ^ br;
    br = ~c.b1 & br;
    r.b2 = c.b2 ^ br;
    br = ~c.b2 & br;
    r.b3 = c.b3 ^ br;
    br = ~c.b3 & br;
    r.b4 = c.b4 ^ br;
    return r;
}

template <bool kFast>
struct Rule3;

template <>
struct Rule3<true> {  // B5/S45
    uint32_t born, sustain;
    // full: the 27-cell sum S; d: the dropped (1,1,1) cell; s: the state.
    // count = S - d, next = (count == 5) | (count == 4) & s: S is 4..6, so
    // next = (S4 == 0 & S3 == 0 & S2 == 1) & (d ? (S10 == 2 | S10 == 1 & s)
    //                                          : (S10 == 1 | S10 == 0 & s)),
    // with S10 the low two bits -- five logic ops, no subtraction.
    __device__ __forceinline__ uint32_t apply(const Count5 &full, uint32_t d, uint32_t s) const {
        const uint32_t p = op3<kNotANotBC>(full.b4, full.b3, full.b2);
        const uint32_t g0 = op3<kNotAAndBOrC>(full.b1, full.b0, s);
        const uint32_t g1 = op3<kTwoOrOneAlive>(full.b1, full.b0, s);
        return p & op3<kMux>(d, g1, g0);
    }
};

template <>
struct Rule3<false> {
    uint32_t born, sustain;
    __device__ __forceinline__ uint32_t apply(const Count5 &full, uint32_t d, uint32_t s) const {
        const Count5 c = sub_bit(full, d);
        uint32_t leaf[32];
#pragma unroll
        for (int v = 0; v < 32; ++v) {
            const uint32_t b = (v < 27 && ((born >> v) & 1u)) ? 0xFFFFFFFFu : 0u;
            const uint32_t k = (v < 27 && ((sustain >> v) & 1u)) ? 0xFFFFFFFFu : 0u;
            leaf[v] = op3<kMux>(s, k | b, b);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) leaf[v] = op3<kMux>(c.b0, leaf[2 * v + 1], leaf[2 * v]);
#pragma unroll
        for (int v = 0; v < 8; ++v) leaf[v] = op3<kMux>(c.b1, leaf[2 * v + 1], leaf[2 * v]);
#pragma unroll
        for (int v = 0; v < 4; ++v) leaf[v] = op3<kMux>(c.b2, leaf[2 * v + 1], leaf[2 * v]);
#pragma unroll
        for (int v = 0; v < 2; ++v) leaf[v] = op3<kMux>(c.b3, leaf[2 * v + 1], leaf[2 * v]);
        return op3<kMux>(c.b4, leaf[1], leaf[0]);
    }
};

// Per-plane products of one lane group: 4-bit y-sums of the YS output
// rows, the cells to the right (x+1) of rows 1..YS, and the raw words.
template <int YS>
struct Plane {
    uint32_t xy[YS][4];
    uint32_t right[YS];  // (x+1) bits of row y+1, for y = 0..YS-1
    uint32_t raw[YS];    // the cells of the output rows
};

template <int YS>
struct Geometry {
    __amdgpu_buffer_rsrc_t in;
    int row_off[YS + 2];  // per-lane byte offset of rows y0-1 .. y0+YS in a plane
    int plane_bytes, D;
    int left_src, right_src;
};

// 32 cells of a row as one word
__device__ __forceinline__ uint32_t load_word(__amdgpu_buffer_rsrc_t in, int voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(in, voff, soff, 0);
}

__device__ __forceinline__ void store_word(uint32_t w, __amdgpu_buffer_rsrc_t out, int voff,
                                           int soff) {
    __builtin_amdgcn_raw_buffer_store_b32(w, out, voff, soff, 0);
}

template <int YS>
__device__ __forceinline__ void load_plane(const Geometry<YS> &g, int z, Plane<YS> &p) {
    uint32_t x0[YS + 2], x1[YS + 2];
    const int soff = wrap(z, g.D) * g.plane_bytes;  // wave-uniform
#pragma unroll
    for (int r = 0; r < YS + 2; ++r) {
        const uint32_t w = load_word(g.in, g.row_off[r], soff);
        const uint32_t wl = lane_read(w, g.left_src), wr = lane_read(w, g.right_src);
        const uint32_t l = (w << 1) | (wl >> 31);  // cells at x - 1
        const uint32_t rr = (w >> 1) | (wr << 31);  // cells at x + 1
        x0[r] = op3<kXor3>(w, l, rr);
        x1[r] = op3<kMaj3>(w, l, rr);
        if (r >= 2) p.right[r - 2] = rr;
        if (r >= 1 && r <= YS) p.raw[r - 1] = w;
    }
#pragma unroll
    for (int y = 0; y < YS; ++y)
        add3_2(x0[y], x1[y], x0[y + 1], x1[y + 1], x0[y + 2], x1[y + 2], p.xy[y]);
}

// Output plane z from the y-sums of planes z-1 (a), z (b), z+1 (c);
// out_off[y] is the row's per-lane offset, or the buffer size (dropped by
// the range check) for rows beyond the field.
template <int YS, bool kFast>
__device__ __forceinline__ void emit_plane(const Plane<YS> &a, const Plane<YS> &b,
                                           const Plane<YS> &c, const Rule3<kFast> &rule,
                                           __amdgpu_buffer_rsrc_t out, const int (&out_off)[YS],
                                           int soff) {
#pragma unroll
    for (int y = 0; y < YS; ++y) {
        const Count5 full = add3_4(a.xy[y], b.xy[y], c.xy[y]);
        const uint32_t next = rule.apply(full, c.right[y], b.raw[y]);
        store_word(next, out, out_off[y], soff);
    }
}

template <int YS, bool kFast>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void life3d_bits_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, int Wd,
                        int H, int D, int y_tiles, int z_tiles, uint32_t born,
                        uint32_t sustain) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(
        static_cast<int>(blockIdx.x) * 4 + (static_cast<int>(threadIdx.x) >> 6));
    if (wave >= y_tiles * z_tiles) return;
    const int G = kWave / Wd;  // lane groups (x-rows) per wave
    const int group = lane / Wd, xw = lane % Wd;
    const int yt = wave % y_tiles, zt = wave / y_tiles;
    const int z_lo = static_cast<int>((static_cast<long long>(zt) * D) / z_tiles);
    const int z_hi = static_cast<int>((static_cast<long long>(zt + 1) * D) / z_tiles);
    constexpr int kUnit = 4;  // bytes per 32-cell lane word
    const unsigned bytes = static_cast<unsigned>(static_cast<size_t>(Wd) * H * D * kUnit);
    const int row_bytes = Wd * kUnit;
    const int y0 = (yt * G + group) * YS;

    Geometry<YS> g;
    g.in = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(src), 0, bytes, kRsrcFlags);
#pragma unroll
    for (int r = 0; r < YS + 2; ++r) g.row_off[r] = wrap(y0 - 1 + r, H) * row_bytes + xw * kUnit;
    g.plane_bytes = H * row_bytes;
    g.D = D;
    g.left_src = group * Wd + (xw + Wd - 1) % Wd;
    g.right_src = group * Wd + (xw + 1) % Wd;
    const __amdgpu_buffer_rsrc_t out = __builtin_amdgcn_make_buffer_rsrc(dst, 0, bytes, kRsrcFlags);
    int out_off[YS];
#pragma unroll
    for (int y = 0; y < YS; ++y)
        out_off[y] = y0 + y < H ? (y0 + y) * row_bytes + xw * kUnit : static_cast<int>(bytes);
    Rule3<kFast> rule{born, sustain};
    const int pb = g.plane_bytes;

    // highest user priority first, one step lower after each quarter of
    // the wave's planes: the waves of a SIMD finish together instead of
    // the oldest first (a SIMD issues from its oldest ready wave)
    __builtin_amdgcn_s_setprio(3);
    const int quarter = (z_hi - z_lo + 3) / 4;
    int prio = 3, next_step = z_lo + quarter;
    Plane<YS> p0, p1, p2;
    load_plane<YS>(g, z_lo - 1, p0);
    load_plane<YS>(g, z_lo, p1);
    int z = z_lo;  // next output plane; plane z+1 is loaded next
    for (; z + 2 < z_hi; z += 3) {
        load_plane<YS>(g, z + 1, p2);
        emit_plane<YS, kFast>(p0, p1, p2, rule, out, out_off, z * pb);
        load_plane<YS>(g, z + 2, p0);
        emit_plane<YS, kFast>(p1, p2, p0, rule, out, out_off, (z + 1) * pb);
        load_plane<YS>(g, z + 3, p1);
        emit_plane<YS, kFast>(p2, p0, p1, rule, out, out_off, (z + 2) * pb);
        if (z + 3 >= next_step) {
            next_step += quarter;
            --prio;
            if (prio == 2)
                __builtin_amdgcn_s_setprio(2);
            else if (prio == 1)
                __builtin_amdgcn_s_setprio(1);
            else if (prio == 0)
                __builtin_amdgcn_s_setprio(0);
        }
    }
    if (z < z_hi) {
        load_plane<YS>(g, z + 1, p2);
        emit_plane<YS, kFast>(p0, p1, p2, rule, out, out_off, z * pb);
        if (z + 1 < z_hi) {
            load_plane<YS>(g, z + 2, p0);
            emit_plane<YS, kFast>(p1, p2, p0, rule, out, out_off, (z + 1) * pb);
        }
    }
}

// One thread per cell: count over the reference's 26 offsets directly.
__global__ __launch_bounds__(256) void life3d_u8_kernel(const uint8_t *__restrict__ src,
                                                        uint8_t *__restrict__ dst, int W, int H,
                                                        int D, uint32_t born, uint32_t sustain) {
    const long long n = static_cast<long long>(W) * H * D;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int x = static_cast<int>(i % W);
        const int y = static_cast<int>((i / W) % H);
        const int z = static_cast<int>(i / (static_cast<long long>(W) * H));
        uint32_t count = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dx == 1 && dy == 1 && dz == 1) continue;  // the dropped offset
                    const int xx = wrap(x + dx, W), yy = wrap(y + dy, H), zz = wrap(z + dz, D);
                    count += src[(static_cast<long long>(zz) * H + yy) * W + xx] & 1u;
                }
        const uint32_t state = src[i] & 1u;
        dst[i] = static_cast<uint8_t>((((born >> count) & 1u) | (((sustain >> count) & 1u) & state)) & 1u);
    }
}

template <int YS, bool kFast>
int launch_bits(const uint32_t *src, uint32_t *dst, int Wd, int H, int D, int z_tiles,
                uint32_t born, uint32_t sustain, hipStream_t s) {
    const int G = kWave / Wd;
    const int y_tiles = (H + G * YS - 1) / (G * YS);
    const long long waves = static_cast<long long>(y_tiles) * z_tiles;
    hipLaunchKernelGGL((life3d_bits_kernel<YS, kFast>),
                       dim3(static_cast<unsigned>((waves + 3) / 4)), dim3(256), 0, s, src, dst,
                       Wd, H, D, y_tiles, z_tiles, born, sustain);
    return xen::check(hipGetLastError(), "life3d_bits_kernel launch");
}

}  // namespace

namespace {

// z ranges per launch when the caller passes 0: about three waves per
// SIMD of the device (measured on MI355X at 1024^3: 46 us per step at
// three waves per SIMD, 53 us at four, 49 us at two), each range at least
// four planes deep.
int auto_z_tiles(int Wd, int height, int depth) {
    static int simds = 0;
    if (!simds) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            simds = 4 * cus;
        else
            simds = 1024;
    }
    const int G = kWave / Wd;
    const long long y_tiles = (height + G * 8 - 1) / (G * 8);
    long long z = (3LL * simds) / y_tiles;
    const long long max_z = depth / 4 > 0 ? depth / 4 : 1;
    if (z > max_z) z = max_z;
    return z < 1 ? 1 : static_cast<int>(z);
}

// Validates a packed field and launches the bit-sliced kernel.
int step3d(const char *name, const void *src, void *dst, int width, int height, int depth,
           uint32_t born, uint32_t sustain, int z_tiles, xen_stream stream) {
    const std::string who(name);
    if (width % 32) return xen::fail(who + ": width must be a multiple of 32", 22);
    const int Wd = width / 32;
    if (Wd < 1 || Wd > 64 || (64 % Wd))
        return xen::fail(who + ": width / 32 must divide 64", 22);
    if (height < 1 || depth < 1) return xen::fail(who + ": empty field", 22);
    const long long unit = 4;
    if (static_cast<long long>(Wd) * height * depth * unit >= (1LL << 31))
        return xen::fail(who + ": field must be smaller than 2 GiB", 22);
    if (src == dst) return xen::fail(who + ": src and dst must differ", 22);
    if (z_tiles <= 0) z_tiles = auto_z_tiles(Wd, height, depth);
    if (z_tiles > depth) z_tiles = depth;
    auto s = reinterpret_cast<hipStream_t>(stream);
    auto in = static_cast<const uint32_t *>(src);
    auto out = static_cast<uint32_t *>(dst);
    const bool fast = (born & 0x7FFFFFFu) == (1u << 5) &&
                      (sustain & 0x7FFFFFFu) == ((1u << 4) | (1u << 5));
    if (fast)
        return launch_bits<8, true>(in, out, Wd, height, depth, z_tiles, born, sustain,
                                            s);
    return launch_bits<8, false>(in, out, Wd, height, depth, z_tiles, born, sustain, s);
}

}  // namespace

extern "C" {

int xen_life3d_bits_step(const void *src, void *dst, int width, int height, int depth,
                         uint32_t born, uint32_t sustain, int z_tiles, xen_stream stream) {
    return step3d("xen_life3d_bits_step", src, dst, width, height, depth, born, sustain,
                         z_tiles, stream);
}

int xen_life3d_u8_direct_step(const void *src, void *dst, int width, int height, int depth,
                              uint32_t born, uint32_t sustain, xen_stream stream) {
    if (width < 1 || height < 1 || depth < 1)
        return xen::fail("xen_life3d_u8_direct_step: empty", 22);
    if (src == dst) return xen::fail("xen_life3d_u8_direct_step: src and dst must differ", 22);
    const long long n = static_cast<long long>(width) * height * depth;
    long long groups = (n + 255) / 256;
    if (groups > 256 * 64) groups = 256 * 64;
    hipLaunchKernelGGL(life3d_u8_kernel, dim3(static_cast<unsigned>(groups)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<const uint8_t *>(src),
                       static_cast<uint8_t *>(dst), width, height, depth, born, sustain);
    return xen::check(hipGetLastError(), "life3d_u8_kernel launch");
}

}  // extern "C"
