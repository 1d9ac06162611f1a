// Bit-sliced persistent batch of Life-like tori of any width (gfx950, wave64).
//
// Many independent automata, each with its own size and its own rule
// next = born(count) | sustain(count) & state (the reference's LifelikeCA,
// examples/game_of_life.py:183-203; count = the eight Moore neighbours on
// a torus, xentica/core/topology/border.py:104-117), advance K steps in one
// launch.  Instance i is stored as rows of pitch_i = ceil(w_i / 32) 32-bit
// words, 32 cells per word with cell x in bit x % 32 of word x / 32; the
// bits past w_i in a row's last word are zero.
//
// Work is described per lane.  A wave's 64 lanes hold one or more segments;
// a segment is one (instance, column tile, row strip) with one lane per word
// column of the tile plus one halo lane on each side.  Lane l of a segment
// holds the "extended column" c: the 32 cells starting at cell 32 c taken
// modulo w (so c = -1 is the 32 cells before the row's start, and the last,
// partial word continues with the row's first cells).  Adjacent lanes hold
// adjacent 32-cell windows of the circular row, so the horizontal neighbour
// of every bit is the torus neighbour, whatever w is.  The halo lanes'
// outer bits take the next segment's (or a zero) word and go stale one bit
// per step from the outside, which never reaches an inner lane for K <= 31.
// Every lane of a wave runs the same rows in lock step: the wave's longest
// strip plus the 2K halo rows; a lane's own strip may be shorter and its
// extra rows are computed and dropped.  Each lane keeps its own running row
// (rows wrap per instance), its own rule (the mux-tree leaves are per-lane
// registers) and its own buffer offsets; loads and stores that a lane does
// not make carry an offset past the batch's end, which the buffer range
// check turns into a zero load and a dropped store.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xentica_hip.h"
#include "xentica_kernels.h"
#include "common.h"
#include "life_common.h"

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerGroup = 4;
constexpr int kRsrcFlags = 0x00020000;

// The rule with per-lane masks: leaf[c] = next state for count9 = c (the
// 3x3 block including the cell); a dead cell has c neighbours, a live one
// c - 1.  alive_next[c] / dead_next[c] are all-ones or zero per lane.
struct LaneRule {
    uint32_t alive_next[10], dead_next[10];

    __device__ __forceinline__ void init(uint32_t born, uint32_t sustain) {
#pragma unroll
        for (int c = 0; c < 10; ++c) {
            dead_next[c] = (c <= 8 && ((born >> c) & 1u)) ? 0xFFFFFFFFu : 0u;
            alive_next[c] =
                (c >= 1 && (((born | sustain) >> (c - 1)) & 1u)) ? 0xFFFFFFFFu : 0u;
        }
    }

    __device__ __forceinline__ uint32_t apply(uint32_t t0, uint32_t u, uint32_t k, uint32_t v,
                                              uint32_t m) const {
        const uint32_t w0 = u ^ k;
        const uint32_t w1 = bitop3<kXorAnd>(v, u, k);
        const uint32_t w2 = bitop3<kAnd3>(v, u, k);
        // count9 counts the cell: 0 only for a dead cell, 9 only for a live
        // one, so those two leaves need no select
        uint32_t leaf[10];
        leaf[0] = dead_next[0];
        leaf[9] = alive_next[9];
#pragma unroll
        for (int c = 1; c < 9; ++c) leaf[c] = bitop3<kMux>(m, alive_next[c], dead_next[c]);
        const uint32_t l0 = bitop3<kMux>(t0, leaf[1], leaf[0]);
        const uint32_t l1 = bitop3<kMux>(t0, leaf[3], leaf[2]);
        const uint32_t l2 = bitop3<kMux>(t0, leaf[5], leaf[4]);
        const uint32_t l3 = bitop3<kMux>(t0, leaf[7], leaf[6]);
        const uint32_t l4 = bitop3<kMux>(t0, leaf[9], leaf[8]);
        const uint32_t m0 = bitop3<kMux>(w0, l1, l0);
        const uint32_t m1 = bitop3<kMux>(w0, l3, l2);
        const uint32_t n0 = bitop3<kMux>(w1, m1, m0);
        return bitop3<kMux>(w2, l4, n0);
    }
};

__device__ __forceinline__ uint32_t batch_word(const HRow &a, const HRow &b, const HRow &c,
                                               const LaneRule &rule) {
    const uint32_t t0 = bitop3<kXor3>(a.h0, b.h0, c.h0);
    const uint32_t k = bitop3<kMaj3>(a.h0, b.h0, c.h0);
    const uint32_t u = bitop3<kXor3>(a.h1, b.h1, c.h1);
    const uint32_t v = bitop3<kMaj3>(a.h1, b.h1, c.h1);
    return rule.apply(t0, u, k, v, b.raw);
}

template <int K>
struct Levels {
    HRow lev[K + 1][3];
};

// Feed input row i (i % 3 == I3); level l produces row i - l when i >= 2l.
template <int K, int I3, bool kAll>
__device__ __forceinline__ void feed_rows(Levels<K> &w, uint32_t input, int i,
                                          const LaneRule &rule) {
    w.lev[0][I3] = hrow(input);
#pragma unroll
    for (int l = 1; l <= K; ++l) {
        if (kAll || i >= 2 * l) {
            const int j3 = ((I3 - l) % 3 + 3) % 3;
            const uint32_t v = batch_word(w.lev[l - 1][(j3 + 2) % 3], w.lev[l - 1][j3],
                                          w.lev[l - 1][(j3 + 1) % 3], rule);
            if (l < K)
                w.lev[l][j3] = hrow(v);
            else
                w.lev[l][j3].raw = v;
        }
    }
}

template <int K, int I3>
__device__ __forceinline__ uint32_t top_row(const Levels<K> &w) {
    return w.lev[K][((I3 - K) % 3 + 3) % 3].raw;
}

// The three words a lane needs to build its extended-column word of one
// row: a = word q, b = word q + 1 (zero past the row), z = word 0 (only
// when the 32 cells run past the row's end).
struct RowLoad {
    uint32_t a, b, z;
};

template <int K>
__global__ __launch_bounds__(kWave * kWavesPerGroup)
void life_batch_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                       unsigned total_bytes, const int4 *__restrict__ insts,
                       const int4 *__restrict__ lanes, const int *__restrict__ wave_rows,
                       int n_waves) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(
        static_cast<int>(blockIdx.x) * kWavesPerGroup + (static_cast<int>(threadIdx.x) >> 6));
    if (wave >= n_waves) return;
    // lane descriptor: instance (-1: idle), extended column, first output
    // row, output rows (0 for halo and idle lanes)
    const int4 d = lanes[static_cast<long long>(wave) * kWave + lane];
    const int L = __builtin_amdgcn_readfirstlane(wave_rows[wave]) + 2 * K;
    const bool idle = d.x < 0;
    // instance record: (base word, pitch, width, height), (born, sustain, -, -)
    const int4 g = idle ? make_int4(0, 1, 32, 1) : insts[2 * d.x];
    const int4 r = idle ? make_int4(0, 0, 0, 0) : insts[2 * d.x + 1];
    const unsigned base = static_cast<unsigned>(g.x) * 4u;
    const unsigned row_bytes = static_cast<unsigned>(g.y) * 4u;
    const int width = g.z, height = g.w;
    // the lane's 32 cells start at cell s of the row: word q, bit o; m
    // cells remain before the row's end
    int s = static_cast<int>((32LL * d.y) % width);
    if (s < 0) s += width;
    const int q = s >> 5, o = s & 31, m = width - s;
    const bool wraps = m < 32;
    const unsigned oor = total_bytes;
    const unsigned off_a = idle ? oor : static_cast<unsigned>(q) * 4u;
    const unsigned off_b = (!idle && q + 1 < g.y) ? static_cast<unsigned>(q + 1) * 4u : oor;
    const unsigned off_z = (!idle && wraps) ? 0u : oor;
    const uint32_t keep_m = wraps ? ((1u << m) - 1u) : 0xFFFFFFFFu;
    const int shift_z = wraps ? m : 0;
    // the last word of a row keeps only its w % 32 valid bits
    const uint32_t keep_out = (d.y == g.y - 1 && (width & 31)) ? ((1u << (width & 31)) - 1u)
                                                               : 0xFFFFFFFFu;

    const __amdgpu_buffer_rsrc_t in_rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(src), 0, total_bytes, kRsrcFlags);
    const __amdgpu_buffer_rsrc_t out_rsrc =
        __builtin_amdgcn_make_buffer_rsrc(dst, 0, total_bytes, kRsrcFlags);

    // running input row (wrapped per instance) and its byte offset
    int in_row = static_cast<int>((static_cast<long long>(d.z) - K) % (idle ? 1 : height));
    if (in_row < 0) in_row += height;
    unsigned in_off = base + static_cast<unsigned>(in_row) * row_bytes;
    // output rows: d.z .. d.z + d.w - 1, stored from feed 2K on
    unsigned out_off = base + static_cast<unsigned>(d.z) * row_bytes + static_cast<unsigned>(d.y) * 4u;
    const int out_rows = d.w;

    auto load_next = [&]() -> RowLoad {
        RowLoad v;
        v.a = __builtin_amdgcn_raw_buffer_load_b32(in_rsrc, off_a == oor ? oor : in_off + off_a, 0, 0);
        v.b = __builtin_amdgcn_raw_buffer_load_b32(in_rsrc, off_b == oor ? oor : in_off + off_b, 0, 0);
        v.z = __builtin_amdgcn_raw_buffer_load_b32(in_rsrc, off_z == oor ? oor : in_off + off_z, 0, 0);
        ++in_row;
        in_off += row_bytes;
        if (in_row == height) {
            in_row = 0;
            in_off = base;
        }
        return v;
    };
    auto word_of = [&](const RowLoad &v) -> uint32_t {
        return (__builtin_amdgcn_alignbit(v.b, v.a, static_cast<uint32_t>(o)) & keep_m) |
               (v.z << shift_z);
    };
    // output row t of the lane's strip
    auto store_row = [&](uint32_t val, int t) {
        const bool keep = t < out_rows;
        __builtin_amdgcn_raw_buffer_store_b32(val & keep_out, out_rsrc, keep ? out_off : oor, 0,
                                              0);
        out_off += row_bytes;
    };

    // highest user priority first, one step lower every two iterations
    // from the first (lower_priority): the waves of a SIMD finish together
    __builtin_amdgcn_s_setprio(3);
    LaneRule rule;
    rule.init(static_cast<uint32_t>(r.x), static_cast<uint32_t>(r.y));
    Levels<K> w;
#pragma unroll
    for (int i = 0; i < 2 * K; ++i) {
        const uint32_t in = word_of(load_next());
        switch (i % 3) {
            case 0: feed_rows<K, 0, false>(w, in, i, rule); break;
            case 1: feed_rows<K, 1, false>(w, in, i, rule); break;
            default: feed_rows<K, 2, false>(w, in, i, rule); break;
        }
    }
    constexpr int P0 = (2 * K) % 3, P1 = (2 * K + 1) % 3, P2 = (2 * K + 2) % 3;
    int i = 2 * K;
    // Six rows per iteration from `cur`; the next six are requested into
    // `nxt` at the top of the iteration and handed over at its end.  The
    // compiler makes that hand-over register copies at the loop head, and a
    // copy waits for its load: these loads were issued a whole iteration
    // earlier, so the wait leaves the stores in flight (on gfx9 `vmcnt`
    // counts stores as well).  Rows past L - 1 are wrapped rows, in bounds
    // and unused.
    RowLoad cur[6], nxt[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) cur[r] = load_next();
    // Every earlier load is consumed or long complete: wait for them here,
    // so that the loop head's waits are counted from the previous iteration
    // alone (the compiler merges the entry path into the loop head's count).
    __builtin_amdgcn_s_waitcnt(kVmcntZero);
    auto feed3 = [&](const RowLoad &x0, const RowLoad &x1, const RowLoad &x2, int row) {
        feed_rows<K, P0, true>(w, word_of(x0), row, rule);
        store_row(top_row<K, P0>(w), row - 2 * K);
        feed_rows<K, P1, true>(w, word_of(x1), row + 1, rule);
        store_row(top_row<K, P1>(w), row + 1 - 2 * K);
        feed_rows<K, P2, true>(w, word_of(x2), row + 2, rule);
        store_row(top_row<K, P2>(w), row + 2 - 2 * K);
    };
    int prio = 3, iters = 0;
    for (; i + 5 < L; i += 6) {
#pragma unroll
        for (int r = 0; r < 6; ++r) nxt[r] = load_next();
        feed3(cur[0], cur[1], cur[2], i);
        feed3(cur[3], cur[4], cur[5], i + 3);
#pragma unroll
        for (int r = 0; r < 6; ++r) cur[r] = nxt[r];
        // steps after iterations 1, 3 and 5: near 30, 55 and 80 % of a
        // 39-row strip's work (1 % faster than after 2, 4 and 6)
        if ((++iters & 1) == 1) lower_priority(prio);
    }
    // the last 0 .. 5 rows
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        if (i + r >= L) break;
        switch ((2 * K + r) % 3) {
            case 0:
                feed_rows<K, 0, true>(w, word_of(cur[r]), i + r, rule);
                store_row(top_row<K, 0>(w), i + r - 2 * K);
                break;
            case 1:
                feed_rows<K, 1, true>(w, word_of(cur[r]), i + r, rule);
                store_row(top_row<K, 1>(w), i + r - 2 * K);
                break;
            default:
                feed_rows<K, 2, true>(w, word_of(cur[r]), i + r, rule);
                store_row(top_row<K, 2>(w), i + r - 2 * K);
                break;
        }
    }
}

template <int K>
int launch_batch(const uint32_t *src, uint32_t *dst, unsigned total_bytes, const int4 *insts,
                 const int4 *lanes, const int *wave_rows, int n_waves, hipStream_t stream) {
    const unsigned groups = static_cast<unsigned>((n_waves + kWavesPerGroup - 1) / kWavesPerGroup);
    hipLaunchKernelGGL((life_batch_kernel<K>), dim3(groups), dim3(kWave * kWavesPerGroup), 0,
                       stream, src, dst, total_bytes, insts, lanes, wave_rows, n_waves);
    return xen::check(hipGetLastError(), "life_batch_kernel launch");
}

// Stored word column x of row y of a padded instance from a uint8 plane:
// one thread per word (the bits past w stay zero).
__global__ __launch_bounds__(256) void pack_rows_kernel(const uint8_t *__restrict__ cells,
                                                        uint32_t *__restrict__ bits, int width,
                                                        int height, int pitch) {
    const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= static_cast<long long>(pitch) * height) return;
    const int y = static_cast<int>(t / pitch), x = static_cast<int>(t % pitch);
    const uint8_t *row = cells + static_cast<long long>(y) * width;
    uint32_t word = 0;
    const int lo = 32 * x;
    const int count = width - lo < 32 ? width - lo : 32;
    for (int b = 0; b < count; ++b) word |= (static_cast<uint32_t>(row[lo + b]) & 1u) << b;
    bits[t] = word;
}

// A padded instance into its uint8 plane and/or int3 colours: one thread
// per cell.
__global__ __launch_bounds__(256) void expand_rows_kernel(const uint32_t *__restrict__ bits,
                                                          uint8_t *__restrict__ cells,
                                                          int *__restrict__ col, int width,
                                                          int height, int pitch, int4 c0, int4 c1) {
    const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= static_cast<long long>(width) * height) return;
    const int y = static_cast<int>(t / width), x = static_cast<int>(t % width);
    const uint32_t bit = (bits[static_cast<long long>(y) * pitch + (x >> 5)] >> (x & 31)) & 1u;
    if (cells) cells[t] = static_cast<uint8_t>(bit);
    if (col) {
        col[3 * t] = bit ? c1.x : c0.x;
        col[3 * t + 1] = bit ? c1.y : c0.y;
        col[3 * t + 2] = bit ? c1.z : c0.z;
    }
}

}  // namespace

extern "C" {

int xen_life_batch_launch(const void *src, void *dst, long long total_bytes, const void *insts,
                          const void *lanes, const int *wave_rows, int n_waves, int depth,
                          xen_stream stream) {
    if (src == dst) return xen::fail("xen_life_batch_launch: src and dst must differ", 22);
    if (total_bytes <= 0 || total_bytes >= (1LL << 31) || (total_bytes & 3))
        return xen::fail("xen_life_batch_launch: the batch must hold 4 .. 2^31 - 4 bytes "
                         "(whole words)", 22);
    if (n_waves < 1) return xen::fail("xen_life_batch_launch: no work", 22);
    auto s = reinterpret_cast<hipStream_t>(stream);
    auto in = static_cast<const uint32_t *>(src);
    auto out = static_cast<uint32_t *>(dst);
    auto is = static_cast<const int4 *>(insts);
    auto ls = static_cast<const int4 *>(lanes);
    const unsigned tb = static_cast<unsigned>(total_bytes);
    switch (depth) {
        case 1: return launch_batch<1>(in, out, tb, is, ls, wave_rows, n_waves, s);
        case 2: return launch_batch<2>(in, out, tb, is, ls, wave_rows, n_waves, s);
        case 3: return launch_batch<3>(in, out, tb, is, ls, wave_rows, n_waves, s);
        case 4: return launch_batch<4>(in, out, tb, is, ls, wave_rows, n_waves, s);
        case 5: return launch_batch<5>(in, out, tb, is, ls, wave_rows, n_waves, s);
        case 6: return launch_batch<6>(in, out, tb, is, ls, wave_rows, n_waves, s);
        case 7: return launch_batch<7>(in, out, tb, is, ls, wave_rows, n_waves, s);
        case 8: return launch_batch<8>(in, out, tb, is, ls, wave_rows, n_waves, s);
        default: return xen::fail("xen_life_batch_launch: depth must be 1..8", 22);
    }
}

int xen_bits_pack_rows(const uint8_t *cells, uint32_t *bits, int width, int height, int pitch,
                       xen_stream stream) {
    if (width < 1 || height < 1 || pitch != (width + 31) / 32)
        return xen::fail("xen_bits_pack_rows: pitch must be ceil(width / 32)", 22);
    const long long words = static_cast<long long>(pitch) * height;
    hipLaunchKernelGGL(pack_rows_kernel, dim3(static_cast<unsigned>((words + 255) / 256)),
                       dim3(256), 0, reinterpret_cast<hipStream_t>(stream), cells, bits, width,
                       height, pitch);
    return xen::check(hipGetLastError(), "pack_rows_kernel launch");
}

int xen_bits_expand_rows(const uint32_t *bits, uint8_t *cells, int32_t *colors, int width,
                         int height, int pitch, const int32_t *color_dead,
                         const int32_t *color_alive, xen_stream stream) {
    if (width < 1 || height < 1 || pitch != (width + 31) / 32)
        return xen::fail("xen_bits_expand_rows: pitch must be ceil(width / 32)", 22);
    if (colors && (!color_dead || !color_alive))
        return xen::fail("xen_bits_expand_rows: colours need both colour values", 22);
    if (!cells && !colors) return 0;
    const int4 c0 = colors ? make_int4(color_dead[0], color_dead[1], color_dead[2], 0)
                           : make_int4(0, 0, 0, 0);
    const int4 c1 = colors ? make_int4(color_alive[0], color_alive[1], color_alive[2], 0)
                           : make_int4(0, 0, 0, 0);
    const long long n = static_cast<long long>(width) * height;
    hipLaunchKernelGGL(expand_rows_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256),
                       0, reinterpret_cast<hipStream_t>(stream), bits, cells, colors, width,
                       height, pitch, c0, c1);
    return xen::check(hipGetLastError(), "expand_rows_kernel launch");
}

}  // extern "C"
