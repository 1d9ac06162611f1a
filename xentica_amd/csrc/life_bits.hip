// Bit-packed Life-like engine for 2-D Moore tori (gfx950, wave64).
//
// State layout: one bit per cell, 32 cells per 32-bit word, x0 fastest
// (cell x of row y is bit x%32 of word y*W + x/32), which is byte-for-byte
// the layout of 64 cells per little-endian uint64.  The step is the
// reference's Life-like rule (examples/game_of_life.py:55-70, 190-197):
// born = (B >> count) & 1, sustained = (S >> count) & 1,
// next = born | sustained & state, count = live cells among the eight
// Moore neighbours on a torus (xentica/core/topology/border.py:104-117).
//
// Kernel structure (one wavefront per tile, no LDS, no barriers):
//   * lane l owns one word column; a wave spans 64 adjacent columns of
//     which the 62 inner ones are output (lanes 0 and 63 are halo: their
//     results go stale one bit per step from the outer edge, so up to 31
//     steps never reach an inner lane);
//   * the columns left over after the whole 62-column tiles (fewer than 62)
//     form a remainder column whose segments are narrower than a wave:
//     interior row segments of it share a wave, one group of rem + 2 lanes
//     each, so a 16384-cell row (512 words = 8 x 62 + 16) runs on 8 1/3
//     waves per row strip instead of 9;
//   * each lane streams the rows of its column through a K-level row
//     pipeline in VGPRs (temporal blocking: K steps per HBM round trip);
//     the strips of a workgroup hand their last rows of each level to the
//     next strip through LDS instead of computing halo rows;
//   * vertical neighbour sums come from registers, horizontal ones from the
//     adjacent lanes through DPP wave shifts plus a funnel shift;
//   * the neighbour count is a bit-sliced carry-save adder over 32 cells at
//     once; B3/S23 takes rows in pairs (a pair sum shared by two output
//     rows, then four 3-input logic ops per output), other rules three rows
//     and a mux tree over the count;
//   * waves step their priority down as they progress, so the waves of a
//     SIMD finish together.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "xentica_hip.h"
#include "xentica_kernels.h"
#include "common.h"
#include "life_common.h"

namespace {

constexpr int kWave = 64;
constexpr int kOutLanes = kWave - 2;
constexpr int kWavesPerGroup = 4;
// At most this many remainder segments share a wave: a wave's loads and
// stores then touch that many rows each, and on MI355X groups of 8 and 16
// made those waves the launch's slowest (2048^2: 0.90 and 1.07 us per step
// against 0.83 at 4 and 0.79 with none shared; 16384^2 gains 6 % at 3)
constexpr int kMaxGroup = 4;
// Depth from which whole tiles run as exchanging groups: the halo rows
// they save grow as K (K - 1), and at depths 1 and 2 the plain strips
// measured faster (16384^2: 15.2 and 15.9 us per launch against 17.6 and
// 18.9)
constexpr int kMinExchangeDepth = 4;
// dword 3 of a raw buffer resource on gfx9-family (CDNA) parts
constexpr int kRsrcFlags = 0x00020000;
// keep several waves per SIMD resident (caps VGPRs at 512 / this)
constexpr int kMinWavesPerSimd = 4;

__device__ __forceinline__ int wrap(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// Pipeline state: lev[l][j % 3] is row j of step level l (0 = input),
// with its horizontal sums for the levels that feed another level.
template <int K>
struct Window {
    HRow lev[K + 1][3];
};

// Feed input row i (with i % 3 == I3) into the pipeline and advance every
// level whose row is valid: level l produces row i - l when i >= 2l.
// kAll asserts i >= 2K (steady state, every level active).
template <int K, bool kFast, int I3, bool kAll>
__device__ __forceinline__ void feed(Window<K> &w, uint32_t input, int i,
                                     const Rule<kFast> &rule) {
    w.lev[0][I3] = hrow(input);
#pragma unroll
    for (int l = 1; l <= K; ++l) {
        if (kAll || i >= 2 * l) {
            const int j3 = ((I3 - l) % 3 + 3) % 3;
            const uint32_t v = life_word<kFast>(w.lev[l - 1][(j3 + 2) % 3], w.lev[l - 1][j3],
                                                w.lev[l - 1][(j3 + 1) % 3], rule);
            if (l < K)
                w.lev[l][j3] = hrow(v);
            else
                w.lev[l][j3].raw = v;  // the output level needs no sums
        }
    }
}

template <int K, int I3>
__device__ __forceinline__ uint32_t level_k_row(const Window<K> &w) {
    return w.lev[K][((I3 - K) % 3 + 3) % 3].raw;
}

// Drain feed D (2 <= D < 2K) of a tile that hands its last rows over to
// its successor: the input rows have ended (the tile loads out + 2 of
// them), level D / 2 takes its row from the successor's first two rows of
// that level (`imported`), the levels above it compute and the levels
// below it are done.  I3 is (out + D) % 3, the slot of the feed's row.
template <int K, bool kFast, int I3, int D>
__device__ __forceinline__ void drain_feed(Window<K> &w, uint32_t imported,
                                           const Rule<kFast> &rule) {
    constexpr int kMid = D / 2;
    w.lev[kMid][((I3 - kMid) % 3 + 3) % 3] = hrow(imported);
#pragma unroll
    for (int l = kMid + 1; l <= K; ++l) {
        const int j3 = ((I3 - l) % 3 + 3) % 3;
        const uint32_t v = life_word<kFast>(w.lev[l - 1][(j3 + 2) % 3], w.lev[l - 1][j3],
                                            w.lev[l - 1][(j3 + 1) % 3], rule);
        if (l < K)
            w.lev[l][j3] = hrow(v);
        else
            w.lev[l][j3].raw = v;
    }
}

// The 2K - 2 drain feeds of a tile whose out % 3 == S, each storing the
// level-K row it completes (output row out + D - 2K).
template <int K, bool kFast, int S, typename Store, int... Ds>
__device__ __forceinline__ void drain_all(Window<K> &w, const uint32_t *imported,
                                          const Rule<kFast> &rule, Store &store, int out,
                                          std::integer_sequence<int, Ds...>) {
    (
        [&] {
            constexpr int D = Ds + 2;
            constexpr int I3 = (S + D) % 3;
            drain_feed<K, kFast, I3, D>(w, imported[Ds], rule);
            store(level_k_row<K, I3>(w), out + D - 2 * K);
        }(),
        ...);
}

// Row pipeline of B3/S23 with rows taken in pairs (pair_life).  Rows of
// every level are numbered from the tile's first input row, and level l
// takes row n of level l - 1 with feed n + l - 1, so the parity of n is
// fixed in every unrolled feed.  Of the rows it has taken, level l keeps
//   after an odd row n:  the pair sum of rows n - 1 and n (q0, q1, q2) and
//                        the horizontal sum of row n (s0, s1);
//   after an even row n: the horizontal sums of rows n (q0, q1) and n - 1
//                        (s0, s1);
// and the state of row n (c).  Row n then completes level l's row n - 1:
// from the single row n - 2 and the new pair for odd n, from the pair kept
// and the single row n for even n.  Six registers a level, against nine
// for three rows with their sums.
template <int K>
struct PairPipe {
    static constexpr int kRows = 10;  // rows fed per steady-state iteration
    uint32_t q0[K + 1], q1[K + 1], q2[K + 1], s0[K + 1], s1[K + 1], c[K + 1];
    uint32_t last = 0;  // level K's row of the latest feed

    // Level l takes row n of level l - 1 and returns its row n - 1 when
    // `out` (n - 1 >= l); `pair`: row n - 1 of level l - 1 is valid.
    __device__ __forceinline__ uint32_t take(int l, const HRow &x, bool odd, bool pair,
                                             bool out) {
        uint32_t v = 0;
        if (odd) {
            uint32_t p0 = 0, p1 = 0, p2 = 0;
            if (pair) pair_sum(q0[l], q1[l], x.h0, x.h1, p0, p1, p2);
            if (out) v = pair_life(s0[l], s1[l], p0, p1, p2, c[l]);
            q0[l] = p0;
            q1[l] = p1;
            q2[l] = p2;
            s0[l] = x.h0;
            s1[l] = x.h1;
        } else {
            if (out) v = pair_life(x.h0, x.h1, q0[l], q1[l], q2[l], c[l]);
            q0[l] = x.h0;
            q1[l] = x.h1;
        }
        c[l] = x.raw;
        return v;
    }
    // Feed input row i (i % 2 == S): level l takes row i - l + 1 of level
    // l - 1 once that row exists (i >= 2l - 2) and completes its own row
    // i - l from i = 2l on.
    template <int S, bool kAll>
    __device__ __forceinline__ void feed(uint32_t input, int i) {
        HRow x = hrow(input);
#pragma unroll
        for (int l = 1; l <= K; ++l) {
            if (!kAll && i < 2 * l - 2) break;
            const uint32_t v =
                take(l, x, ((S + l + 1) & 1) != 0, kAll || i >= 2 * l - 1, kAll || i >= 2 * l);
            if (!kAll && i < 2 * l) break;
            if (l < K)
                x = hrow(v);
            else
                last = v;
        }
    }
    // level m's row completed by feed r (m < K): level m + 1 took it
    __device__ __forceinline__ uint32_t level_row(int m) const { return c[m + 1]; }
    // Drain feed D of a tile that hands its last rows over (as drain_feed;
    // S = (out + D) % 2, the parity of the feed's row).
    template <int S, int D>
    __device__ __forceinline__ void drain(uint32_t imported) {
        constexpr int kMid = D / 2;
        HRow x = hrow(imported);
#pragma unroll
        for (int l = kMid + 1; l <= K; ++l) {
            const uint32_t v = take(l, x, ((S + l + 1) & 1) != 0, true, true);
            if (l < K)
                x = hrow(v);
            else
                last = v;
        }
    }
};

// Rows requested ahead of the row being fed: the steady state loads the
// rows of its next iteration (nine for three-row slots, ten for pairs), so
// a wave reads up to row L + kAhead - 1 of its window of L rows.
constexpr int kAhead = 10;
static_assert(PairPipe<1>::kRows <= kAhead, "rows requested ahead");

// The 2K - 2 drain feeds of a B3/S23 tile whose out % 2 == S, each
// storing the level-K row it completes (output row out + D - 2K).
template <int K, int S, typename Store, int... Ds>
__device__ __forceinline__ void drain_pairs(PairPipe<K> &pipe, const uint32_t *imported,
                                            Store &store, int out,
                                            std::integer_sequence<int, Ds...>) {
    (
        [&] {
            constexpr int D = Ds + 2;
            pipe.template drain<(S + D) % 2, D>(imported[Ds]);
            store(pipe.last, out + D - 2 * K);
        }(),
        ...);
}

// Wave plan of one launch.  Waves [0, full * strips) are the whole
// 62-column tiles over `strips` row strips:
//   * xgroups == 0: balanced strips, each tile computing K halo rows above
//     and below its own rows (independent of every other wave);
//   * xgroups > 0: per column tile, xgroups balanced groups of R rows, each
//     run by one workgroup of kWavesPerGroup waves: its first G - 1 strips
//     have hx = (R + K - 1) / G rows and hand their last rows over to the
//     next strip of the group (see the kernel), the last strip keeps the
//     halo rows and gets R - (G - 1) hx, about K - 1 rows fewer, so that
//     every wave of the group does the same work.
// The rem < 62 columns left of the row form the remainder column, cut
// into `segs` row segments of seg_rows rows (the last one shorter);
// with xgroups > 0, seg_rows = hx - (K - 1), the rows whose halo work
// equals an exchanging strip's.  Segments [e_lo, e_hi) are interior: their
// window of rows, including the rows requested ahead, lies inside the
// field, so they need no row wrap, and they are packed `group` (at most
// kMaxGroup) to a wave, each in rem + 2 lanes at its own row offset.  The
// other segments run one to a wave.
struct Plan {
    int full, strips, rem, group, seg_rows, segs, e_lo, e_hi, xgroups;
    long long waves;
};

// Rows of each of a group's first kWavesPerGroup - 1 strips for a group of
// R rows at depth K (see Plan).  Rounded down, the last strip can do up to
// 6 % more work than the others; rounding up balanced them better but
// measured slower for B3/S23 at 16384^2 (3.15 against 3.08 us per step),
// while B36/S23 gained 2 %.
__host__ __device__ inline int exchange_rows(int R, int K) {
    return (R + K - 1) / kWavesPerGroup;
}

// Whether `groups` balanced groups of the H rows can run as exchanging
// workgroups at depth K: every strip but a group's last has at least 2K
// rows, the last at least one.
bool exchange_fits(int H, int K, int groups) {
    if (groups < 1 || groups > H) return false;
    const int rmin = H / groups;
    const int rmax = H % groups ? rmin + 1 : rmin;
    for (int R = rmin; R <= rmax; ++R) {
        const int hx = exchange_rows(R, K);
        if (hx < 2 * K || R - (kWavesPerGroup - 1) * hx < 1) return false;
    }
    return true;
}

// The plan for `strips` row strips of the whole tiles (1 <= strips <= H);
// at depths of at least kMinExchangeDepth, a multiple of kWavesPerGroup
// whose groups fit runs as exchanging groups.
Plan make_plan(int W, int H, int K, int strips) {
    Plan p{};
    p.full = W / kOutLanes;
    p.rem = W - p.full * kOutLanes;
    p.strips = strips;
    if (p.full > 0 && K >= kMinExchangeDepth && strips % kWavesPerGroup == 0 &&
        exchange_fits(H, K, strips / kWavesPerGroup))
        p.xgroups = strips / kWavesPerGroup;
    p.waves = static_cast<long long>(p.full) * strips;
    if (p.rem == 0) return p;
    p.group = kWave / (p.rem + 2);
    if (p.group > kMaxGroup) p.group = kMaxGroup;
    if (p.xgroups > 0)
        p.seg_rows = exchange_rows(H / p.xgroups, K) - (K - 1);
    else
        p.seg_rows = (H + strips - 1) / strips;
    p.segs = (H + p.seg_rows - 1) / p.seg_rows;
    if (p.group >= 2) {
        // interior: j * seg_rows >= K and (j + 1) * seg_rows + K + kAhead <= H
        p.e_lo = (K + p.seg_rows - 1) / p.seg_rows;
        if (p.e_lo > p.segs) p.e_lo = p.segs;
        const int hi = H - K - kAhead >= 0 ? (H - K - kAhead) / p.seg_rows : 0;
        p.e_hi = hi < p.e_lo ? p.e_lo : (hi > p.segs ? p.segs : hi);
    }
    const int shared = p.e_hi - p.e_lo;
    p.waves += (p.segs - shared) + (shared + p.group - 1) / (p.group > 0 ? p.group : 1);
    return p;
}

// One launch = K steps over the whole field.  Wave w owns a tile of
// output word columns [col_lo, col_lo + ncols) and output rows
// [row_lo, row_lo + out) (see Plan; a shared remainder wave owns nsegs
// such tiles, segment g at rows row_lo + g * seg_rows).  It streams the
// rows from row_lo - K on through a K-level row pipeline: when input row i
// arrives, level l computes row i - l from rows i-l-1 .. i-l+1 of level
// l-1, so only three rows per level are live.
//   * A halo tile loads L = out + 2K rows, level l computing rows
//     [l, L - l): the rows past out + l - 1 only feed this tile's own last
//     output rows, and the next strip computes them as well.
//   * An exchanging tile (a workgroup's strips but the last) loads out + 2
//     rows and computes only its rows [l, out + l) of each level.  Every
//     tile writes its first two rows of each level 1 .. K-1 (rows l and
//     l + 1, complete after the first 2K input rows) to LDS; after a
//     barrier, the exchanging tile's last 2K - 2 feeds take the rows
//     out + l and out + l + 1 of level l from the next strip's, which are
//     the same rows of the field.
template <int K, bool kFast>
__global__ __launch_bounds__(kWave * kWavesPerGroup)
__attribute__((amdgpu_waves_per_eu(kMinWavesPerSimd)))
void life_bits_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                      int W, int H, Plan plan, uint32_t born, uint32_t sustain) {
    constexpr int kLevels = K > 1 ? K - 1 : 1;  // levels handed over
    __shared__ uint32_t xch[kWavesPerGroup][kLevels][2][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    // wave-uniform tile coordinates live in scalar registers
    const int wig = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int wave = static_cast<int>(blockIdx.x) * kWavesPerGroup + wig;
    // waves past the plan still take part in the workgroup's barrier
    const bool live = wave < plan.waves;
    const int whole = plan.full * plan.strips;
    int col_lo = 0, ncols = 1, row_lo = 0, out = 1, seg_lanes = kWave, nsegs = 1;
    bool exchange = false;
    if (!live) {
        nsegs = 0;
    } else if (wave < whole && plan.xgroups > 0) {
        const int tx = wave / plan.strips;
        const int rest = wave - tx * plan.strips;
        const int j = rest / kWavesPerGroup;  // the workgroup's row group
        const int g0 = static_cast<int>((static_cast<long long>(j) * H) / plan.xgroups);
        const int g1 = static_cast<int>((static_cast<long long>(j + 1) * H) / plan.xgroups);
        const int hx = exchange_rows(g1 - g0, K);
        col_lo = tx * kOutLanes;
        ncols = kOutLanes;
        row_lo = g0 + wig * hx;
        exchange = wig < kWavesPerGroup - 1;
        out = exchange ? hx : g1 - row_lo;
    } else if (wave < whole) {
        const int tx = wave % plan.full;
        const int ty = wave / plan.full;
        col_lo = tx * kOutLanes;
        ncols = kOutLanes;
        // balanced split of the H rows over the strips
        row_lo = static_cast<int>((static_cast<long long>(ty) * H) / plan.strips);
        out = static_cast<int>((static_cast<long long>(ty + 1) * H) / plan.strips) - row_lo;
    } else {
        const int r = wave - whole;
        const int shared = plan.e_hi - plan.e_lo;
        const int solo = plan.segs - shared;
        col_lo = plan.full * kOutLanes;
        ncols = plan.rem;
        int seg;
        if (r < solo) {
            seg = r < plan.e_lo ? r : plan.e_hi + (r - plan.e_lo);
        } else {
            seg = plan.e_lo + (r - solo) * plan.group;
            seg_lanes = plan.rem + 2;
            nsegs = plan.e_hi - seg < plan.group ? plan.e_hi - seg : plan.group;
        }
        row_lo = seg * plan.seg_rows;
        out = (H - row_lo < plan.seg_rows ? H - row_lo : plan.seg_rows);
    }
    const int L = exchange ? out + 2 : out + 2 * K;  // input rows fed
    const int row0 = row_lo - K;

    // Buffer resources: scalar base and row offsets, one per-lane column
    // offset (plus the segment's row offset in a shared wave).
    // Out-of-range offsets are dropped by the hardware range check, which
    // is how masked-off stores and idle lanes are expressed.
    const unsigned bytes = static_cast<unsigned>(static_cast<size_t>(W) * H * 4u);
    const __amdgpu_buffer_rsrc_t in_rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(src), 0, bytes, kRsrcFlags);
    const __amdgpu_buffer_rsrc_t out_rsrc =
        __builtin_amdgcn_make_buffer_rsrc(dst, 0, bytes, kRsrcFlags);
    const int row_bytes = W * 4;
    const int seg = lane / seg_lanes;
    const int sl = lane - seg * seg_lanes;  // lane within its segment
    const bool active = seg < nsegs;
    const int seg_off = seg * plan.seg_rows * row_bytes;
    const int col = wrap(col_lo - 1 + sl, W);
    const int lane_off = active ? col * 4 + seg_off : static_cast<int>(bytes);
    const int out_col = col_lo - 1 + sl;
    const bool lane_out = active && sl >= 1 && sl <= ncols;
    const int out_voff = lane_out ? out_col * 4 + seg_off : static_cast<int>(bytes);
    const int out_row0 = row_lo;
    const int out_rows = out;

    // Input rows are requested strictly in order (row0, row0 + 1, ... on
    // the torus), so the scalar byte offset of the next one advances by a
    // row and wraps to 0 at the field's end: three scalar instructions a
    // row instead of a signed remainder.
    const int field_bytes = H * row_bytes;
    int next_off = wrap(row0, H) * row_bytes;
    auto load_next = [&]() -> uint32_t {
        const uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(in_rsrc, lane_off, next_off, 0);
        next_off += row_bytes;
        next_off = next_off == field_bytes ? 0 : next_off;
        return v;
    };
    // output row r = i - 2K of the tile, fed at input row i
    auto store_row = [&](uint32_t val, int r) {
        const bool keep = r < out_rows;
        __builtin_amdgcn_raw_buffer_store_b32(val, out_rsrc,
                                              keep ? out_voff : static_cast<int>(bytes),
                                              keep ? (out_row0 + r) * row_bytes : 0, 0);
    };

    // Every wave starts at the highest user priority and lowers it by one
    // after each steady-state iteration (lower_priority), so the three waves
    // of a SIMD finish close together instead of oldest first (16384^2: the
    // last wave of a SIMD ended 25.4 us after the launch's first wave
    // started instead of 28.8, and a step took 3.08 us instead of 3.30 on
    // one box).
    __builtin_amdgcn_s_setprio(3);
    if constexpr (kFast) {
        // B3/S23: rows in pairs, ten rows per steady-state iteration (an
        // even count keeps each feed's row parity fixed); otherwise as for
        // any rule below
        constexpr int kRows = PairPipe<K>::kRows;
        PairPipe<K> pipe;
        int i = 2 * K;
        uint32_t cur[kRows], nxt[kRows];
        uint32_t *const mine = &xch[wig][0][0][lane];
        if (live) {
            uint32_t pro[2 * K];
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) pro[r] = load_next();
#pragma unroll
            for (int r = 0; r < kRows; ++r) cur[r] = load_next();
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) {
                if (r % 2 == 0)
                    pipe.template feed<0, false>(pro[r], r);
                else
                    pipe.template feed<1, false>(pro[r], r);
                if (r >= 2) {
                    const int m = r / 2;
                    mine[((m - 1) * 2 + (r & 1)) * kWave] = pipe.level_row(m);
                }
            }
        }
        __syncthreads();
        if (!live) return;
        // rows row (even) and row + 1
        auto feed2 = [&](uint32_t x0, uint32_t x1, int row) {
            pipe.template feed<0, true>(x0, row);
            store_row(pipe.last, row - 2 * K);
            pipe.template feed<1, true>(x1, row + 1);
            store_row(pipe.last, row + 1 - 2 * K);
        };
        __builtin_amdgcn_s_waitcnt(kVmcntZero);
        int prio = 3;
        for (; i + kRows - 1 < L; i += kRows) {
#pragma unroll
            for (int r = 0; r < kRows; ++r) nxt[r] = load_next();
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < kRows; r += 2) {
                feed2(cur[r], cur[r + 1], i + r);
                // the step down after six of the ten rows puts the three
                // steps of an exchanging strip near a quarter, a half and
                // three quarters of its work (2 % faster at 16384^2 than
                // at the end of the iteration)
                if (r == 4) lower_priority(prio);
            }
#pragma unroll
            for (int r = 0; r < kRows; ++r) cur[r] = nxt[r];
        }
        // the last 0 .. kRows - 1 rows
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            if (i + r >= L) break;
            if (r % 2 == 0)
                pipe.template feed<0, true>(cur[r], i + r);
            else
                pipe.template feed<1, true>(cur[r], i + r);
            store_row(pipe.last, i + r - 2 * K);
        }
        if constexpr (K > 1) {
            if (!exchange) return;
            uint32_t imported[2 * K - 2];
            const uint32_t *const next = &xch[wig + 1][0][0][lane];
#pragma unroll
            for (int d = 0; d < 2 * K - 2; ++d) imported[d] = next[d * kWave];
            using Feeds = std::make_integer_sequence<int, 2 * K - 2>;
            if (out % 2 == 0)
                drain_pairs<K, 0>(pipe, imported, store_row, out, Feeds{});
            else
                drain_pairs<K, 1>(pipe, imported, store_row, out, Feeds{});
        }
    } else {
        Rule<kFast> rule{born, sustain};
        Window<K> w;
        // The steady state feeds nine rows per iteration from the registers
        // `cur`; at the top of each iteration the nine rows of the next one are
        // requested into `nxt`, and `cur = nxt` moves them over at the end.  The
        // compiler turns that hand-over (and, in practice, any other order of
        // loads across the back edge) into register copies at the loop head,
        // and a copy waits for its load: here those loads were issued a whole
        // iteration (72 row-levels of work) earlier, so the wait is normally
        // already satisfied and the stores stay in flight.  The prologue's 2K
        // rows and the first nine are all requested before the first feed.
        // Rows past L-1 are wrapped rows: in bounds, loaded and never fed.
        constexpr int P0 = (2 * K) % 3, P1 = (2 * K + 1) % 3, P2 = (2 * K + 2) % 3;
        int i = 2 * K;
        uint32_t cur[9], nxt[9];
        uint32_t *const mine = &xch[wig][0][0][lane];
        if (live) {
            uint32_t pro[2 * K];
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) pro[r] = load_next();
#pragma unroll
            for (int r = 0; r < 9; ++r) cur[r] = load_next();
            // prologue: rows 0 .. 2K-1 fill the pipeline (no output yet); feed r
            // completes row r - l of each level l <= r / 2, the first two rows of
            // level r / 2 at r = 2l and 2l + 1: those go to LDS
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) {
                switch (r % 3) {
                    case 0: feed<K, kFast, 0, false>(w, pro[r], r, rule); break;
                    case 1: feed<K, kFast, 1, false>(w, pro[r], r, rule); break;
                    default: feed<K, kFast, 2, false>(w, pro[r], r, rule); break;
                }
                if (r >= 2) {
                    const int m = r / 2;
                    mine[((m - 1) * 2 + (r & 1)) * kWave] = w.lev[m][((r - m) % 3 + 3) % 3].raw;
                }
            }
        }
        __syncthreads();
        if (!live) return;
        // rows row .. row+2 (slots P0, P1, P2 by row mod 3), each output row
        // r = row - 2K stored as its level-K row leaves the pipeline
        auto feed3 = [&](uint32_t x0, uint32_t x1, uint32_t x2, int row) {
            feed<K, kFast, P0, true>(w, x0, row, rule);
            store_row(level_k_row<K, P0>(w), row - 2 * K);
            feed<K, kFast, P1, true>(w, x1, row + 1, rule);
            store_row(level_k_row<K, P1>(w), row + 1 - 2 * K);
            feed<K, kFast, P2, true>(w, x2, row + 2, rule);
            store_row(level_k_row<K, P2>(w), row + 2 - 2 * K);
        };
        // Every load so far has been consumed or is long complete: wait for them
        // here, so that the loop head's waits are counted from the previous
        // iteration alone (the compiler merges the entry path into the loop
        // head's count, and the prologue's load order would otherwise make each
        // hand-over wait for the previous iteration's stores as well).
        __builtin_amdgcn_s_waitcnt(kVmcntZero);
        int prio = 3;
        for (; i + 8 < L; i += 9) {
#pragma unroll
            for (int r = 0; r < 9; ++r) nxt[r] = load_next();  // rows i + 9 .. i + 17
            // keep the nine loads at the top of the iteration: the scheduler
            // otherwise sinks them towards their use, and the hand-over at the
            // loop head then waits on loads issued late in the iteration
            __builtin_amdgcn_sched_barrier(0);
            feed3(cur[0], cur[1], cur[2], i);
            feed3(cur[3], cur[4], cur[5], i + 3);
            feed3(cur[6], cur[7], cur[8], i + 6);
#pragma unroll
            for (int r = 0; r < 9; ++r) cur[r] = nxt[r];
            lower_priority(prio);
        }
        // the last 0 .. 8 rows
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            if (i + r >= L) break;
            switch ((2 * K + r) % 3) {
                case 0:
                    feed<K, kFast, 0, true>(w, cur[r], i + r, rule);
                    store_row(level_k_row<K, 0>(w), i + r - 2 * K);
                    break;
                case 1:
                    feed<K, kFast, 1, true>(w, cur[r], i + r, rule);
                    store_row(level_k_row<K, 1>(w), i + r - 2 * K);
                    break;
                default:
                    feed<K, kFast, 2, true>(w, cur[r], i + r, rule);
                    store_row(level_k_row<K, 2>(w), i + r - 2 * K);
                    break;
            }
        }
        if constexpr (K > 1) {
            if (!exchange) return;
            // drain: the next strip's first two rows of each level 1 .. K-1, in
            // the order the feeds D = 2 .. 2K-1 take them (level D / 2, row D % 2)
            uint32_t imported[2 * K - 2];
            const uint32_t *const next = &xch[wig + 1][0][0][lane];
#pragma unroll
            for (int d = 0; d < 2 * K - 2; ++d) imported[d] = next[d * kWave];
            using Feeds = std::make_integer_sequence<int, 2 * K - 2>;
            switch (out % 3) {
                case 0: drain_all<K, kFast, 0>(w, imported, rule, store_row, out, Feeds{}); break;
                case 1: drain_all<K, kFast, 1>(w, imported, rule, store_row, out, Feeds{}); break;
                default: drain_all<K, kFast, 2>(w, imported, rule, store_row, out, Feeds{}); break;
            }
        }
    }
}

template <int K, bool kFast>
int launch_k(const uint32_t *src, uint32_t *dst, int W, int H, const Plan &plan,
             uint32_t born, uint32_t sustain, hipStream_t stream) {
    const unsigned groups =
        static_cast<unsigned>((plan.waves + kWavesPerGroup - 1) / kWavesPerGroup);
    hipLaunchKernelGGL((life_bits_kernel<K, kFast>), dim3(groups), dim3(kWave * kWavesPerGroup),
                       0, stream, src, dst, W, H, plan, born, sustain);
    return xen::check(hipGetLastError(), "life_bits_kernel launch");
}

template <bool kFast>
int launch_any_k(int k, const uint32_t *src, uint32_t *dst, int W, int H, const Plan &plan,
                 uint32_t born, uint32_t sustain, hipStream_t stream) {
    switch (k) {
        case 1: return launch_k<1, kFast>(src, dst, W, H, plan, born, sustain, stream);
        case 2: return launch_k<2, kFast>(src, dst, W, H, plan, born, sustain, stream);
        case 3: return launch_k<3, kFast>(src, dst, W, H, plan, born, sustain, stream);
        case 4: return launch_k<4, kFast>(src, dst, W, H, plan, born, sustain, stream);
        case 5: return launch_k<5, kFast>(src, dst, W, H, plan, born, sustain, stream);
        case 6: return launch_k<6, kFast>(src, dst, W, H, plan, born, sustain, stream);
        case 7: return launch_k<7, kFast>(src, dst, W, H, plan, born, sustain, stream);
        case 8: return launch_k<8, kFast>(src, dst, W, H, plan, born, sustain, stream);
        default: return xen::fail("life_bits: temporal depth must be 1..8", 22);
    }
}

// Modelled cost of one wave of a plan, in level rows per level: a halo
// strip of S rows streams S + 2K rows; an exchanging group's waves do the
// work of hx - (K - 1) output rows of a halo strip (the last strip of a
// group and the remainder segments are sized to match).
double wave_cost(const Plan &p, int H, int K) {
    if (p.xgroups > 0)
        return static_cast<double>(exchange_rows(H / p.xgroups, K)) + K + 1.0;
    return static_cast<double>(H) / static_cast<double>(p.strips) + 2.0 * K;
}

// Row strips per launch when the caller passes 0.  A launch is bound by
// VALU issue; shorter strips add halo work while more waves per SIMD hide
// more latency.  For w = 1, 2, 3 waves per SIMD of the device it takes
// the most strips whose plan has at most w waves per SIMD (so that no CU
// receives one workgroup more than the others, at most one strip per
// row), once over all strip counts and once over the counts that run as
// exchanging groups, and models the time as
//   max(1, waves / SIMDs) * wave_cost / issue(min(w, ceil(waves / SIMDs))),
// with issue(w) the share of a SIMD's VALU slots that w waves keep busy,
// measured on MI355X at depth 8: 0.55, 0.85 and 1.  The smallest wins.
// Measured optima of the model with balanced halo strips: 16384^2 three
// waves per SIMD (4.6 us per step against 5.6 at 1.5 waves); 4096^2 and
// 2048^2 one wave per SIMD (1.27 and 0.83 us per step, against 1.89 and
// 1.78 with strips of at least 4K rows); 512^2 and 640x360 a strip per row
// (0.74 and 0.76 us against 1.78 and 1.80).  Results are kept per field
// shape and depth.
int auto_strips(int W, int H, int depth) {
    static int simds = 0;
    if (!simds) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            simds = 4 * cus;
        else
            simds = 1024;
    }
    thread_local int last_w = -1, last_h = -1, last_k = -1, last_s = 1;
    if (W == last_w && H == last_h && depth == last_k) return last_s;
    constexpr double kIssue[4] = {0.0, 0.55, 0.85, 1.0};
    const Plan one = make_plan(W, H, depth, 1);
    // waves per strip, for the first guess
    const double per_strip = one.full + (one.rem ? 1.0 / (one.group > 1 ? one.group : 1) : 0.0);
    // the most groups whose strips can have 2K rows
    const int max_groups = H / (2 * depth * kWavesPerGroup - depth + 1);
    long long best = 1;
    double best_time = 0.0;
    bool first = true;
    for (int w = 1; w <= 3; ++w) {
        const long long target = static_cast<long long>(w) * simds;
        long long strips = static_cast<long long>(target / per_strip) + 1;
        if (strips > H) strips = H;
        while (strips > 1 && make_plan(W, H, depth, static_cast<int>(strips)).waves > target)
            --strips;
        long long cands[2] = {strips, 0};
        long long xs = strips / kWavesPerGroup * kWavesPerGroup;
        if (xs > static_cast<long long>(max_groups) * kWavesPerGroup)
            xs = static_cast<long long>(max_groups) * kWavesPerGroup;
        for (; xs >= kWavesPerGroup; xs -= kWavesPerGroup) {
            const Plan p = make_plan(W, H, depth, static_cast<int>(xs));
            if (p.xgroups > 0 && p.waves <= target) {
                cands[1] = xs;
                break;
            }
        }
        for (long long c : cands) {
            if (c < 1) continue;
            const Plan p = make_plan(W, H, depth, static_cast<int>(c));
            double per_simd = static_cast<double>(p.waves) / simds;
            if (per_simd < 1.0) per_simd = 1.0;
            // the issue share of the waves a SIMD actually holds
            long long held = (p.waves + simds - 1) / simds;
            if (held > w) held = w;
            if (held < 1) held = 1;
            const double time = per_simd * wave_cost(p, H, depth) / kIssue[held];
            if (first || time < best_time) {
                best = c;
                best_time = time;
                first = false;
            }
        }
    }
    last_w = W;
    last_h = H;
    last_k = depth;
    last_s = static_cast<int>(best);
    return last_s;
}

}  // namespace

extern "C" {

int xen_life_bits_launch(const void *src, void *dst, int width_words, int height,
                         uint32_t born, uint32_t sustain, int depth, int strips,
                         xen_stream stream) {
    if (width_words < 1 || height < 1)
        return xen::fail("xen_life_bits_launch: empty field", 22);
    if (src == dst) return xen::fail("xen_life_bits_launch: src and dst must differ", 22);
    if (static_cast<long long>(width_words) * height * 4 >= (1LL << 31))
        return xen::fail("xen_life_bits_launch: field must be smaller than 2 GiB", 22);
    if (depth < 1 || depth > 8) return xen::fail("xen_life_bits_launch: depth must be 1..8", 22);
    int rows = strips > 0 ? strips : auto_strips(width_words, height, depth);
    if (rows > height) rows = height;
    const Plan plan = make_plan(width_words, height, depth, rows);
    auto s = reinterpret_cast<hipStream_t>(stream);
    auto in = static_cast<const uint32_t *>(src);
    auto out = static_cast<uint32_t *>(dst);
    const bool fast = (born & 0x1FFu) == 0x8u && (sustain & 0x1FFu) == 0xCu;
    if (fast)
        return launch_any_k<true>(depth, in, out, width_words, height, plan, born, sustain, s);
    return launch_any_k<false>(depth, in, out, width_words, height, plan, born, sustain, s);
}

int xen_life_bits_run(void *buf_a, void *buf_b, int width_words, int height,
                      uint32_t born, uint32_t sustain, long long steps, int depth,
                      int strips, int *result_in_b, xen_stream stream) {
    if (depth < 1 || depth > 8) return xen::fail("xen_life_bits_run: depth must be 1..8", 22);
    void *cur = buf_a, *nxt = buf_b;
    int in_b = 0;
    long long left = steps;
    while (left > 0) {
        const int k = left >= depth ? depth : static_cast<int>(left);
        if (int rc = xen_life_bits_launch(cur, nxt, width_words, height, born, sustain, k,
                                          strips, stream))
            return rc;
        void *t = cur;
        cur = nxt;
        nxt = t;
        in_b ^= 1;
        left -= k;
    }
    *result_in_b = in_b;
    return 0;
}

}  // extern "C"
