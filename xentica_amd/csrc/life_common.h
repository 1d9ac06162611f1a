// Device helpers shared by the bit-sliced Life-like kernels (gfx950,
// wave64): DPP lane shifts, v_bitop3_b32 truth tables, the horizontal
// 3-sums of a row word, and the rule applied to the bit-sliced count.
//
// Bit layout: 32 cells per 32-bit word, cell x of a word in bit x; lane l
// of a wave holds one word column, so the cells left and right of a word
// come from lanes l - 1 and l + 1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bitop3.h"

namespace {

// Value of lane (i - 1) in lane i (DPP wave_shr:1).
__device__ __forceinline__ uint32_t from_left_lane(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x138, 0xF, 0xF, true));
}

// Value of lane (i + 1) in lane i (DPP wave_shl:1).
__device__ __forceinline__ uint32_t from_right_lane(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x130, 0xF, 0xF, true));
}

// v_bitop3_b32: result bit = imm[(a << 2) | (b << 1) | c], i.e. the
// immediate is the function evaluated on a = 0xF0, b = 0xCC, c = 0xAA
// (the convention hipcc itself uses when it selects the instruction).
template <uint8_t kTable>
__device__ __forceinline__ uint32_t bitop3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, kTable);
}
// s_waitcnt immediate of gfx9-family parts: vmcnt 0, expcnt and lgkmcnt at
// their maxima (not waited on)
constexpr int kVmcntZero = 0x0F70;
constexpr uint8_t kXor3 = bitop3_table([](bool a, bool b, bool c) { return a ^ b ^ c; });
constexpr uint8_t kMaj3 = bitop3_table([](bool a, bool b, bool c) { return (a + b + c) >= 2; });
constexpr uint8_t kMux = bitop3_table([](bool a, bool b, bool c) { return a ? b : c; });
// count9 = t0 + 2W, W = u + k + 2v: e1 = (W == 1), e2 = (W == 2)
constexpr uint8_t kW1 = bitop3_table([](bool u, bool k, bool v) { return (u != k) && !v; });
constexpr uint8_t kW2 = bitop3_table(
    [](bool u, bool k, bool v) { return (u && k && !v) || (!u && !k && v); });
constexpr uint8_t kE2AliveNotT0 = bitop3_table([](bool e2, bool m, bool t0) { return e2 && m && !t0; });
constexpr uint8_t kAndOr2 = bitop3_table([](bool t0, bool e1, bool f) { return (t0 && e1) || f; });
constexpr uint8_t kXorAnd = bitop3_table([](bool v, bool u, bool k) { return v != (u && k); });
constexpr uint8_t kAnd3 = bitop3_table([](bool v, bool u, bool k) { return v && u && k; });

// One step down the wave's user priority (s_setprio takes an immediate),
// from 3 to 0 and no further.  A SIMD issues from its oldest ready wave
// first, so of several waves with equal work the oldest finishes long
// before the youngest, which then runs alone and cannot keep the VALU
// busy; kernels start every wave at priority 3 and lower it as the wave
// progresses, so a wave that is ahead yields to the waves behind it.
__device__ __forceinline__ void lower_priority(int &prio) {
    --prio;
    if (prio == 2)
        __builtin_amdgcn_s_setprio(2);
    else if (prio == 1)
        __builtin_amdgcn_s_setprio(1);
    else if (prio == 0)
        __builtin_amdgcn_s_setprio(0);
}

// A row word with its horizontal 3-sums: h0 + 2 h1 = (x-1) + x + (x+1).
struct HRow {
    uint32_t raw, h0, h1;
};

__device__ __forceinline__ HRow hrow(uint32_t w) {
    const uint32_t wl = from_left_lane(w), wr = from_right_lane(w);
    const uint32_t l = (w << 1) | (wl >> 31);  // cells at x - 1
    const uint32_t r = (w >> 1) | (wr << 31);  // cells at x + 1
    return {w, bitop3<kXor3>(l, w, r), bitop3<kMaj3>(l, w, r)};
}

// Rule application on the 3x3 block sum (the cell included):
// count9 = t0 + 2 (u + k + 2 v) from the three rows' horizontal sums, and
// the neighbour count is count9 - state.  kFast selects B3/S23, where
// next = (count9 == 3) | (state & count9 == 4).  Otherwise the rule masks
// are runtime values and the lookup is a mux tree over the bits of count9
// with the state muxed in at the leaves.
template <bool kFast>
struct Rule;

template <>
struct Rule<true> {
    uint32_t born, sustain;
    __device__ __forceinline__ uint32_t apply(uint32_t t0, uint32_t u, uint32_t k, uint32_t v,
                                              uint32_t m) const {
        const uint32_t e1 = bitop3<kW1>(u, k, v);
        const uint32_t e2 = bitop3<kW2>(u, k, v);
        return bitop3<kAndOr2>(t0, e1, bitop3<kE2AliveNotT0>(e2, m, t0));
    }
};

template <>
struct Rule<false> {
    uint32_t born, sustain;
    __device__ __forceinline__ uint32_t apply(uint32_t t0, uint32_t u, uint32_t k, uint32_t v,
                                              uint32_t m) const {
        const uint32_t w0 = u ^ k;
        const uint32_t w1 = bitop3<kXorAnd>(v, u, k);
        const uint32_t w2 = bitop3<kAnd3>(v, u, k);
        // leaf[c]: next state for count9 = c; a dead cell has c neighbours,
        // a live one c - 1 and is kept when that is a birth OR a survival
        // count (next = born | sustained & state).  count9 counts the cell,
        // so count9 = 0 only for a dead cell and count9 = 9 only for a live
        // one: those two leaves need no select.
        uint32_t leaf[10];
#pragma unroll
        for (int c = 0; c < 10; ++c) {
            const uint32_t d = (c <= 8 && ((born >> c) & 1u)) ? 0xFFFFFFFFu : 0u;
            const uint32_t a =
                (c >= 1 && (((born | sustain) >> (c - 1)) & 1u)) ? 0xFFFFFFFFu : 0u;
            leaf[c] = c == 0 ? d : c == 9 ? a : bitop3<kMux>(m, a, d);
        }
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

// New state of one word (32 cells) from the horizontal sums of its rows
// up (a), mid (b) and down (c).
template <bool kFast>
__device__ __forceinline__ uint32_t life_word(const HRow &a, const HRow &b, const HRow &c,
                                              const Rule<kFast> &rule) {
    const uint32_t t0 = bitop3<kXor3>(a.h0, b.h0, c.h0);
    const uint32_t k = bitop3<kMaj3>(a.h0, b.h0, c.h0);
    const uint32_t u = bitop3<kXor3>(a.h1, b.h1, c.h1);
    const uint32_t v = bitop3<kMaj3>(a.h1, b.h1, c.h1);
    return rule.apply(t0, u, k, v, b.raw);
}

// B3/S23 over rows taken in pairs.  Output rows 2q and 2q + 1 share the
// sum of the horizontal sums of rows 2q and 2q + 1,
//   P = p0 + 2 p1 + 4 p2  (0 .. 6),
// and add one more row's horizontal sum a0 + 2 a1: row 2q - 1 for output
// 2q, row 2q + 2 for output 2q + 1.  With count9 = a0 + 2 a1 + P (the cell
// included) the new state is (count9 == 3) | (m & count9 == 4), m the
// cell's state.  The four operations below compute it; they were found by
// an exhaustive search over circuits of three-input logic ops, and
// `pair_rule_holds` checks them against the rule for every input that can
// occur: P <= 6, and m = 1 implies P >= 1 (the pair holds the cell itself).
// A pair costs 4 ops and each output 4, so a row costs 6 + 2 + 4 = 12
// VALU ops with its horizontal sums against 6 + 8 = 14 for three rows.
constexpr uint8_t kOneOf3 = bitop3_table([](bool a, bool b, bool c) { return a + b + c == 1; });
constexpr uint8_t kTwoOf3 = bitop3_table([](bool a, bool b, bool c) { return a + b + c == 2; });
constexpr uint8_t kPairTwos =
    bitop3_table([](bool a1, bool p1, bool p2) { return (a1 != p1) || (a1 && p1 && p2); });
constexpr uint8_t kZNotBoth =
    bitop3_table([](bool p2, bool o2, bool z) { return z && !(p2 && o2); });

constexpr bool table_bit(uint8_t t, bool a, bool b, bool c) {
    return (t >> ((a << 2) | (b << 1) | c)) & 1;
}

constexpr bool pair_rule_holds() {
    for (int a = 0; a < 4; ++a)
        for (int P = 0; P <= 6; ++P)
            for (int m = 0; m < 2; ++m) {
                if (m && P == 0) continue;
                const bool a0 = a & 1, a1 = a >> 1, p0 = P & 1, p1 = (P >> 1) & 1, p2 = P >> 2;
                const bool o1 = table_bit(kOneOf3, a0, p0, m);
                const bool o2 = table_bit(kPairTwos, a1, p1, p2);
                const bool z = table_bit(kTwoOf3, m, o1, o2);
                const int c9 = a + P;
                if (table_bit(kZNotBoth, p2, o2, z) != (c9 == 3 || (m && c9 == 4))) return false;
            }
    return true;
}
static_assert(pair_rule_holds(), "paired B3/S23 circuit");

// p0 + 2 p1 + 4 p2 = (r0 + 2 r1) + (s0 + 2 s1): the pair sum of two rows'
// horizontal sums.
__device__ __forceinline__ void pair_sum(uint32_t r0, uint32_t r1, uint32_t s0, uint32_t s1,
                                         uint32_t &p0, uint32_t &p1, uint32_t &p2) {
    const uint32_t c = r0 & s0;
    p0 = r0 ^ s0;
    p1 = bitop3<kXor3>(r1, s1, c);
    p2 = bitop3<kMaj3>(r1, s1, c);
}

// New B3/S23 state from a row's horizontal sum (a0, a1), the pair sum
// (p0, p1, p2) and the cell's state m.
__device__ __forceinline__ uint32_t pair_life(uint32_t a0, uint32_t a1, uint32_t p0, uint32_t p1,
                                              uint32_t p2, uint32_t m) {
    const uint32_t o1 = bitop3<kOneOf3>(a0, p0, m);
    const uint32_t o2 = bitop3<kPairTwos>(a1, p1, p2);
    const uint32_t z = bitop3<kTwoOf3>(m, o1, o2);
    return bitop3<kZNotBoth>(p2, o2, z);
}

}  // namespace
