// Truth tables of gfx950's three-input logic op (v_bitop3_b32), computed
// from the boolean function so that no immediate is written by hand.
// Result bit = table[(a << 2) | (b << 1) | c], i.e. the immediate is the
// function evaluated on a = 0xF0, b = 0xCC, c = 0xAA -- the operand order
// hipcc itself uses when it selects the instruction.
#pragma once
#include <stdint.h>

template <typename F>
constexpr uint8_t bitop3_table(F f) {
    uint8_t table = 0;
    for (int i = 0; i < 8; ++i)
        if (f((i >> 2) & 1, (i >> 1) & 1, i & 1)) table |= static_cast<uint8_t>(1u << i);
    return table;
}

static_assert(bitop3_table([](bool a, bool b, bool c) { return a ^ b ^ c; }) == 0x96, "xor3");
static_assert(bitop3_table([](bool a, bool, bool) { return a; }) == 0xF0, "operand a");
static_assert(bitop3_table([](bool, bool b, bool) { return b; }) == 0xCC, "operand b");
static_assert(bitop3_table([](bool, bool, bool c) { return c; }) == 0xAA, "operand c");
