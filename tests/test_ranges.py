"""
The interval analysis behind the fused kernel's colour plan.

* Soundness against a C compiler: random expressions over variables of
  every integer type, each variable spanning a contiguous interval of 16
  values (some at its type's limits), are compiled with gcc and evaluated
  on every combination; the analysis' interval must contain every result.
* Closed forms: single operations whose C result is fixed by the standard
  (promotions, truncating division and remainder, shifts, conversions).
* Plans of the example models: which colour effects may skip the colour of
  a launch's earlier steps, and with which stored-colour interval.
"""
import os
import random
from decimal import Decimal
import shutil
import subprocess

import numpy as np
import pytest

from xentica_amd.runtime import ranges
from xentica_amd.runtime.ranges import Value

_T = {name: ranges.type_named(name) for name in (
    "unsigned char", "signed char", "short", "unsigned short", "int",
    "unsigned int", "long", "unsigned long", "float")}


def _ev(text, **names):
    return ranges.Analysis(names).evaluate(text)


def _iv(value):
    return (value.lo, value.hi)


@pytest.mark.parametrize("text, names, want", [
    ("(x & 7)", {"x": Value.full(_T["int"])}, (0, 7)),
    ("(x % 8)", {"x": Value.full(_T["int"])}, (-7, 7)),
    ("(x % 8)", {"x": Value.full(_T["unsigned int"])}, (0, 7)),
    ("(x % 8)", {"x": Value(_T["int"], 0, 5)}, (0, 5)),
    ("(-7 / 2)", {}, (-3, -3)),
    ("(-7 % 3)", {}, (-1, -1)),
    ("(-5 >> 1)", {}, (-3, -3)),
    ("(1u << 31)", {}, (1 << 31, 1 << 31)),
    ("(1 << 31)", {}, (-(1 << 31), (1 << 31) - 1)),
    ("(x - 1)", {"x": Value(_T["unsigned int"], 0, 3)}, (0, (1 << 32) - 1)),
    ("(x - 1)", {"x": Value(_T["unsigned char"], 0, 3)}, (-1, 2)),
    ("((unsigned char) (x))", {"x": Value(_T["int"], 250, 260)}, (0, 255)),
    ("((unsigned char) (x))", {"x": Value(_T["int"], 5, 9)}, (5, 9)),
    ("(((1) < (x)) ? (1) : (x))", {"x": Value(_T["int"], -4, 9)}, (-4, 9)),
    ("(x | 1)", {"x": Value(_T["int"], 4, 6)}, (4, 7)),
    ("(~x)", {"x": Value(_T["unsigned char"], 0, 255)}, (-256, -1)),
    ("(x * 255)", {"x": Value(_T["unsigned char"], 0, 1)}, (0, 255)),
    ("((unsigned int) ((x * 2.5f)))", {"x": Value(_T["int"], 0, 4)},
     (0, 10)),
    ("(0xFFFFFFFF + 1)", {}, (0, (1 << 32) - 1)),
    ("(4294967295 + 1)", {}, (1 << 32, 1 << 32)),
    ("(010 + 0)", {}, (8, 8)),
])
def test_closed_forms(text, names, want):
    got = _ev(text, **names)
    assert got.lo <= want[0] and want[1] <= got.hi
    if want[0] == want[1] or text.startswith("(x"):
        assert _iv(got) == want, text


def test_unknown_constructs():
    with pytest.raises(ranges.Unknown):
        _ev("(* reinterpret_cast<float*>(&(t)))", t=Value.full(_T["int"]))
    with pytest.raises(ranges.Unknown):
        _ev("(y + 1)")
    analysis = ranges.Analysis({})
    analysis.run("int a = 3;\nint b = a * 2;\nb += 1;\n"
                 "unsigned char c = 300;\nfloat f = 0.0f;\n"
                 "f = (* reinterpret_cast<float*>(&(a)));\n")
    assert _iv(analysis.env["b"]) == (7, 7)
    assert _iv(analysis.env["c"]) == (0, 255)
    assert analysis.env["f"].hi == float("inf")
    with pytest.raises(ranges.GiveUp):
        ranges.Analysis({}).run("int a = 0;\nif (a) { a = 1; }\n")
    with pytest.raises(ranges.GiveUp):
        ranges.Analysis({}).run("a = 1;\n")


_LIMITS = {
    "unsigned char": (0, 255), "signed char": (-128, 127),
    "short": (-32768, 32767), "unsigned short": (0, 65535),
    "int": (-2 ** 31, 2 ** 31 - 1), "unsigned int": (0, 2 ** 32 - 1),
    "long": (-2 ** 63, 2 ** 63 - 1), "unsigned long": (0, 2 ** 64 - 1),
}


def _interval(rng, tname):
    lo, hi = _LIMITS[tname]
    start = rng.choice([lo, hi - 15, 0, -8, 100, 240, 65530])
    start = min(max(start, lo), hi - 15)
    return start, start + 15


_CONSTS = ["0", "1", "2", "7", "255", "3", "65535", "4294967295u", "1ull",
           "2147483647", "0x80000000", "9"]
_CASTS = ["unsigned char", "signed char", "short", "unsigned short", "int",
          "unsigned int", "long", "unsigned long"]


def _expr(rng, depth):
    if depth == 0 or rng.random() < 0.2:
        return rng.choice(["x", "y", "z", "x", "y", rng.choice(_CONSTS)])
    pick = rng.random()
    if pick < 0.12:
        return "(%s(%s))" % (rng.choice(["-", "~", "!"]), _expr(rng, depth - 1))
    if pick < 0.24:
        return "((%s) (%s))" % (rng.choice(_CASTS), _expr(rng, depth - 1))
    if pick < 0.32:
        return "((%s) ? (%s) : (%s))" % (_expr(rng, depth - 1),
                                         _expr(rng, depth - 1),
                                         _expr(rng, depth - 1))
    if pick < 0.42:
        return "(%s %s %s)" % (_expr(rng, depth - 1), rng.choice(["/", "%"]),
                               rng.choice(["1", "2", "3", "7", "16", "255",
                                           "-2", "-7", "3u"]))
    if pick < 0.52:
        return "(%s %s %d)" % (_expr(rng, depth - 1), rng.choice(["<<", ">>"]),
                               rng.randint(0, 7))
    if pick < 0.58:
        return "((unsigned int) ((%s * 0.5f)))" % _expr(rng, depth - 1)
    if pick < 0.64:
        return "(((float) (%s)) / %s)" % (
            _expr(rng, depth - 1),
            rng.choice(["65535", "360", "3", "7.0f", "((float) (18))"]))
    op = rng.choice(["+", "-", "*", "&", "|", "^", "<", "==", ">=", "&&",
                     "||"])
    return "(%s %s %s)" % (_expr(rng, depth - 1), op, _expr(rng, depth - 1))


def _c_int128(value):
    if value < 0:
        return "(-(__int128) %dULL)" % (-value)
    return "((__int128) %dULL)" % value


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_sound_against_gcc(tmp_path):
    rng = random.Random(20251203)
    cases = []
    while len(cases) < 300:
        types = [rng.choice(list(_LIMITS)) for _ in range(3)]
        spans = [_interval(rng, t) for t in types]
        text = _expr(rng, 3)
        names = {n: Value(_T[t], lo, hi)
                 for n, t, (lo, hi) in zip("xyz", types, spans)}
        try:
            got = ranges.Analysis(names).evaluate(text)
        except ranges.Unknown:
            continue
        cases.append((types, spans, text, got))
    body = ["#include <stdio.h>", "#include <math.h>", "int main(void) {"]
    for idx, (types, spans, text, _) in enumerate(cases):
        (tx, ty, tz), ((xl, xh), (yl, yh), (zl, zh)) = types, spans
        body.append(
            "{ long double mn = INFINITY, mx = -INFINITY; int nan = 0;\n"
            "for (__int128 a = %s; a <= %s; ++a)\n"
            "for (__int128 b = %s; b <= %s; ++b)\n"
            "for (__int128 c = %s; c <= %s; ++c) {\n"
            "  %s x = (%s) a; %s y = (%s) b; %s z = (%s) c;\n"
            "  long double v = (long double) (%s);\n"
            "  if (v != v) nan = 1; else { if (v < mn) mn = v; "
            "if (v > mx) mx = v; }\n"
            "  (void) x; (void) y; (void) z; }\n"
            "printf(\"%d %%d %%.21Le %%.21Le\\n\", nan, mn, mx); }"
            % (_c_int128(xl), _c_int128(xh), _c_int128(yl), _c_int128(yh),
               _c_int128(zl), _c_int128(zh), tx, tx, ty, ty, tz, tz, text,
               idx))
    body.append("return 0; }")
    src = tmp_path / "fuzz.c"
    src.write_text("\n".join(body))
    exe = tmp_path / "fuzz"
    subprocess.run(["gcc", "-O0", "-fwrapv", "-w", "-o", str(exe), str(src)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True,
                         text=True, env=dict(os.environ)).stdout.split("\n")
    checked = 0
    for line in out:
        if not line:
            continue
        idx, nan, mn, mx = line.split()
        types, spans, text, got = cases[int(idx)]
        if int(nan):
            assert got.ctype.floating and got.hi == float("inf"), text
            continue
        # 22 significant digits print every long double integer exactly
        low, high = Decimal(mn), Decimal(mx)
        assert Decimal(got.lo) <= low and high <= Decimal(got.hi), \
            (text, types, spans, got, mn, mx)
        checked += 1
    assert checked == len(cases) == 300


def test_color_plans_of_examples():
    from examples.game_of_life import (GameOfLife, GameOfLifeColor,
                                       GOLExperiment, GOLExperimentColor)
    from examples.noisetv import NoiseTV, NoiseTVExperiment
    from examples.sand_heights import SandHeights, SandHeightsExperiment
    from examples.workloads import LifeLikeRule, batch_experiments
    from xentica_amd.runtime.fused import FusedProgram
    full = ((0, 255), ) * 3
    # state in {0, 1} times 255, fades 255: stored colours in [0, 255]
    assert FusedProgram(GameOfLife(GOLExperiment)).color_plan == full
    assert FusedProgram(
        LifeLikeRule(batch_experiments(count=1)[0])).color_plan == full
    # a fade of 10 below the channel's span of 255: every step
    assert FusedProgram(
        GameOfLifeColor(GOLExperimentColor)).color_plan is None
    # heights 0..15 times 255 span 3825 > 255: every step
    assert FusedProgram(
        SandHeights(SandHeightsExperiment)).color_plan is None
    # the colour advances the generator: every step, every cell
    assert FusedProgram(NoiseTV(NoiseTVExperiment)).color_plan is None


def test_plan_admits_and_bounds_after():
    from examples.game_of_life import GameOfLife, GOLExperiment
    from xentica_amd.runtime.fused import FusedProgram, _union
    program = FusedProgram(GameOfLife(GOLExperiment))
    assert program.color_values == ((0, 255), ) * 3
    assert program.admits(((0, 0), ) * 3)
    assert program.admits(((0, 255), ) * 3)
    assert not program.admits(None)
    assert not program.admits(((0, 256), (0, 0), (0, 0)))
    assert not program.admits(((0, 0), (-1, 0), (0, 0)))
    # a launch from admitted colours leaves the new values' bounds; from
    # unknown or outside colours, unknown bounds
    assert program.bounds_after(((0, 0), ) * 3, True) == ((0, 255), ) * 3
    assert program.bounds_after(((9, 300), ) * 3, False) is None
    assert program.bounds_after(None, False) is None
    assert _union([((0, 1), (2, 3), (4, 5)), ((-1, 0), (2, 9), (4, 4))]) \
        == ((-1, 1), (2, 9), (4, 5))
    assert _union([((0, 1), ) * 3, None]) is None


def test_colour_bounds_follow_host_writes():
    from examples.game_of_life import GameOfLife, GOLExperiment
    model = GameOfLife(GOLExperiment)
    colors = model.gpu.arrays.colors
    assert colors.bounds == ((0, 0), ) * 3
    data = np.zeros(model.cells_num * 3, dtype=np.int32)
    data[4] = 700
    data[3 * 17 + 2] = -3
    colors.set(data)
    assert colors.bounds == ((0, 0), (0, 700), (-3, 0))


def test_stores_only_into_arrays():
    word = ranges.type_named("unsigned char")
    analysis = ranges.Analysis({}, {"_mp": (word, 2)}, {})
    analysis.run("int a = 1;\n(_mp[_cur ^ 1])[_li] = a;\ncol[i] = a;\n")
    assert _iv(analysis.env["a"]) == (1, 1)
    # a store through a local's address would change the local
    for code in ("int a = 1;\n(* reinterpret_cast<float*>(&(a))) = 2.0f;\n",
                 "int a = 1;\nother[0] = a;\n",
                 "int a = 1;\n*(&a) = 5;\n"):
        with pytest.raises(ranges.GiveUp):
            ranges.Analysis({}, {"_mp": (word, 2)}, {}).run(code)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_selects_equal_products_against_gcc(tmp_path):
    """Expressions written out again with 0/1 products as selects give the
    original's value on every input (gcc evaluates both)."""
    rng = random.Random(777)
    cases = []
    while len(cases) < 300:
        types = [rng.choice(list(_LIMITS)) for _ in range(3)]
        spans = [rng.choice([(0, 1), (0, 1), (0, 2), _interval(rng, t),
                             _interval(rng, t)]) for t in types]
        text = _expr(rng, 3)
        if "*" not in text:
            continue
        names = {n: Value(_T[t], lo, hi)
                 for n, t, (lo, hi) in zip("xyz", types, spans)}
        try:
            _, out = ranges.Analysis(names).evaluate(text, selects=True)
        except ranges.Unknown:
            continue
        cases.append((types, spans, text, out))
    assert sum(" ? (" in out for *_, out in cases) > 100
    assert sum("_xen_fdiv(" in out for *_, out in cases) > 20
    body = ["#include <stdio.h>", "#include <math.h>",
            "#define __device__", "#define __forceinline__ static inline",
            ranges.FDIV_HELPER, "int main(void) {"]
    for idx, (types, spans, text, out) in enumerate(cases):
        (tx, ty, tz), ((xl, xh), (yl, yh), (zl, zh)) = types, spans
        body.append(
            "{ long bad = 0;\n"
            "for (__int128 a = %s; a <= %s; ++a)\n"
            "for (__int128 b = %s; b <= %s; ++b)\n"
            "for (__int128 c = %s; c <= %s; ++c) {\n"
            "  %s x = (%s) a; %s y = (%s) b; %s z = (%s) c;\n"
            "  long double u = (long double) (%s);\n"
            "  long double v = (long double) (%s);\n"
            "  if (u != v && (u == u || v == v)) ++bad;\n"
            "  (void) x; (void) y; (void) z; }\n"
            "printf(\"%d %%ld\\n\", bad); }"
            % (_c_int128(xl), _c_int128(xh), _c_int128(yl), _c_int128(yh),
               _c_int128(zl), _c_int128(zh), tx, tx, ty, ty, tz, tz, text,
               out, idx))
    body.append("return 0; }")
    src = tmp_path / "selects.c"
    src.write_text("\n".join(body))
    exe = tmp_path / "selects"
    subprocess.run(["gcc", "-O0", "-fwrapv", "-w", "-o", str(exe), str(src),
                    "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True,
                         text=True).stdout.split()
    results = dict(zip(out[::2], out[1::2]))
    assert len(results) == len(cases)
    bad = {idx: cases[int(idx)][2:] for idx, count in results.items()
           if count != "0"}
    assert not bad


def test_rewrite_statements():
    analysis = ranges.Analysis({"rule": Value.full(_T["int"])})
    code = ("unsigned char m = 300 & 1;\nunsigned int g = 70000;\n"
            "unsigned int b = 0;\nb += ((m * (g > b)) * (g - b));\n"
            "float f = 0.0f;\nf = (m * 2.5f);\n")
    out, complete = analysis.rewrite(code)
    assert complete
    assert out == (
        "\nunsigned char m = (300 & 1);\nunsigned int g = 70000;\n"
        "unsigned int b = 0;\n"
        "b += ((((m ? (int) ((g > b)) : (int) 0)) ? (unsigned int) ((g - b))"
        " : (unsigned int) 0));\n"
        # float products stay products
        "float f = 0.0f;\nf = ((m * 2.5f));\n")
    out, complete = ranges.Analysis({}).rewrite(
        "int a = 1;\nif (a) { a = 2; }\nint b = (a * 3);\n")
    assert not complete and out.endswith("int b = (a * 3);\n")


def _rn32(value):
    """A Fraction rounded to the nearest float32 (ties to even)."""
    from fractions import Fraction
    if value == 0:
        return Fraction(0)
    sign = -1 if value < 0 else 1
    mag = abs(value)
    exp = mag.numerator.bit_length() - mag.denominator.bit_length()
    if Fraction(2) ** exp > mag:
        exp -= 1
    ulp = Fraction(2) ** (max(exp, -126) - 23)
    whole, rem = divmod((mag / ulp).numerator, (mag / ulp).denominator)
    twice, den = 2 * rem, (mag / ulp).denominator
    if twice > den or (twice == den and whole % 2):
        whole += 1
    return sign * whole * ulp


def _fdiv_exact_slow(c, lo, hi, recip):
    """The division check with exact rational arithmetic throughout."""
    from fractions import Fraction
    c = _rn32(Fraction(c))
    r = Fraction(recip)
    for x in range(lo, hi + 1):
        q = _rn32(x * r)
        got = _rn32(_rn32(x - q * c) * r + q)
        if got != _rn32(Fraction(x) / c):
            return False
    return True


@pytest.mark.parametrize("c, lo, hi", [(65535, 0, 4095), (360, 0, 4095),
                                       (3, -2048, 2047), (7, 0, 4095),
                                       (0.1, 0, 1023)])
def test_division_check_against_exact_arithmetic(c, lo, hi):
    """The vectorised check of _xen_fdiv agrees with exact rational
    arithmetic: it passes the rounded reciprocal (one fused correction
    absorbs even an ulp of error in it) and fails one 2^-12 off."""
    ok, recip = ranges.division_by_constant_exact(c, lo, hi)
    assert ok and _fdiv_exact_slow(c, lo, hi, recip)
    r32 = np.float32(1.0) / np.float32(c)
    off = float(r32 * np.float32(1 + 2.0 ** -12))
    bad, _ = ranges.division_by_constant_exact(c, lo, hi, recip=off)
    assert not bad and not _fdiv_exact_slow(c, lo, hi, off)


def test_division_emits_the_divisor_it_checked():
    """A divisor that float32 cannot hold is written into the kernel as the
    float32 value the exactness check used, not as its decimal."""
    flt = ranges.CTYPES[("float", )]
    left = ranges.Value(flt, 0, 4095, integral=True)
    c = 16777217.0  # 2^24 + 1 rounds to 2^24 in float32
    right = ranges.Value(flt, c, c)
    text = ranges._constant_division(left, "x", right, flt)
    assert text is not None
    divisor = text.split(",")[1].strip()
    assert np.float32(float(divisor.rstrip("f"))) == np.float32(c)
    assert float(divisor.rstrip("f")) == float(np.float32(c))
