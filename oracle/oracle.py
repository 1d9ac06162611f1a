"""
CPU oracle of the reference's semantics -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker; the
product never does.  It restates, independently of the product code:

* the seed streams: ``LocalRandom`` = stdlib ``random.Random(word)``
  seeding a NumPy legacy ``RandomState`` with its first 32 random bits
  (``xentica/seeds/random.py:40-44``); ``RandInt`` = ``randint(lo, hi+1,
  count)`` (``random.py:81-96``); ``BigBang`` area clamping and flat-order
  fill (``xentica/seeds/patterns.py:187-235``); ``PrimordialSoup``
  (``patterns.py:262-274``); packing without masking
  (``xentica/core/base.py:510-524``);
* the neighbourhood offset lists and slot pairing
  (``xentica/core/topology/neighborhood.py:159-190``), including the
  reference's Moore quirk in 1-D and in 3-D and up (centre kept, last
  offset dropped);
* the step itself in C (``oracle/bsca_oracle.c``), in the reference's
  plain emit/absorb structure with buffer planes.

Pinning (what checks each part; see ``tests/test_oracle.py``):
  seeds      -- ``tests/seeds/test_random.py:34,42,71`` vectors and the
                1648433356 seed CRC (``tests/seeds/test_patterns.py:155-164``)
  lifelike   -- golden CRCs 2981695958 (Life), 1098273940 (static border),
                2742543959 (6-D von Neumann), 4016906756 (interactive rule),
                plus an independent NumPy roll-sum Life and glider/blinker
                closed forms
  color_life -- golden CRC 532957133
  sands      -- golden CRC 1639048270 and the closed form (roll by -1,-1)
  noisetv    -- golden CRC 2773894957 and the closed form (LCG twice)
  crc32      -- ``zlib.crc32``
  energy_life -- no reference vector (the model has no reference
                counterpart): pinned by exact energy conservation, the
                min/max principle of its flux, single-spike closed forms,
                reduction to B3/S23 Life (uniform genome, zero thresholds,
                no mutation) checked against the NumPy roll-sum Life, and
                the xorshift32 period 2**32 - 1 (Marsaglia 2003)
  3-D Moore quirk -- parity unpinned by any reference vector; pinned by
                the offset list read from ``neighborhood.py:165-167`` and a
                brute-force 3-D check in the tests.
"""
import ctypes
import itertools
import os
import random
import subprocess

import numpy as np

__all__ = [
    'library', 'local_random', 'moore_deltas', 'von_neumann_deltas',
    'moore_rev', 'von_neumann_rev', 'big_bang', 'primordial_soup',
    'pack', 'Field', 'lifelike', 'color_life', 'shifting_sands', 'noisetv',
    'crc32', 'golly2int', 'numpy_life_step', 'WIDE', 'ENERGY_BITS',
    'energy_life',
    'energy_life_field', 'xorshift32_period', 'energy_fields',
]

#: 128-bit little-endian word (low 64 bits first).
WIDE = np.dtype([('lo', '<u8'), ('hi', '<u8')])

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def library():
    """Load ``liboracle.so``, building it with make when it is missing."""
    global _LIB  # pylint: disable=global-statement
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        lib = ctypes.CDLL(path)
        p = ctypes.c_void_p
        i64, i32, u32 = ctypes.c_int64, ctypes.c_int, ctypes.c_uint32
        lib.oracle_lifelike.argtypes = [
            p, p, i32, p, i32, p, p, i32, i64, u32, u32, i32, i32, i32,
            i32, i32, p]
        lib.oracle_color_life.argtypes = [
            p, p, i64, i64, p, p, i32, i32, i32, i32, i32]
        lib.oracle_shifting_sands.argtypes = [
            p, p, i64, i64, p, p, i32, i32, i32, i32, i32]
        lib.oracle_noisetv.argtypes = [p, p, i64, i32, i32, i32, i32, i32]
        lib.oracle_energy_life.argtypes = [
            p, i32, p, i64, i64, p, p, u32, u32, u32, i32, i32, i32, i32,
            i32, p]
        lib.oracle_energy_life.restype = i32
        lib.oracle_xorshift32_period.argtypes = [u32]
        lib.oracle_xorshift32_period.restype = ctypes.c_uint64
        lib.oracle_crc32.argtypes = [u32, p, i64]
        lib.oracle_crc32.restype = u32
        for name in ('oracle_lifelike', 'oracle_color_life',
                     'oracle_shifting_sands', 'oracle_noisetv'):
            getattr(lib, name).restype = i32
        _LIB = lib
    return _LIB


def _ptr(arr):
    return None if arr is None else arr.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- seeds

def local_random(word):
    """``(random.Random(word), RandomState(first 32 bits))``."""
    std = random.Random(word)
    return std, np.random.RandomState(std.getrandbits(32))


def golly2int(rule):
    """``"B3/S23"`` -> birth bits 0-8 + survival bits 9-17 (sum of bits)."""
    digits = "".join(ch for ch in rule if ch.isdigit() or ch == "/")
    born, sus = digits.split("/", 1)
    return sum(1 << (int(d) + 9) for d in sus) + sum(1 << int(d) for d in born)


def pack(values, widths):
    """Sum ``value << shift`` over fields in declaration order (no mask)."""
    val, shift = 0, 0
    for value, width in zip(values, widths):
        if value is not None:
            val = val + (value << shift)
        shift += width
    return val


def primordial_soup(size, draws, rng):
    """
    Draw every field for the whole board.

    ``draws`` maps field name -> callable(rng, count) returning values (or
    a constant); fields are drawn in sorted name order.
    """
    count = int(np.prod(size))
    return {name: draws[name](rng, count) if callable(draws[name])
            else draws[name] for name in sorted(draws)}, count


def big_bang(size, pos, area, draws, rng):
    """Clamp the area, return ``(flat indices, drawn fields)``."""
    pos = list(pos)
    for axis in range(len(size)):
        if pos[axis] + area[axis] >= size[axis]:
            pos[axis] = size[axis] - area[axis]
        pos[axis] = max(0, pos[axis])
    grids = np.meshgrid(*[np.arange(p, p + a) for p, a in zip(pos, area)],
                        indexing='ij')
    flat = np.zeros(grids[0].shape, dtype=np.int64)
    stride = 1
    for axis in range(len(size)):
        flat += grids[axis] * stride
        stride *= size[axis]
    indices = np.sort(flat.ravel())
    count = int(np.prod(area))
    values = {name: draws[name](rng, count) if callable(draws[name])
              else draws[name] for name in sorted(draws)}
    return indices, values


def randint(lo, hi):
    """A ``RandInt(lo, hi)`` draw on the NumPy stream."""
    return lambda rng, count: rng[1].randint(lo, hi + 1, count)


# ---------------------------------------------------------------- topology

def moore_deltas(ndim):
    """Offsets exactly as the reference lists them (quirk included)."""
    deltas = [d for d in itertools.product([-1, 0, 1], repeat=ndim)
              if d != (0, 0)]
    return deltas[:3 ** ndim - 1]


def moore_rev(ndim):
    """``rev_index_map`` of the Moore neighbourhood."""
    return list(range(3 ** ndim - 1))[::-1]


def von_neumann_deltas(ndim):
    """``+e0, -e0, +e1, -e1, ...``."""
    out = []
    for axis in range(ndim):
        for sign in (1, -1):
            out.append(tuple(sign if a == axis else 0 for a in range(ndim)))
    return out


def von_neumann_rev(ndim):
    """Pairs ``2i <-> 2i+1``."""
    rev = []
    for axis in range(ndim):
        rev += [2 * axis + 1, 2 * axis]
    return rev


# ---------------------------------------------------------------- steps

class Field:
    """A field in the reference's plain layout: ``n*(k+1)+1`` words."""

    def __init__(self, size, k, dtype, state):
        self.size = tuple(int(s) for s in size)
        self.n = int(np.prod(self.size))
        self.k = k
        self.fld = np.zeros(self.n * (k + 1) + 1, dtype=dtype)
        self.fld[:self.n] = state
        self.col = np.zeros(3 * self.n, dtype=np.int32)

    @property
    def cells(self):
        """Plane 0 (the main state)."""
        return self.fld[:self.n]


def lifelike(field, deltas, rev, born, sustain, steps, border_static=False,
             border_value=0, fade=(255, 255, 1), colors=True, threads=0,
             hashes=False):
    """Life-like steps in place; returns per-step CRCs when asked."""
    ndim = len(field.size)
    sizes = np.array(field.size, dtype=np.int64)
    dl = np.array(deltas, dtype=np.int8).reshape(-1)
    rv = np.array(rev, dtype=np.int32)
    out = np.zeros(max(steps, 1), dtype=np.uint32) if hashes else None
    status = library().oracle_lifelike(
        _ptr(field.fld), _ptr(field.col) if colors else None, ndim,
        _ptr(sizes), len(rev), _ptr(dl), _ptr(rv), int(border_static),
        int(border_value), born & 0xFFFFFFFF, sustain & 0xFFFFFFFF,
        fade[0], fade[1], fade[2], steps, threads, _ptr(out))
    assert status == 0
    return out[:steps] if hashes else None


def color_life(field, steps, fade=(255, 10, 1), threads=0):
    """GameOfLifeColor steps in place."""
    dl = np.array(moore_deltas(2), dtype=np.int8).reshape(-1)
    rv = np.array(moore_rev(2), dtype=np.int32)
    assert library().oracle_color_life(
        _ptr(field.fld), _ptr(field.col), field.size[0], field.size[1],
        _ptr(dl), _ptr(rv), fade[0], fade[1], fade[2], steps, threads) == 0


def shifting_sands(field, steps, fade=(255, 255, 1), threads=0):
    """ShiftingSands steps in place."""
    dl = np.array(moore_deltas(2), dtype=np.int8).reshape(-1)
    rv = np.array(moore_rev(2), dtype=np.int32)
    assert library().oracle_shifting_sands(
        _ptr(field.fld), _ptr(field.col), field.size[0], field.size[1],
        _ptr(dl), _ptr(rv), fade[0], fade[1], fade[2], steps, threads) == 0


def noisetv(field, steps, fade=(255, 255, 1), threads=0):
    """NoiseTV steps in place."""
    assert library().oracle_noisetv(
        _ptr(field.fld), _ptr(field.col), field.n, fade[0], fade[1],
        fade[2], steps, threads) == 0


def energy_life(field, steps, sustain=32, birth=128, mutation=16,
                fade=(255, 255, 1), colors=True, threads=0, hashes=False):
    """EnergyLife steps in place on a field of 64-bit words (the model's
    layout) or :data:`WIDE` words (its 128-bit variant)."""
    dl = np.array(moore_deltas(2), dtype=np.int8).reshape(-1)
    rv = np.array(moore_rev(2), dtype=np.int32)
    out = np.zeros(max(steps, 1), dtype=np.uint32) if hashes else None
    wide = field.fld.dtype == WIDE
    assert wide or field.fld.dtype == np.uint64
    assert library().oracle_energy_life(
        _ptr(field.fld), int(wide), _ptr(field.col) if colors else None,
        field.size[0], field.size[1], _ptr(dl), _ptr(rv), sustain, birth,
        mutation, fade[0], fade[1], fade[2], steps, threads, _ptr(out)) == 0
    return out[:steps] if hashes else None


def xorshift32_period(seed):
    """Steps until xorshift32 (13, 17, 5) returns to ``seed`` (0: never)."""
    return int(library().oracle_xorshift32_period(seed))


# EnergyLife's field layouts (examples/energy_life.py): energy width;
# genome (18 bits), mode (1) and the xorshift32 state (32) follow it
ENERGY_BITS = {False: 8, True: 16}


def energy_fields(words):
    """``(energy, genome, mode, rng)`` arrays of EnergyLife words (uint64,
    or :data:`WIDE` for the 128-bit variant)."""
    wide = words.dtype == WIDE
    eb = ENERGY_BITS[wide]
    if wide:
        lo = words['lo'].astype(np.uint64)
        hi = words['hi'].astype(np.uint64)
        rng = ((lo >> np.uint64(eb + 19)) | (hi << np.uint64(64 - eb - 19)))
    else:
        lo = words.astype(np.uint64)
        rng = lo >> np.uint64(eb + 19)
    energy = (lo & np.uint64((1 << eb) - 1)).astype(np.int64)
    genome = ((lo >> np.uint64(eb)) & np.uint64(0x3FFFF)).astype(np.int64)
    mode = ((lo >> np.uint64(eb + 18)) & np.uint64(1)).astype(np.int64)
    rng = (rng & np.uint64(0xFFFFFFFF)).astype(np.int64)
    return energy, genome, mode, rng


def energy_words(energy, genome, mode, rng, wide=False):
    """Pack non-negative in-range fields into EnergyLife words."""
    energy, genome, mode, rng = (np.asarray(v, dtype=np.uint64)
                                 for v in (energy, genome, mode, rng))
    eb = ENERGY_BITS[wide]
    lo = (energy | (genome << np.uint64(eb)) | (mode << np.uint64(eb + 18))
          | (rng << np.uint64(eb + 19)))
    if not wide:
        return lo
    out = np.zeros(energy.shape, dtype=WIDE)
    out['lo'] = lo
    out['hi'] = rng >> np.uint64(64 - eb - 19)
    return out


def energy_life_field(size, word=42, wide=False):
    """
    Step-0 state of ``EnergyLifeExperiment`` (``wide``: of the 128-bit
    variant): ``PrimordialSoup`` drawing, in sorted name order, energy
    ``RandInt(0, 255)``, genome ``RandInt(0, 2**18 - 1)``, mode
    ``RandInt(0, 1)`` and generator ``RandInt(1, 2**32 - 1)``.
    """
    rng = local_random(word)
    vals, _ = primordial_soup(size, {
        "energy": randint(0, 255), "genome": randint(0, 2 ** 18 - 1),
        "mode": randint(0, 1), "rng": randint(1, 2 ** 32 - 1)}, rng)
    words = energy_words(vals["energy"], vals["genome"], vals["mode"],
                         vals["rng"], wide)
    return Field(size, 8, WIDE if wide else np.uint64, words)


def crc32(data):
    """CRC-32 of a byte buffer (the oracle's own table-driven C code)."""
    data = np.ascontiguousarray(data).view(np.uint8)
    return int(library().oracle_crc32(0, _ptr(data), data.size))


def numpy_life_step(state, born=8, sustain=12):
    """
    Independent 2-D torus Life-like step by neighbour roll-sum.

    Uses ``state[x1, x0]`` layout of a ``(w1, w0)`` array; an independent
    check of the C oracle, not a restatement of the reference's layout.
    """
    count = np.zeros(state.shape, dtype=np.int64)
    for d1 in (-1, 0, 1):
        for d0 in (-1, 0, 1):
            if d0 or d1:
                count += np.roll(np.roll(state, -d1, axis=0), -d0, axis=1)
    is_born = (born >> count) & 1
    is_sus = (sustain >> count) & 1
    return ((is_born | (is_sus & state)) & 1).astype(state.dtype)


# ------------------------------------------------ the reference's experiments

def gol_experiment_field(k=8):
    """
    Step-0 state of ``GOLExperiment`` (``examples/game_of_life.py:206-232``).

    640x360, BigBang at (320, 180) of 100x100, ``RandInt(0, 1)``, word
    ``"OBEY XENTICA"``.
    """
    rng = local_random("OBEY XENTICA")
    size = (640, 360)
    idx, vals = big_bang(size, (320, 180), (100, 100),
                         {"state": randint(0, 1)}, rng)
    state = np.zeros(int(np.prod(size)), dtype=np.int64)
    state[idx] = pack([vals["state"]], [1])
    field = Field(size, k, np.uint8, state.astype(np.uint8))
    return field


def soup_field(size, word, draws, widths, names, k, dtype):
    """Step-0 state of a ``PrimordialSoup`` experiment."""
    rng = local_random(word)
    vals, _ = primordial_soup(size, draws, rng)
    packed = pack([vals.get(name) for name in names], widths)
    state = np.zeros(int(np.prod(size)), dtype=np.int64) + packed
    return Field(size, k, dtype, state.astype(dtype))
