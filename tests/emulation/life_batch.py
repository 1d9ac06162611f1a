"""
CPU emulation of ``life_batch_kernel`` (``xentica_amd/csrc/life_batch.hip``).

Test infrastructure: it follows the kernel's per-lane set-up (the start
cell, word and bit of a lane's extended column, the words it loads and the
buffer offsets, past-the-end offsets for loads and stores a lane does not
make), its running per-lane rows, the K-level row pipeline with its
prologue, steady state by threes and remainder, and DPP lane shifts with a
zero at the wave's edges, with a wave's 64 lanes as NumPy vectors.  It
asserts that every load and store a lane makes is in bounds and that no
word is stored twice, so the kernel's plan and indexing can be checked on
a machine without a GPU.
"""
import numpy as np

from tests.emulation.life_bits import hrow, M32, maj

ONES = np.uint64(0xFFFFFFFF)


def rule_word(a, b, c, dead_next, alive_next):
    """New middle row (per-lane masks: dead_next[c9], alive_next[c9])."""
    t0 = a[1] ^ b[1] ^ c[1]
    k = maj(a[1], b[1], c[1])
    u = a[2] ^ b[2] ^ c[2]
    v = maj(a[2], b[2], c[2])
    m = b[0]
    w0 = u ^ k
    w1 = v ^ (u & k)
    w2 = v & u & k
    leaf = [(m & alive_next[c9]) | ((~m & M32) & dead_next[c9])
            for c9 in range(10)]
    mux = lambda s, x, y: (s & x) | ((~s & M32) & y)  # noqa: E731
    l0 = mux(t0, leaf[1], leaf[0])
    l1 = mux(t0, leaf[3], leaf[2])
    l2 = mux(t0, leaf[5], leaf[4])
    l3 = mux(t0, leaf[7], leaf[6])
    l4 = mux(t0, leaf[9], leaf[8])
    m0 = mux(w0, l1, l0)
    m1 = mux(w0, l3, l2)
    n0 = mux(w1, m1, m0)
    return mux(w2, l4, n0)


def kernel(src, total_bytes, insts, lanes, wave_rows, K):
    """One launch; returns the destination words (unwritten words keep a
    marker) and a per-word store count."""
    n_words = total_bytes // 4
    dst = np.full(n_words, 0xDEADBEEF, dtype=np.uint64)
    stores = np.zeros(n_words, dtype=np.int64)
    src = src.astype(np.uint64)
    oor = total_bytes

    def load(offsets):
        offsets = np.asarray(offsets, dtype=np.int64)
        ok = offsets < oor
        assert np.all(offsets[ok] >= 0) and np.all(offsets[ok] % 4 == 0)
        out = np.zeros(len(offsets), dtype=np.uint64)
        out[ok] = src[offsets[ok] // 4]
        return out

    for wave in range(len(wave_rows)):
        d = lanes[wave].astype(np.int64)
        L = int(wave_rows[wave]) + 2 * K
        idle = d[:, 0] < 0
        inst = np.where(idle, 0, d[:, 0])
        g = np.where(idle[:, None], np.array([0, 1, 32, 1]),
                     insts[2 * inst][:, :4])
        r = np.where(idle[:, None], 0, insts[2 * inst + 1][:, :4])
        base, pitch, width, height = (g[:, j].astype(np.int64)
                                      for j in range(4))
        row_bytes = pitch * 4
        s = np.fmod(32 * d[:, 1], width)
        s = np.where(s < 0, s + width, s)
        q, o, m = s >> 5, s & 31, width - s
        wraps = m < 32
        off_a = np.where(idle, oor, q * 4)
        off_b = np.where(~idle & (q + 1 < pitch), (q + 1) * 4, oor)
        off_z = np.where(~idle & wraps, 0, oor)
        keep_m = np.where(wraps, (np.uint64(1) << m.astype(np.uint64))
                          - np.uint64(1), ONES)
        shift_z = np.where(wraps, m, 0).astype(np.uint64)
        keep_out = np.where((d[:, 1] == pitch - 1) & (width % 32 != 0),
                            (np.uint64(1) << (width % 32).astype(np.uint64))
                            - np.uint64(1), ONES)
        in_row = np.fmod(d[:, 2] - K, np.where(idle, 1, height))
        in_row = np.where(in_row < 0, in_row + height, in_row)
        in_off = base * 4 + in_row * row_bytes
        out_off = base * 4 + d[:, 2] * row_bytes + d[:, 1] * 4
        out_rows = d[:, 3]
        born = r[:, 0].astype(np.uint64)
        sustain = r[:, 1].astype(np.uint64)
        dead_next, alive_next = [], []
        for c9 in range(10):
            dead_next.append(np.where(
                (c9 <= 8) & (((born >> np.uint64(min(c9, 31))) & 1) == 1),
                ONES, np.uint64(0)))
            alive_next.append(np.where(
                (c9 >= 1) & ((((born | sustain) >> np.uint64(max(c9 - 1, 0)))
                              & 1) == 1), ONES, np.uint64(0)))
        state = {"in_row": in_row, "in_off": in_off, "out_off": out_off}

        def load_next():
            off = state["in_off"]
            a = load(np.where(off_a == oor, oor, off + off_a))
            b = load(np.where(off_b == oor, oor, off + off_b))
            z = load(np.where(off_z == oor, oor, off + off_z))
            state["in_row"] = state["in_row"] + 1
            state["in_off"] = state["in_off"] + row_bytes
            wrapped = state["in_row"] == height
            state["in_row"] = np.where(wrapped, 0, state["in_row"])
            state["in_off"] = np.where(wrapped, base * 4, state["in_off"])
            return a, b, z

        def word_of(v):
            a, b, z = v
            both = (b << np.uint64(32)) | a
            aligned = (both >> o.astype(np.uint64)) & M32
            return ((aligned & keep_m) | (z << shift_z)) & M32

        def store_row(val, t):
            keep = t < out_rows
            off = state["out_off"]
            for lane in np.nonzero(keep)[0]:
                word_off = int(off[lane])
                assert 0 <= word_off < oor and word_off % 4 == 0, (wave, lane)
                dst[word_off // 4] = int(val[lane] & keep_out[lane])
                stores[word_off // 4] += 1
            state["out_off"] = state["out_off"] + row_bytes

        lev = [[None] * 3 for _ in range(K + 1)]

        def feed(i3, val, i, steady):
            lev[0][i3] = hrow(val)
            for lv in range(1, K + 1):
                if steady or i >= 2 * lv:
                    j3 = (i3 - lv) % 3
                    new = rule_word(lev[lv - 1][(j3 + 2) % 3],
                                    lev[lv - 1][j3],
                                    lev[lv - 1][(j3 + 1) % 3], dead_next,
                                    alive_next)
                    lev[lv][j3] = hrow(new) if lv < K else (new, None, None)

        def top(i3):
            return lev[K][(i3 - K) % 3][0]

        for i in range(2 * K):
            feed(i % 3, word_of(load_next()), i, False)
        p0, p1, p2 = (2 * K) % 3, (2 * K + 1) % 3, (2 * K + 2) % 3
        i = 2 * K
        va, vb, vc = load_next(), load_next(), load_next()
        while i + 2 < L:
            na, nb, nc = load_next(), load_next(), load_next()
            feed(p0, word_of(va), i, True)
            store_row(top(p0), i - 2 * K)
            feed(p1, word_of(vb), i + 1, True)
            store_row(top(p1), i + 1 - 2 * K)
            feed(p2, word_of(vc), i + 2, True)
            store_row(top(p2), i + 2 - 2 * K)
            va, vb, vc = na, nb, nc
            i += 3
        if i < L:
            feed(p0, word_of(va), i, True)
            store_row(top(p0), i - 2 * K)
            if i + 1 < L:
                feed(p1, word_of(vb), i + 1, True)
                store_row(top(p1), i + 1 - 2 * K)
    return dst, stores
