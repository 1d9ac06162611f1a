"""
CPU emulation of ``life_bits_kernel`` (``xentica_amd/csrc/life_bits.hip``).

Test infrastructure: it mirrors the kernel's control flow line by line
(the wave plan: whole 62-column tiles over balanced row strips and the
remainder column's segments, interior ones sharing a wave; the K-level
row pipeline -- rows in threes for any rule, in pairs for B3/S23 -- with
its prologue, steady state of nine (pairs: ten) rows requested a whole
iteration ahead and remainder, the row wrapping, DPP wave shifts with
zero fill at the wave edges, and stores dropped by the buffer range
check) with the 64 lanes of a wavefront as
NumPy vectors, so the kernel's termination, bounds and results can be
checked on a machine without a GPU.
"""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)

def wrap(v, n):
    v = v % n
    return v

def maj(a, b, c):
    return (a & b) | (c & (a | b))

def hrow(w):
    """A row word with its horizontal 3-sums (h0 + 2 h1 = left + self + right)."""
    # DPP wave_shr:1 / wave_shl:1 with bound_ctrl: zero at the wave edges
    wl = np.concatenate([[0], w[:-1]]).astype(np.uint64)
    wr = np.concatenate([w[1:], [0]]).astype(np.uint64)
    l = ((w << 1) | (wl >> 31)) & M32
    r = ((w >> 1) | (wr << 31)) & M32
    return (w, l ^ w ^ r, maj(l, w, r))

def life_word(a, b, c, born, sustain, fast):
    """New middle row from the (raw, h0, h1) of rows up, mid, down."""
    t0 = a[1] ^ b[1] ^ c[1]
    k = maj(a[1], b[1], c[1])
    u = a[2] ^ b[2] ^ c[2]
    v = maj(a[2], b[2], c[2])
    alive = b[0]
    # count9 (3x3 block with the cell) = t0 + 2 (u + k + 2 v)
    if fast:
        e1 = (u ^ k) & (~v & M32)                                # W == 1
        e2 = (u & k & (~v & M32)) | ((~u & M32) & (~k & M32) & v)  # W == 2
        f = e2 & alive & (~t0 & M32)
        return (t0 & e1) | f
    w0 = u ^ k
    w1 = v ^ (u & k)
    w2 = v & u & k
    res = np.zeros_like(alive)
    for c9 in range(10):
        eq = np.full_like(alive, M32)
        for bit, cb in enumerate((t0, w0, w1, w2)):
            eq &= cb if (c9 >> bit) & 1 else (~cb & M32)
        dead_next = M32 if c9 <= 8 and (born >> c9) & 1 else np.uint64(0)
        n = c9 - 1
        live_next = M32 if 0 <= n <= 8 and ((born | sustain) >> n) & 1 \
            else np.uint64(0)
        res |= eq & ((alive & live_next) | ((~alive & M32) & dead_next))
    return res

def pair_sum(r0, r1, s0, s1):
    """``pair_sum``: p0 + 2 p1 + 4 p2 = (r0 + 2 r1) + (s0 + 2 s1)."""
    c = r0 & s0
    return r0 ^ s0, r1 ^ s1 ^ c, maj(r1, s1, c)


def pair_life(a0, a1, p0, p1, p2, m):
    """``pair_life``: B3/S23 from a row's horizontal sum (a0, a1), the
    pair sum (p0, p1, p2) and the cell's state m, in the kernel's four
    three-input operations."""
    o1 = (a0 ^ p0 ^ m) & (~maj(a0, p0, m) & M32)           # one of three
    o2 = (a1 ^ p1) | (a1 & p1 & p2)
    z = (~(m ^ o1 ^ o2) & M32) & maj(m, o1, o2)           # two of three
    return z & (~(p2 & o2) & M32)


OUT_LANES = 62
AHEAD = 10        # kAhead: the most rows a wave requests past its window
ROWS_TRIPLE = 9   # rows fed per steady-state iteration, rows in threes
ROWS_PAIR = 10    # the same, B3/S23 rows in pairs
MAX_GROUP = 4
G = 4           # waves per workgroup
MIN_EXCHANGE_DEPTH = 4


def exchange_fits(H, K, groups):
    """``exchange_fits``: every strip of a group but the last has at least
    2K rows, the last at least one."""
    if groups < 1 or groups > H:
        return False
    rmin = H // groups
    rmax = rmin + 1 if H % groups else rmin
    for R in range(rmin, rmax + 1):
        hx = (R + K - 1) // G
        if hx < 2 * K or R - (G - 1) * hx < 1:
            return False
    return True


def make_plan(W, H, K, strips):
    """``make_plan``: whole 62-column tiles over row strips (exchanging
    groups of G strips where they fit), and the remainder column's
    segments, interior ones packed ``group`` to a wave."""
    p = dict(full=W // OUT_LANES, rem=W % OUT_LANES, strips=strips, group=0,
             seg_rows=0, segs=0, e_lo=0, e_hi=0, xgroups=0)
    if p["full"] > 0 and K >= MIN_EXCHANGE_DEPTH and strips % G == 0 and \
            exchange_fits(H, K, strips // G):
        p["xgroups"] = strips // G
    p["waves"] = p["full"] * strips
    if p["rem"] == 0:
        return p
    p["group"] = min(64 // (p["rem"] + 2), MAX_GROUP)
    if p["xgroups"]:
        s = p["seg_rows"] = (H // p["xgroups"] + K - 1) // G - (K - 1)
    else:
        s = p["seg_rows"] = -(-H // strips)
    p["segs"] = -(-H // s)
    if p["group"] >= 2:
        e_lo = min(-(-K // s), p["segs"])
        hi = (H - K - AHEAD) // s if H - K - AHEAD >= 0 else 0
        p["e_lo"], p["e_hi"] = e_lo, min(max(hi, e_lo), p["segs"])
    shared = p["e_hi"] - p["e_lo"]
    p["waves"] += (p["segs"] - shared) + -(-shared // max(p["group"], 1))
    return p


def wave_tile(p, wave, H, K):
    """(col_lo, ncols, row_lo, out, seg_lanes, nsegs, exchange) of a wave."""
    whole = p["full"] * p["strips"]
    if wave < whole and p["xgroups"]:
        tx = wave // p["strips"]
        rest = wave - tx * p["strips"]
        j, wig = rest // G, rest % G
        g0, g1 = j * H // p["xgroups"], (j + 1) * H // p["xgroups"]
        hx = (g1 - g0 + K - 1) // G
        row_lo = g0 + wig * hx
        exchange = wig < G - 1
        out = hx if exchange else g1 - row_lo
        return (tx * OUT_LANES, OUT_LANES, row_lo, out, 64, 1, exchange)
    if wave < whole:
        tx, ty = wave % p["full"], wave // p["full"]
        row_lo = ty * H // p["strips"]
        return (tx * OUT_LANES, OUT_LANES, row_lo,
                (ty + 1) * H // p["strips"] - row_lo, 64, 1, False)
    r = wave - whole
    shared = p["e_hi"] - p["e_lo"]
    solo = p["segs"] - shared
    seg_lanes, nsegs = 64, 1
    if r < solo:
        seg = r if r < p["e_lo"] else p["e_hi"] + (r - p["e_lo"])
    else:
        seg = p["e_lo"] + (r - solo) * p["group"]
        seg_lanes = p["rem"] + 2
        nsegs = min(p["e_hi"] - seg, p["group"])
    row_lo = seg * p["seg_rows"]
    return (p["full"] * OUT_LANES, p["rem"], row_lo,
            min(H - row_lo, p["seg_rows"]), seg_lanes, nsegs, False)


class Wave:
    """One wavefront of the kernel: ``prologue`` (the first 2K feeds and
    the rows handed over) and ``finish`` (after the workgroup barrier)."""

    def __init__(self, src, dst, W, H, plan, K, born, sustain, writes, wave):
        self.src, self.dst, self.W, self.H, self.K = src, dst, W, H, K
        self.born, self.sustain, self.writes = born, sustain, writes
        self.fast = (born & 0x1FF) == 8 and (sustain & 0x1FF) == 12
        (col_lo, ncols, row_lo, out, seg_lanes, nsegs,
         self.exchange) = wave_tile(plan, wave, H, K)
        assert 1 <= ncols <= OUT_LANES and out >= 1
        if self.exchange:
            assert out >= 2 * K
        self.wave, self.out, self.row_lo = wave, out, row_lo
        self.nbytes = W * H * 4
        self.row_bytes = W * 4
        lane = np.arange(64)
        seg = lane // seg_lanes
        sl = lane - seg * seg_lanes
        self.active = seg < nsegs
        self.nsegs = nsegs
        seg_off = seg * plan["seg_rows"] * self.row_bytes
        col = wrap(col_lo - 1 + sl, W)
        self.lane_off = np.where(self.active, col * 4 + seg_off, self.nbytes)
        out_col = col_lo - 1 + sl
        lane_out = self.active & (sl >= 1) & (sl <= ncols)
        self.out_voff = np.where(lane_out, out_col * 4 + seg_off, self.nbytes)
        self.L = out + 2 if self.exchange else out + 2 * K
        self.row0 = row_lo - K
        self.lev = [[None] * 3 for _ in range(K + 1)]
        self.rows = ROWS_PAIR if self.fast else ROWS_TRIPLE
        zero = np.zeros(64, dtype=np.uint64)
        # PairPipe: q0..q2, s0, s1, c of each level 1 .. K
        self.pq = [[zero, zero, zero] for _ in range(K + 1)]
        self.ps = [[zero, zero] for _ in range(K + 1)]
        self.pc = [zero for _ in range(K + 1)]
        self.last = None
        self.exports = {}

    def load_row(self, i):
        soff = wrap(self.row0 + i, self.H) * self.row_bytes
        off = self.lane_off + soff
        # every load of an active lane is in the field (no reliance on the
        # range check's treatment of the scalar offset); idle lanes carry a
        # vector offset past the end and read zero
        assert np.all(off[self.active] < self.nbytes), (self.wave, i)
        val = np.zeros(64, dtype=np.uint64)
        val[self.active] = self.src[off[self.active] // 4]
        if self.nsegs > 1:
            # a shared wave's segments never wrap their rows, fed or not
            assert 0 <= self.row0 + i < self.H
        return val

    def store_row(self, val, r):
        keep = r < self.out
        assert keep
        voff = self.out_voff
        soff = (self.row_lo + r) * self.row_bytes
        for ln in range(64):
            if voff[ln] < self.nbytes:   # range check on the vector offset
                off = voff[ln] + soff
                assert off < self.nbytes, (self.wave, r, ln)
                self.dst[off // 4] = val[ln]
                self.writes[off // 4] += 1

    def compute(self, l, i):
        """Level l's row i - l from level l - 1's rows around it."""
        K = self.K
        j3 = (i - l) % 3
        lv = self.lev[l - 1]
        v = life_word(lv[(j3 + 2) % 3], lv[j3], lv[(j3 + 1) % 3], self.born,
                      self.sustain, self.fast)
        self.lev[l][j3] = hrow(v) if l < K else (v, None, None)

    def take(self, l, x, odd, pair, out):
        """``PairPipe::take``: level l takes a row of level l - 1 (odd: an
        odd row n) and returns its own row n - 1 when ``out``."""
        v = None
        if odd:
            zero = np.zeros(64, dtype=np.uint64)
            p = pair_sum(self.pq[l][0], self.pq[l][1], x[1], x[2]) if pair \
                else (zero, zero, zero)
            if out:
                v = pair_life(self.ps[l][0], self.ps[l][1], *p, self.pc[l])
            self.pq[l] = list(p)
            self.ps[l] = [x[1], x[2]]
        else:
            if out:
                v = pair_life(x[1], x[2], *self.pq[l], self.pc[l])
            self.pq[l] = [x[1], x[2], self.pq[l][2]]
        self.pc[l] = x[0]
        return v

    def feed_pairs(self, i, inp, all_levels):
        """``PairPipe::feed``: level l takes row i - l + 1 of level l - 1
        from i = 2l - 2 on and completes its row i - l from i = 2l on."""
        K = self.K
        x = hrow(inp)
        for l in range(1, K + 1):
            if not all_levels and i < 2 * l - 2:
                break
            v = self.take(l, x, (i - l + 1) % 2 == 1,
                          all_levels or i >= 2 * l - 1, all_levels or i >= 2 * l)
            if not all_levels and i < 2 * l:
                break
            if l < K:
                x = hrow(v)
            else:
                self.last = v

    def output(self, i):
        """The level-K row completed by the feed of row i."""
        if self.fast:
            return self.last
        return self.lev[self.K][(i - self.K) % 3][0]

    def feed(self, i, inp, all_levels):
        assert i < self.L
        if self.fast:
            self.feed_pairs(i, inp, all_levels)
            return
        self.lev[0][i % 3] = hrow(inp)
        for l in range(1, self.K + 1):
            if all_levels or i >= 2 * l:
                self.compute(l, i)

    def prologue(self):
        K = self.K
        self.pro = [self.load_row(r) for r in range(2 * K)]
        self.i = 2 * K
        self.cur = [self.load_row(self.i + r) for r in range(self.rows)]
        for r in range(2 * K):
            self.feed(r, self.pro[r], False)
            if r >= 2:
                m = r // 2
                row = self.pc[m + 1] if self.fast else \
                    self.lev[m][(r - m) % 3][0]
                self.exports[(m, r & 1)] = row.copy()

    def finish(self, successor):
        K, L = self.K, self.L
        i, cur = self.i, self.cur
        iters = 0
        rows = self.rows
        while i + rows - 1 < L:
            nxt = [self.load_row(i + rows + r) for r in range(rows)]
            for q in range(rows):
                self.feed(i + q, cur[q], True)
                self.store_row(self.output(i + q), i + q - 2 * K)
            cur = nxt
            i += rows
            iters += 1
            assert iters < 100000
        for r in range(rows):
            if i + r >= L:
                break
            self.feed(i + r, cur[r], True)
            self.store_row(self.output(i + r), i + r - 2 * K)
        if not self.exchange or K == 1:
            return
        # drain: feed D = 2 .. 2K-1 takes level D // 2's row from the next
        # strip's exports; the levels above compute, the output is stored
        assert successor is not None
        out = self.out
        for D in range(2, 2 * K):
            i = out + D
            mid = D // 2
            x = hrow(successor.exports[(mid, D % 2)])
            if self.fast:
                for l in range(mid + 1, K + 1):
                    v = self.take(l, x, (i - l + 1) % 2 == 1, True, True)
                    if l < K:
                        x = hrow(v)
                    else:
                        self.last = v
            else:
                self.lev[mid][(i - mid) % 3] = x
                for l in range(mid + 1, K + 1):
                    self.compute(l, i)
            self.store_row(self.output(i), i - 2 * K)


def kernel(src, W, H, plan, K, born, sustain, writes):
    dst = np.full(W * H, 0xDEADBEEF, dtype=np.uint64)
    for first in range(0, plan["waves"], G):
        waves = [Wave(src, dst, W, H, plan, K, born, sustain, writes, w)
                 for w in range(first, min(first + G, plan["waves"]))]
        for wv in waves:
            wv.prologue()
        # __syncthreads: every wave of the workgroup has exported its rows
        for g, wv in enumerate(waves):
            succ = waves[g + 1] if g + 1 < len(waves) else None
            if wv.exchange:
                # the successor is the next strip of the same column tile
                assert succ is not None and succ.row_lo == wv.row_lo + wv.out
            wv.finish(succ)
    assert not np.any(dst == 0xDEADBEEF), "unwritten words"
    return dst


def wave_cost(p, H, K):
    """``wave_cost``: modelled level rows per level of one wave."""
    if p["xgroups"]:
        return (H // p["xgroups"] + K - 1) // G + K + 1.0
    return H / p["strips"] + 2.0 * K


def auto_strips(W, H, depth, simds=1024):
    """``auto_strips``: the strip count of the smallest modelled time."""
    issue = (0.0, 0.55, 0.85, 1.0)
    one = make_plan(W, H, depth, 1)
    per_strip = one["full"] + (1.0 / max(one["group"], 1) if one["rem"] else 0)
    max_groups = H // (2 * depth * G - depth + 1)
    best, best_time = 1, None
    for w in (1, 2, 3):
        target = w * simds
        strips = min(int(target / per_strip) + 1, H)
        while strips > 1 and make_plan(W, H, depth, strips)["waves"] > target:
            strips -= 1
        cands = [strips, 0]
        xs = min(strips // G * G, max_groups * G)
        while xs >= G:
            p = make_plan(W, H, depth, xs)
            if p["xgroups"] and p["waves"] <= target:
                cands[1] = xs
                break
            xs -= G
        for c in cands:
            if c < 1:
                continue
            p = make_plan(W, H, depth, c)
            held = min(w, max(1, -(-p["waves"] // simds)))
            t = max(1.0, p["waves"] / simds) * wave_cost(p, H, depth) / issue[held]
            if best_time is None or t < best_time:
                best, best_time = c, t
    return best


def pack(cells, W, H):
    bits = np.zeros(W * H, dtype=np.uint64)
    c = cells.reshape(H, W, 32).astype(np.uint64)
    for b in range(32):
        bits |= (c[:, :, b].reshape(-1) << np.uint64(b))
    return bits

def unpack(bits, W, H):
    out = np.zeros((W * H, 32), dtype=np.uint8)
    for b in range(32):
        out[:, b] = (bits >> np.uint64(b)) & np.uint64(1)
    return out.reshape(-1)



def run(cells, width, height, depth, born, sustain, steps, strips=0):
    """Run the emulated kernel launch sequence; returns the uint8 plane."""
    W, H = width // 32, height
    bits = pack(cells, W, H)
    left = steps
    while left > 0:
        k = min(depth, left)
        rows = strips if strips > 0 else auto_strips(W, H, k)
        plan = make_plan(W, H, k, min(rows, H))
        writes = np.zeros(W * H, dtype=np.int64)
        bits = kernel(bits, W, H, plan, k, born, sustain, writes)
        assert np.all(writes == 1), "a word stored twice or not at all"
        left -= k
    return unpack(bits, W, H)
