"""
CPU emulation of ``life_bits_kernel`` (``xentica_amd/csrc/life_bits.hip``).

Test infrastructure: it mirrors the kernel's control flow line by line
(the wave plan: whole 62-column tiles over balanced row strips and the
remainder column's segments, interior ones sharing a wave; the K-level
row pipeline with its prologue, steady state of nine rows requested a
whole iteration ahead and remainder, the row wrapping, DPP wave shifts
with zero fill at the wave edges, and stores dropped by the buffer range
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

OUT_LANES = 62
AHEAD = 9
MAX_GROUP = 4


def make_plan(W, H, K, strips):
    """``make_plan``: whole 62-column tiles over balanced row strips, and the
    remainder column's segments, interior ones packed ``group`` to a wave."""
    p = dict(full=W // OUT_LANES, rem=W % OUT_LANES, strips=strips, group=0,
             seg_rows=0, segs=0, e_lo=0, e_hi=0)
    p["waves"] = p["full"] * strips
    if p["rem"] == 0:
        return p
    p["group"] = min(64 // (p["rem"] + 2), MAX_GROUP)
    s = p["seg_rows"] = -(-H // strips)
    p["segs"] = -(-H // s)
    if p["group"] >= 2:
        e_lo = min(-(-K // s), p["segs"])
        hi = (H - K - AHEAD) // s if H - K - AHEAD >= 0 else 0
        p["e_lo"], p["e_hi"] = e_lo, min(max(hi, e_lo), p["segs"])
    shared = p["e_hi"] - p["e_lo"]
    p["waves"] += (p["segs"] - shared) + -(-shared // max(p["group"], 1))
    return p


def wave_tile(p, wave, H):
    """(col_lo, ncols, row_lo, out, seg_lanes, nsegs) of a wave."""
    whole = p["full"] * p["strips"]
    if wave < whole:
        tx, ty = wave % p["full"], wave // p["full"]
        row_lo = ty * H // p["strips"]
        return (tx * OUT_LANES, OUT_LANES, row_lo,
                (ty + 1) * H // p["strips"] - row_lo, 64, 1)
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
            min(H - row_lo, p["seg_rows"]), seg_lanes, nsegs)


def kernel(src, W, H, plan, K, born, sustain, writes):
    fast = (born & 0x1FF) == 8 and (sustain & 0x1FF) == 12
    dst = np.full(W * H, 0xDEADBEEF, dtype=np.uint64)
    nbytes = W * H * 4
    row_bytes = W * 4
    lane = np.arange(64)
    for wave in range(plan["waves"]):
        col_lo, ncols, row_lo, out, seg_lanes, nsegs = wave_tile(plan, wave, H)
        assert 1 <= ncols <= OUT_LANES and out >= 1
        seg = lane // seg_lanes
        sl = lane - seg * seg_lanes
        active = seg < nsegs
        seg_off = seg * plan["seg_rows"] * row_bytes
        col = wrap(col_lo - 1 + sl, W)
        lane_off = np.where(active, col * 4 + seg_off, nbytes)
        out_col = col_lo - 1 + sl
        lane_out = active & (sl >= 1) & (sl <= ncols)
        out_voff = np.where(lane_out, out_col * 4 + seg_off, nbytes)
        L = out + 2 * K; row0 = row_lo - K
        def load_row(i, fed=True):
            soff = wrap(row0 + i, H) * row_bytes
            off = lane_off + soff
            # every load of an active lane is in the field (no reliance on
            # the range check's treatment of the scalar offset); idle lanes
            # carry a vector offset past the end and read zero
            assert np.all(off[active] < nbytes), (wave, i)
            val = np.zeros(64, dtype=np.uint64)
            val[active] = src[off[active] // 4]
            if fed and nsegs > 1:
                # a shared wave's segments never wrap their rows
                assert 0 <= row0 + i < H
            return val
        def store_row(val, r):
            keep = r < out
            voff = out_voff if keep else np.full(64, nbytes)
            soff = (row_lo + r) * row_bytes if keep else 0
            for ln in range(64):
                if voff[ln] < nbytes:   # range check on the vector offset
                    off = voff[ln] + soff
                    assert off < nbytes, (wave, r, ln)
                    dst[off // 4] = val[ln]
                    writes[off // 4] += 1
        lev = [[None] * 3 for _ in range(K + 1)]
        def feed(i, inp, all_levels):
            assert i < L
            lev[0][i % 3] = hrow(inp)
            for l in range(1, K + 1):
                if all_levels or i >= 2 * l:
                    j3 = (i - l) % 3
                    v = life_word(lev[l-1][(j3 + 2) % 3], lev[l-1][j3], lev[l-1][(j3 + 1) % 3], born, sustain, fast)
                    lev[l][j3] = hrow(v) if l < K else (v, None, None)
        # prologue: 2K rows, then the first AHEAD rows requested
        pro = [load_row(r) for r in range(2 * K)]
        i = 2 * K
        cur = [load_row(i + r, fed=i + r < L) for r in range(AHEAD)]
        for r in range(2 * K):
            feed(r, pro[r], False)
        iters = 0
        while i + AHEAD - 1 < L:
            nxt = [load_row(i + AHEAD + r, fed=i + AHEAD + r < L)
                   for r in range(AHEAD)]
            for q in range(AHEAD):
                feed(i + q, cur[q], True)
                store_row(lev[K][(i + q - K) % 3][0], i + q - 2 * K)
            cur = nxt
            i += AHEAD; iters += 1
            assert iters < 100000
        for r in range(AHEAD):
            if i + r >= L:
                break
            feed(i + r, cur[r], True)
            store_row(lev[K][(i + r - K) % 3][0], i + r - 2 * K)
    assert not np.any(dst == 0xDEADBEEF), "unwritten words"
    return dst


def auto_strips(W, H, depth, simds=1024):
    """``auto_strips``: the strip count of the smallest modelled time."""
    issue = (0.0, 0.55, 0.85, 1.0)
    one = make_plan(W, H, depth, 1)
    per_strip = one["full"] + (1.0 / max(one["group"], 1) if one["rem"] else 0)
    best, best_time = 1, None
    for w in (1, 2, 3):
        target = w * simds
        strips = min(int(target / per_strip) + 1, H)
        while strips > 1 and make_plan(W, H, depth, strips)["waves"] > target:
            strips -= 1
        waves = make_plan(W, H, depth, strips)["waves"]
        t = max(1.0, waves / simds) * (H / strips + 2 * depth) / issue[w]
        if best_time is None or t < best_time:
            best, best_time = strips, t
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
