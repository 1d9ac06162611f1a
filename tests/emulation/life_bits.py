"""
CPU emulation of ``life_bits_kernel`` (``xentica_amd/csrc/life_bits.hip``).

Test infrastructure: it mirrors the kernel's control flow line by line
(balanced column/row tiles, the K-level row pipeline with its prologue,
steady state unrolled by three and remainder, the wave-uniform row
wrapping, DPP wave shifts with zero fill at the wave edges, and stores
dropped by the buffer range check) with the 64 lanes of a wavefront as
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

def kernel(src, W, H, tiles_x, tiles_y, K, born, sustain):
    fast = (born & 0x1FF) == 8 and (sustain & 0x1FF) == 12
    dst = np.full(W * H, 0xDEADBEEF, dtype=np.uint64)
    nbytes = W * H * 4
    lane = np.arange(64)
    for wave in range(tiles_x * tiles_y):
        tx, ty = wave % tiles_x, wave // tiles_x
        col_lo = tx * W // tiles_x; col_hi = (tx + 1) * W // tiles_x
        assert col_hi - col_lo <= 62
        col = wrap(col_lo - 1 + lane, W)
        row_lo = ty * H // tiles_y; row_hi = (ty + 1) * H // tiles_y
        out = row_hi - row_lo; assert out >= 1
        L = out + 2 * K; row0 = row_lo - K
        out_col = col_lo - 1 + lane
        lane_out = (lane >= 1) & (out_col < col_hi)
        out_voff = np.where(lane_out, out_col * 4, nbytes)
        def load_row(i):
            soff = wrap(row0 + i, H) * W * 4
            off = col * 4 + soff
            assert np.all(off < nbytes)
            return src[off // 4].copy()
        def store_row(val, r):
            keep = r < out
            voff = out_voff if keep else np.full(64, nbytes)
            soff = (row_lo + r) * W * 4 if keep else 0
            for ln in range(64):
                if voff[ln] < nbytes:   # range check on the vector offset
                    off = voff[ln] + soff
                    assert off < nbytes, (wave, r, ln)
                    dst[off // 4] = val[ln]
        lev = [[None] * 3 for _ in range(K + 1)]
        def feed(i, inp, all_levels):
            lev[0][i % 3] = hrow(inp)
            for l in range(1, K + 1):
                if all_levels or i >= 2 * l:
                    j3 = (i - l) % 3
                    v = life_word(lev[l-1][(j3 + 2) % 3], lev[l-1][j3], lev[l-1][(j3 + 1) % 3], born, sustain, fast)
                    lev[l][j3] = hrow(v) if l < K else (v, None, None)
        for i in range(2 * K):
            feed(i, load_row(i), False)
        i = 2 * K
        a, b, c = load_row(i), load_row(i + 1), load_row(i + 2)
        na, nb, nc = load_row(i + 3), load_row(i + 4), load_row(i + 5)
        iters = 0
        while i + 2 < L:
            fa, fb, fc = load_row(i + 6), load_row(i + 7), load_row(i + 8)
            for q, v in enumerate((a, b, c)):
                feed(i + q, v, True)
                store_row(lev[K][(i + q - K) % 3][0], i + q - 2 * K)
            a, b, c = na, nb, nc
            na, nb, nc = fa, fb, fc
            i += 3; iters += 1
            assert iters < 100000
        if i < L:
            feed(i, a, True); store_row(lev[K][(i - K) % 3][0], i - 2 * K)
            if i + 1 < L:
                feed(i + 1, b, True); store_row(lev[K][(i + 1 - K) % 3][0], i + 1 - 2 * K)
    assert not np.any(dst == 0xDEADBEEF), "unwritten words"
    return dst

def auto_strips(tiles_x, H, depth, simds=1024):
    strips = (4 * simds) // tiles_x
    max_strips = H // (4 * depth) if H // (4 * depth) > 0 else 1
    strips = min(strips, max_strips)
    return max(1, strips)

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
    tiles_x = (W + 61) // 62
    tiles_y = strips if strips > 0 else auto_strips(tiles_x, H, depth)
    tiles_y = min(tiles_y, H)
    bits = pack(cells, W, H)
    left = steps
    while left > 0:
        k = min(depth, left)
        bits = kernel(bits, W, H, tiles_x, tiles_y, k, born, sustain)
        left -= k
    return unpack(bits, W, H)
