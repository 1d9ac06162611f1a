"""
CPU oracle of the neural-CA surrogate -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` may import
this module, and only as the checker.  The surrogate has no counterpart in
the reference; this restates its definition (the header of
``include/xentica_nca.h``) in float64 NumPy:

* :func:`encode` -- 16 sub-block means per 16x16 block of a fine field;
* :func:`forward` -- 3x3 depthwise perception (torus) + 1x1 MLP
  16 -> 64 -> 16 with ReLU;
* :func:`backward` -- gradients of ``sum(dy * y)`` for every parameter.

With ``rounding=True`` both passes apply the kernels' roundings (fp32 fused
multiply-adds for the perception in tap order, bf16 round-to-nearest-even
of the MFMA operands: perception output, weights, hidden activations,
upstream gradient and ``dZ1``), so the GPU can be compared within fp32
summation-order tolerance; with ``rounding=False`` they are the exact
mathematics, which :func:`numeric_gradient` (central differences) pins.
"""
import numpy as np

__all__ = ['PARAM_SHAPES', 'encode', 'forward', 'backward', 'loss',
           'numeric_gradient', 'bf16', 'init_params', 'flatten', 'unflatten']

#: Parameter names, shapes and order of the flat gradient vector.
PARAM_SHAPES = (('dw', (16, 9)), ('bdw', (16, )), ('w1', (64, 16)),
                ('b1', (64, )), ('w2', (16, 64)), ('b2', (16, )))


def bf16(values):
    """Round to bfloat16 (nearest even), returned as float32."""
    bits = np.ascontiguousarray(values, dtype=np.float32).view(np.uint32)
    bits = bits.astype(np.uint64)
    bits = (bits + 0x7FFF + ((bits >> 16) & 1)) >> 16 << 16
    return bits.astype(np.uint32).view(np.float32)


def init_params(seed=0):
    """A random parameter set (float32 arrays)."""
    rng = np.random.default_rng(seed)
    scale = {'dw': 0.4, 'bdw': 0.1, 'w1': 0.3, 'b1': 0.1, 'w2': 0.2,
             'b2': 0.1}
    return {name: (rng.standard_normal(shape) * scale[name]).astype(
        np.float32) for name, shape in PARAM_SHAPES}


def flatten(grads):
    """Concatenate in the kernel's parameter order."""
    return np.concatenate([np.ravel(grads[name]) for name, _ in PARAM_SHAPES])


def unflatten(vector):
    """Split a flat vector in the kernel's parameter order."""
    out, pos = {}, 0
    for name, shape in PARAM_SHAPES:
        size = int(np.prod(shape))
        out[name] = np.asarray(vector[pos:pos + size]).reshape(shape)
        pos += size
    return out


def encode(fine, mask=0xFF):
    """
    ``fine`` uint8 ``[batch, 16h, 16w]`` -> ``(x [batch, h, w, 16],
    block_mean [batch, h, w])``; channel ``4 * sy + sx`` is the mean of
    the 4x4 sub-block (sy, sx) of each 16x16 block.
    """
    fine = np.asarray(fine) & mask
    batch, fh, fw = fine.shape
    h, w = fh // 16, fw // 16
    x = np.zeros((batch, h, w, 16))
    for sy in range(4):
        for sx in range(4):
            total = np.zeros((batch, h, w))
            for r in range(4):
                for c in range(4):
                    total += fine[:, 4 * sy + r::16, 4 * sx + c::16]
            x[..., 4 * sy + sx] = total / 16
    return x, x.mean(axis=-1)


def _taps(x):
    """The 9 torus neighbours of every cell, tap (dy+1)*3 + (dx+1)."""
    return [np.roll(np.roll(x, -dy, axis=1), -dx, axis=2)
            for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _perceive(x, params, rounding):
    taps = _taps(x)
    if not rounding:
        p = params['bdw'].astype(np.float64) + sum(
            taps[t] * params['dw'][:, t].astype(np.float64) for t in range(9))
        return p, taps
    p = np.broadcast_to(params['bdw'].astype(np.float32),
                        x.shape).astype(np.float32)
    for t in range(9):
        # fp32 fused multiply-add: exact product, one rounding of the sum
        p = (params['dw'][:, t].astype(np.float64)
             * taps[t].astype(np.float32).astype(np.float64)
             + p.astype(np.float64)).astype(np.float32)
    return p, taps


def forward(x, params, rounding=True):
    """``y [..., 16]`` and the intermediates ``(p, z1, h)``."""
    x = np.asarray(x, dtype=np.float32 if rounding else np.float64)
    round_ = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    p, _ = _perceive(x, params, rounding)
    pb = round_(p).astype(np.float64)
    w1 = round_(params['w1']).astype(np.float64)
    w2 = round_(params['w2']).astype(np.float64)
    z1 = pb @ w1.T + params['b1']
    if rounding:
        z1 = z1.astype(np.float32)
    h = round_(np.maximum(z1, 0)).astype(np.float64)
    y = h @ w2.T + params['b2']
    return y, (p, z1, h)


def backward(x, dy, params, rounding=True):
    """Gradients of ``sum(dy * y)`` as a dict of float64 arrays."""
    x = np.asarray(x, dtype=np.float32 if rounding else np.float64)
    round_ = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    y, (p, z1, h) = forward(x, params, rounding)
    del y
    _, taps = _perceive(x, params, rounding)
    pb = round_(p).astype(np.float64)
    w1 = round_(params['w1']).astype(np.float64)
    w2 = round_(params['w2']).astype(np.float64)
    dy = np.asarray(dy, dtype=np.float32 if rounding else np.float64)
    dyb = round_(dy).astype(np.float64)
    dh = dyb @ w2
    dz = np.where(z1 > 0, dh, 0.0)
    if rounding:
        dz = dz.astype(np.float32)
    dzb = round_(dz).astype(np.float64)
    dp = dzb @ w1
    flat = lambda a: a.reshape(-1, a.shape[-1])  # noqa: E731
    grads = {
        'b2': flat(dy.astype(np.float64)).sum(axis=0),
        'w2': flat(dyb).T @ flat(h),
        'b1': flat(dz.astype(np.float64)).sum(axis=0),
        'w1': flat(dzb).T @ flat(pb),
        'bdw': flat(dp).sum(axis=0),
        'dw': np.stack([flat(dp * taps[t]).sum(axis=0) for t in range(9)],
                       axis=1),
    }
    return grads


def loss(x, dy, params):
    """``sum(dy * y)`` of the exact (unrounded) network."""
    y, _ = forward(x, params, rounding=False)
    return float(np.sum(np.asarray(dy, np.float64) * y))


def numeric_gradient(x, dy, params, eps=1e-6):
    """Central differences of :func:`loss` for every parameter entry."""
    base = {k: np.asarray(v, np.float64).copy() for k, v in params.items()}
    grads = {}
    for name, shape in PARAM_SHAPES:
        grad = np.zeros(shape)
        for idx in np.ndindex(*shape):
            old = base[name][idx]
            base[name][idx] = old + eps
            up = loss(x, dy, base)
            base[name][idx] = old - eps
            down = loss(x, dy, base)
            base[name][idx] = old
            grad[idx] = (up - down) / (2 * eps)
        grads[name] = grad
    return grads
