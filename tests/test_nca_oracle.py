"""
The surrogate's CPU oracle, pinned without a GPU.

* bf16 rounding against PyTorch's own float32 -> bfloat16 conversion;
* the encoder against a reshape-based block mean;
* the exact backward against central finite differences (float64);
* the tap convention against the exact model: the shift parameters map
  the encoding at step t to the encoding 16 steps later, which is the
  field rolled by one block (ShiftingSands moves it by (-1, -1) a step).
"""
import numpy as np
import pytest
import torch

from oracle import nca_oracle as nca


def test_bf16_matches_torch():
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.standard_normal(4096).astype(np.float32) * 10,
        # exact ties between two bf16 values round to even
        (np.arange(1, 200, dtype=np.float32) * 2 ** -8 + 1).astype(np.float32),
    ])
    want = torch.from_numpy(vals).to(torch.bfloat16).float().numpy()
    assert np.array_equal(nca.bf16(vals), want)


def test_encode_is_block_of_sub_block_means():
    rng = np.random.default_rng(2)
    fine = rng.integers(0, 256, (2, 48, 64), dtype=np.uint8)
    x, mean = nca.encode(fine, mask=0x0F)
    sub = (fine & 0x0F).reshape(2, 3, 4, 4, 4, 4, 4).astype(np.float64)
    # axes: batch, by, sy, r, bx, sx, c
    want = sub.mean(axis=(3, 6)).transpose(0, 1, 3, 2, 4).reshape(2, 3, 4, 16)
    assert np.array_equal(x, want)
    assert np.allclose(mean, (fine & 0x0F).reshape(2, 3, 16, 4, 16)
                       .mean(axis=(2, 4)))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 8, (1, 3, 4, 16))
    dy = rng.standard_normal((1, 3, 4, 16))
    params = {k: v.astype(np.float64) for k, v in nca.init_params(4).items()}
    analytic = nca.backward(x, dy, params, rounding=False)
    numeric = nca.numeric_gradient(x, dy, params)
    for name, _ in nca.PARAM_SHAPES:
        assert np.allclose(analytic[name], numeric[name], rtol=1e-6,
                           atol=1e-6), name


def test_rounded_backward_stays_close_to_exact():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 8, (1, 8, 8, 16)).astype(np.float32)
    dy = rng.standard_normal((1, 8, 8, 16)).astype(np.float32)
    params = nca.init_params(6)
    exact = nca.backward(x, dy, params, rounding=False)
    rounded = nca.backward(x, dy, params, rounding=True)
    # bf16 operands perturb each term by about 2**-8; the sums cancel, so
    # compare directions rather than elements
    for name, _ in nca.PARAM_SHAPES:
        a, b = np.ravel(rounded[name]), np.ravel(exact[name])
        assert a @ b / np.linalg.norm(a) / np.linalg.norm(b) > 0.99, name


def shift_params():
    """Parameters of the exact 16-step map of the sand model."""
    params = {name: np.zeros(shape, np.float32)
              for name, shape in nca.PARAM_SHAPES}
    params['dw'][:, 8] = 1.0           # tap (dy, dx) = (+1, +1)
    params['w1'][:16, :] = np.eye(16)
    params['w2'][:, :16] = np.eye(16)
    return params


@pytest.mark.parametrize("rounding", [False, True])
def test_shift_parameters_reproduce_the_exact_model(rounding):
    rng = np.random.default_rng(7)
    fine = rng.integers(0, 9, (1, 64, 80), dtype=np.uint8)
    # 16 ShiftingSands steps move the field by (-16, -16)
    later = np.roll(fine, (-16, -16), axis=(1, 2))
    x, _ = nca.encode(fine)
    target, _ = nca.encode(later)
    y, _ = nca.forward(x, shift_params(), rounding=rounding)
    assert np.array_equal(y, target)
