"""
xen_nca_backward_noalloc (static device workspace, no allocation) against
xen_nca_backward (caller's workspace), both called through the C ABI: the
gradients must be bit-identical, on two calls in a row, the second on a
smaller grid than the first.
"""
import ctypes

import numpy as np
import pytest

from oracle import nca_oracle

pytestmark = pytest.mark.gpu

#: XEN_NCA_PARAMS of include/xentica_nca.h
PARAMS = 2288

_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_i = ctypes.c_int


def _bind():
    from xentica_amd.runtime import hip
    lib = hip.library()
    lib.xen_nca_partial_rows.argtypes = [_ll, _i, _i]
    lib.xen_nca_partial_rows.restype = _ll
    lib.xen_nca_backward.argtypes = [_vp, _vp, _ll, _i, _i, _vp, _vp, _vp,
                                     _vp, _vp, _vp, _vp, _ll, _vp, _vp]
    lib.xen_nca_backward.restype = _i
    lib.xen_nca_backward_noalloc.argtypes = [_vp, _vp, _ll, _i, _i, _vp, _vp,
                                             _vp, _vp, _vp, _vp, _vp, _vp]
    lib.xen_nca_backward_noalloc.restype = _i
    return hip, lib


def _params(seed):
    import torch
    params = nca_oracle.init_params(seed)
    out = []
    for name, _ in nca_oracle.PARAM_SHAPES:
        tensor = torch.from_numpy(params[name]).cuda()
        if name in ("w1", "w2"):
            tensor = tensor.to(torch.bfloat16)
        out.append(tensor.contiguous())
    return out


def _grads(hip, lib, x, dy, params, noalloc):
    import torch
    batch, h, w, _ = x.shape
    grads = torch.full((PARAMS, ), float("nan"),
                       device="cuda", dtype=torch.float32)
    ptrs = [_vp(t.data_ptr()) for t in params]
    stream = _vp(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = lib.xen_nca_backward_noalloc(
            _vp(x.data_ptr()), _vp(dy.data_ptr()), batch, h, w, *ptrs,
            _vp(grads.data_ptr()), stream)
    else:
        rows = int(lib.xen_nca_partial_rows(batch, h, w))
        assert rows > 0
        partials = torch.empty(rows * PARAMS, device="cuda",
                               dtype=torch.float32)
        rc = lib.xen_nca_backward(
            _vp(x.data_ptr()), _vp(dy.data_ptr()), batch, h, w, *ptrs,
            _vp(partials.data_ptr()), rows, _vp(grads.data_ptr()), stream)
    hip.check(rc, "backward (noalloc=%s)" % noalloc)
    torch.cuda.synchronize()
    return grads.cpu().numpy()


@pytest.mark.parametrize("shapes", [((4, 32, 64), (1, 8, 16)),
                                    ((1, 256, 256), (2, 16, 32))])
def test_noalloc_backward_equals_workspace_backward(shapes):
    """Two calls in a row (the second on a smaller grid, so stale rows of
    the first call stay in the workspace): bit-identical gradients."""
    import torch
    hip, lib = _bind()
    rng = np.random.default_rng(7)
    params = _params(3)
    for batch, h, w in shapes:
        print("noalloc backward grid", (batch, h, w), flush=True)
        x = torch.from_numpy(rng.uniform(0, 8, (batch, h, w, 16)).astype(
            np.float32)).cuda()
        dy = torch.from_numpy(rng.standard_normal((batch, h, w, 16)).astype(
            np.float32)).cuda()
        want = _grads(hip, lib, x, dy, params, noalloc=False)
        got = _grads(hip, lib, x, dy, params, noalloc=True)
        assert np.all(np.isfinite(want))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
            (batch, h, w)
