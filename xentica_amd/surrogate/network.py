"""
The surrogate network on the hand-written kernels (``include/xentica_nca.h``).

:class:`NCAFunction` is the ``torch.autograd.Function`` whose forward and
backward are ``xen_nca_forward`` and ``xen_nca_backward`` (bf16 MFMA with
fp32 accumulation; the gradient of the bf16 copies of ``w1``/``w2`` is
applied to their fp32 masters).  PyTorch only provides device memory,
streams and the optimiser; there is no PyTorch or CPU path for the
network, and the kernels raise if the library is missing.

Shapes: ``x``/``y`` ``[batch, h, w, 16]`` float32 on the GPU, coarse grid
a torus per sample, ``batch * h * w`` a multiple of 16.
"""
import ctypes

import torch

from xentica_amd.core.exceptions import XenticaException
from xentica_amd.runtime import hip

__all__ = ['NCAFunction', 'NCASurrogate', 'encode', 'forward_raw', 'predict',
           'PARAM_SHAPES', 'PARAMS']

CHANNELS = 16
HIDDEN = 64
#: (name, shape) in the kernels' parameter order.
PARAM_SHAPES = (('dw', (CHANNELS, 9)), ('bdw', (CHANNELS, )),
                ('w1', (HIDDEN, CHANNELS)), ('b1', (HIDDEN, )),
                ('w2', (CHANNELS, HIDDEN)), ('b2', (CHANNELS, )))
PARAMS = sum(int(torch.Size(shape).numel()) for _, shape in PARAM_SHAPES)

_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_i = ctypes.c_int
_u32 = ctypes.c_uint32
_SIGNATURES = {
    'xen_nca_encode': ([_vp, _ll, _i, _i, _u32, _vp, _vp, _vp], _i),
    'xen_nca_forward': ([_vp, _ll, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                         _vp, _vp, _vp], _i),
    'xen_nca_predict': ([_vp, _ll, _i, _i, _u32, _vp, _vp, _vp, _vp, _vp,
                         _vp, _vp, _vp, _vp], _i),
    'xen_nca_partial_rows': ([_ll, _i, _i], _ll),
    'xen_nca_backward': ([_vp, _vp, _ll, _i, _i, _vp, _vp, _vp, _vp, _vp,
                          _vp, _vp, _ll, _vp, _vp], _i),
}
_BOUND = False


def _lib():
    global _BOUND  # pylint: disable=global-statement
    lib = hip.library()
    if not _BOUND:
        for name, (argtypes, restype) in _SIGNATURES.items():
            func = getattr(lib, name)
            func.argtypes = argtypes
            func.restype = restype
        _BOUND = True
    return lib


def _stream(stream):
    """HIP stream handle: a :class:`hip.Stream`, or torch's current one."""
    if stream is not None:
        return stream.handle
    return _vp(torch.cuda.current_stream().cuda_stream)


def _dev(tensor, dtype, name):
    if not tensor.is_cuda or tensor.dtype != dtype or \
            not tensor.is_contiguous():
        raise XenticaException(
            "%s must be a contiguous %s tensor on the GPU" % (name, dtype))
    return _vp(tensor.data_ptr())


def _grid(x):
    if x.dim() != 4 or x.shape[-1] != CHANNELS:
        raise XenticaException("x must be [batch, h, w, 16]")
    return int(x.shape[0]), int(x.shape[1]), int(x.shape[2])


def encode(fine_ptr, batch, h, w, x, block_mean=None, value_mask=0xFF,
           stream=None):
    """
    Coarse-grain a fine uint8 field at device address ``fine_ptr``
    (``[batch, 16h, 16w]``) into ``x`` ``[batch, h, w, 16]`` and, when
    given, ``block_mean`` ``[batch, h, w]``.
    """
    if tuple(x.shape) != (batch, h, w, CHANNELS):
        raise XenticaException("x has the wrong shape")
    mean = None
    if block_mean is not None:
        if tuple(block_mean.shape) != (batch, h, w):
            raise XenticaException("block_mean has the wrong shape")
        mean = _dev(block_mean, torch.float32, "block_mean")
    hip.check(_lib().xen_nca_encode(
        _vp(fine_ptr), batch, h, w, value_mask,
        _dev(x, torch.float32, "x"), mean, _stream(stream)), "nca encode")


def forward_raw(x, params, y, pred=None, stream=None):
    """
    One forward launch: ``params`` = (dw, bdw, w1 bf16, b1, w2 bf16, b2)
    device tensors; writes ``y`` and, when given, ``pred``.
    """
    batch, h, w = _grid(x)
    dw, bdw, w1, b1, w2, b2 = params
    hip.check(_lib().xen_nca_forward(
        _dev(x, torch.float32, "x"), batch, h, w,
        _dev(dw, torch.float32, "dw"), _dev(bdw, torch.float32, "bdw"),
        _dev(w1, torch.bfloat16, "w1"), _dev(b1, torch.float32, "b1"),
        _dev(w2, torch.bfloat16, "w2"), _dev(b2, torch.float32, "b2"),
        _dev(y, torch.float32, "y"),
        _dev(pred, torch.float32, "pred") if pred is not None else None,
        _stream(stream)), "nca forward")


def predict(fine_ptr, batch, h, w, params, y=None, pred=None, value_mask=0xFF,
            stream=None):
    """
    :func:`encode` and :func:`forward_raw` in one launch, without ``x``:
    the fine uint8 field at ``fine_ptr`` (``[batch, 16h, 16w]``) to ``y``
    ``[batch, h, w, 16]`` and/or ``pred`` ``[batch, h, w]``, bit for bit
    what the two launches give.  ``h`` must be a multiple of 8 and ``w``
    of 16.
    """
    if y is None and pred is None:
        raise XenticaException("predict needs y or pred")
    if y is not None and tuple(y.shape) != (batch, h, w, CHANNELS):
        raise XenticaException("y has the wrong shape")
    if pred is not None and tuple(pred.shape) != (batch, h, w):
        raise XenticaException("pred has the wrong shape")
    dw, bdw, w1, b1, w2, b2 = params
    hip.check(_lib().xen_nca_predict(
        _vp(fine_ptr), batch, h, w, value_mask,
        _dev(dw, torch.float32, "dw"), _dev(bdw, torch.float32, "bdw"),
        _dev(w1, torch.bfloat16, "w1"), _dev(b1, torch.float32, "b1"),
        _dev(w2, torch.bfloat16, "w2"), _dev(b2, torch.float32, "b2"),
        _dev(y, torch.float32, "y") if y is not None else None,
        _dev(pred, torch.float32, "pred") if pred is not None else None,
        _stream(stream)), "nca predict")


class NCAFunction(torch.autograd.Function):
    """``y = network(x)`` with the hand-written forward and backward."""

    @staticmethod
    def forward(ctx, x, dw, bdw, w1, b1, w2, b2):
        # pylint: disable=arguments-differ
        x = x.contiguous()
        w1b = w1.detach().to(torch.bfloat16).contiguous()
        w2b = w2.detach().to(torch.bfloat16).contiguous()
        params = (dw.detach().contiguous(), bdw.detach().contiguous(), w1b,
                  b1.detach().contiguous(), w2b, b2.detach().contiguous())
        y = torch.empty_like(x)
        forward_raw(x, params, y)
        ctx.save_for_backward(x, *params)
        return y

    @staticmethod
    def backward(ctx, dy):
        # pylint: disable=arguments-differ
        x, dw, bdw, w1b, b1, w2b, b2 = ctx.saved_tensors
        dy = dy.contiguous().float()
        batch, h, w = _grid(x)
        lib = _lib()
        rows = int(lib.xen_nca_partial_rows(batch, h, w))
        if rows < 0:
            raise XenticaException("nca backward: invalid grid")
        partials = torch.empty(rows * PARAMS, device=x.device,
                               dtype=torch.float32)
        grads = torch.empty(PARAMS, device=x.device, dtype=torch.float32)
        hip.check(lib.xen_nca_backward(
            _dev(x, torch.float32, "x"), _dev(dy, torch.float32, "dy"),
            batch, h, w, _dev(dw, torch.float32, "dw"),
            _dev(bdw, torch.float32, "bdw"), _dev(w1b, torch.bfloat16, "w1"),
            _dev(b1, torch.float32, "b1"), _dev(w2b, torch.bfloat16, "w2"),
            _dev(b2, torch.float32, "b2"), _vp(partials.data_ptr()), rows,
            _vp(grads.data_ptr()), _stream(None)), "nca backward")
        out, pos = [], 0
        for _, shape in PARAM_SHAPES:
            size = int(torch.Size(shape).numel())
            out.append(grads[pos:pos + size].view(shape))
            pos += size
        return (None, *out)


class NCASurrogate(torch.nn.Module):
    """
    3x3 depthwise perception (16 channels) + 1x1 MLP 16 -> 64 -> 16.

    Input: the 16 sub-block means of every coarse cell at step t; output:
    the same 16 channels predicted 16 steps later; the coarse observable
    (block mean height) is the mean of the output channels.

    The trainable parameters act on standardised values
    ``(x - shift) / scale``; the standardisation is folded into the
    kernels' parameters (``dw / scale``, ``bdw - shift / scale * sum(dw)``,
    ``w2 * scale``, ``b2 * scale + shift``), so the device reads raw
    encodings and writes raw heights with no extra passes.
    """

    def __init__(self, seed=0, device="cuda"):
        super().__init__()
        gen = torch.Generator().manual_seed(seed)
        scale = {'dw': 0.3, 'bdw': 0.0, 'w1': 0.25, 'b1': 0.0, 'w2': 0.12,
                 'b2': 0.0}
        for name, shape in PARAM_SHAPES:
            value = torch.randn(shape, generator=gen) * scale[name]
            self.register_parameter(
                name, torch.nn.Parameter(value.to(device)))
        self.register_buffer("shift", torch.zeros((), device=device))
        self.register_buffer("scale", torch.ones((), device=device))

    def set_normalization(self, shift, scale):
        """Standardise inputs and outputs by ``shift`` and ``scale``."""
        self.shift.fill_(float(shift))
        self.scale.fill_(float(scale))

    def tensors(self):
        """Trainable parameters in the kernels' order."""
        return [getattr(self, name) for name, _ in PARAM_SHAPES]

    def kernel_tensors(self):
        """Kernel parameters with the standardisation folded in."""
        dw, bdw, w1, b1, w2, b2 = self.tensors()
        inv = 1.0 / self.scale
        return [dw * inv, bdw - self.shift * inv * dw.sum(dim=1), w1, b1,
                w2 * self.scale, b2 * self.scale + self.shift]

    def forward(self, x):  # pylint: disable=arguments-differ
        """Differentiable prediction of the 16 channels 16 steps ahead."""
        return NCAFunction.apply(x, *self.kernel_tensors())

    def inference_params(self):
        """Device tensors for :func:`forward_raw` (bf16 weight copies)."""
        with torch.no_grad():
            dw, bdw, w1, b1, w2, b2 = self.kernel_tensors()
            return (dw.detach().contiguous(), bdw.detach().contiguous(),
                    w1.detach().to(torch.bfloat16).contiguous(),
                    b1.detach().contiguous(),
                    w2.detach().to(torch.bfloat16).contiguous(),
                    b2.detach().contiguous())
