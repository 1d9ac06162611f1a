"""
ctypes binding of the hand-written kernels (``include/xentica_kernels.h``).

Pointers are integer device addresses (``Buffer.ptr`` /
``DeviceArray.device_ptr()``); streams are :class:`hip.Stream` or ``None``.
"""
import ctypes

import numpy as np

from xentica_amd.runtime import hip
from xentica_amd.runtime.crc import combine_partials

__all__ = [
    'life_bits_launch', 'life_bits_run', 'life3d_bits_step',
    'life3d_u8_direct_step', 'bits_pack', 'bits_unpack',
    'crc32_u8', 'crc32_bits', 'CRC_CHUNK', 'low_bytes', 'bits_expand',
    'field_sum',
]

_vp = ctypes.c_void_p
_ll = ctypes.c_longlong
_u32 = ctypes.c_uint32
_i = ctypes.c_int

_SIGNATURES = {
    'xen_life_bits_launch': [_vp, _vp, _i, _i, _u32, _u32, _i, _i, _vp],
    'xen_life_bits_run': [_vp, _vp, _i, _i, _u32, _u32, _ll, _i, _i,
                          ctypes.POINTER(_i), _vp],
    'xen_life3d_bits_step': [_vp, _vp, _i, _i, _i, _u32, _u32, _i, _vp],
    'xen_life3d_u8_direct_step': [_vp, _vp, _i, _i, _i, _u32, _u32, _vp],
    'xen_bits_pack': [_vp, _vp, _ll, _vp],
    'xen_bits_unpack': [_vp, _vp, _ll, _vp],
    'xen_low_bytes': [_vp, _vp, _ll, _vp],
    'xen_bits_expand': [_vp, _vp, _vp, _ll, _vp, _vp, _vp],
    'xen_crc32_u8_partials': [_vp, _ll, _ll, _vp, _vp],
    'xen_crc32_bits_partials': [_vp, _ll, _ll, _vp, _vp],
    'xen_field_sum_partials': [_vp, _ll, _i, _i, _i, _vp, _i, _vp],
}
_BOUND = False

#: Bytes per CRC chunk (one thread each); a multiple of 32.
CRC_CHUNK = 4096


def _lib():
    global _BOUND  # pylint: disable=global-statement
    lib = hip.library()
    if not _BOUND:
        for name, argtypes in _SIGNATURES.items():
            func = getattr(lib, name)
            func.argtypes = argtypes
            func.restype = ctypes.c_int
        _BOUND = True
    return lib


def _s(stream):
    return stream.handle if stream is not None else None


def life_bits_launch(src, dst, width_words, height, born, sustain, depth,
                     strips=0, stream=None):
    """One launch of ``depth`` Life-like steps on a packed torus."""
    hip.check(_lib().xen_life_bits_launch(
        src, dst, width_words, height, born, sustain, depth, strips,
        _s(stream)), "xen_life_bits_launch")


def life_bits_run(buf_a, buf_b, width_words, height, born, sustain, steps,
                  depth, strips=0, stream=None):
    """Run ``steps`` steps; returns True when the result is in ``buf_b``."""
    in_b = _i(0)
    hip.check(_lib().xen_life_bits_run(
        buf_a, buf_b, width_words, height, born, sustain, steps, depth,
        strips, ctypes.byref(in_b), _s(stream)), "xen_life_bits_run")
    return bool(in_b.value)


def life3d_bits_step(src, dst, size, born, sustain, z_tiles=0, stream=None):
    """One step of a packed 3-D Life-like torus of ``size = (w, h, d)``."""
    hip.check(_lib().xen_life3d_bits_step(
        src, dst, size[0], size[1], size[2], born, sustain, z_tiles,
        _s(stream)), "xen_life3d_bits_step")


def life3d_u8_direct_step(src, dst, size, born, sustain, stream=None):
    """One step of a uint8 3-D Life-like torus (direct 26-load stencil)."""
    hip.check(_lib().xen_life3d_u8_direct_step(
        src, dst, size[0], size[1], size[2], born, sustain, _s(stream)),
        "xen_life3d_u8_direct_step")


def bits_pack(cells, bits, n_cells, stream=None):
    """uint8 plane (0/1 per cell) -> packed bits, on the device."""
    hip.check(_lib().xen_bits_pack(cells, bits, n_cells, _s(stream)),
              "xen_bits_pack")


def bits_unpack(bits, cells, n_cells, stream=None):
    """Packed bits -> uint8 plane, on the device."""
    hip.check(_lib().xen_bits_unpack(bits, cells, n_cells, _s(stream)),
              "xen_bits_unpack")


def bits_expand(bits, cells, colors, n_cells, color_dead=(0, 0, 0),
                color_alive=(0, 0, 0), stream=None):
    """Packed bits -> uint8 plane and/or int3 colours by state (device
    pointers; ``cells`` or ``colors`` may be None)."""
    dead = (ctypes.c_int32 * 3)(*color_dead)
    alive = (ctypes.c_int32 * 3)(*color_alive)
    hip.check(_lib().xen_bits_expand(bits, cells, colors, n_cells, dead,
                                     alive, _s(stream)), "xen_bits_expand")


def _partials_buffer(n_chunks):
    return hip.Buffer(max(1, n_chunks) * 4)


def low_bytes(src_ptr, dst_ptr, count, stream=None):
    """``dst[i] = uint8(src[i])`` for ``count`` int32 values on the device."""
    hip.check(_lib().xen_low_bytes(src_ptr, dst_ptr, count, _s(stream)),
              "xen_low_bytes")


def crc32_u8(data_ptr, nbytes, stream=None, chunk=CRC_CHUNK):
    """``binascii.crc32`` of ``nbytes`` device bytes."""
    n_chunks = (nbytes + chunk - 1) // chunk
    parts = _partials_buffer(n_chunks)
    hip.check(_lib().xen_crc32_u8_partials(
        data_ptr, nbytes, chunk, parts.ptr, _s(stream)), "crc32 u8")
    host = np.empty(n_chunks, dtype=np.uint32)
    parts.download(host, stream)
    return combine_partials(host, chunk, nbytes)


def crc32_bits(bits_ptr, n_cells, stream=None, chunk=CRC_CHUNK):
    """``binascii.crc32`` of the uint8 expansion of a packed plane."""
    n_chunks = (n_cells + chunk - 1) // chunk
    parts = _partials_buffer(n_chunks)
    hip.check(_lib().xen_crc32_bits_partials(
        bits_ptr, n_cells, chunk, parts.ptr, _s(stream)), "crc32 bits")
    host = np.empty(n_chunks, dtype=np.uint32)
    parts.download(host, stream)
    return combine_partials(host, chunk, n_cells)


#: Blocks (partial sums) of :func:`field_sum`.
FIELD_SUM_BLOCKS = 1024
_FIELD_PARTIALS = []


def field_sum(words_ptr, n, word_bytes, shift, width, stream=None):
    """Sum over ``n`` container words at ``words_ptr`` of the field
    ``(word >> shift) & (2**width - 1)``, computed on the device (an exact
    integer)."""
    if not _FIELD_PARTIALS:
        _FIELD_PARTIALS.append(hip.Buffer(8 * FIELD_SUM_BLOCKS))
    parts = _FIELD_PARTIALS[0]
    hip.check(_lib().xen_field_sum_partials(
        words_ptr, n, word_bytes, shift, width, parts.ptr, FIELD_SUM_BLOCKS,
        _s(stream)), "xen_field_sum_partials")
    host = np.empty(FIELD_SUM_BLOCKS, dtype=np.uint64)
    parts.download(host, stream)
    return int(host.sum(dtype=np.uint64))
