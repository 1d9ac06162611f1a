"""
CRC-32 of device-resident planes, combined on the host.

The device kernels (``xen_crc32_*_partials``) return the zero-initialised,
un-inverted CRC-32 register of every fixed-size chunk.  CRC-32 is linear
over GF(2), so the register of a concatenation is
``shift(reg(A), len(B)) ^ reg(B)``, where ``shift`` advances a register
through zero bytes -- the method zlib's ``crc32_combine`` publishes.  The
standard checksum (``binascii.crc32``, used by every reference test,
``tests/core/test_base.py:39-41``) then applies the 0xFFFFFFFF initial
value and final inversion.  Chunk registers are merged pairwise in a
vectorised tree.
"""
import functools

import numpy as np

__all__ = ['combine_partials', 'shift_register', 'raw_crc32']

_POLY = 0xEDB88320


def _advance_one_zero_byte(reg):
    for _ in range(8):
        reg = (reg >> 1) ^ (_POLY if reg & 1 else 0)
    return reg


@functools.lru_cache(maxsize=None)
def _zero_byte_columns():
    """Operator of one zero byte, as the images of the 32 unit registers."""
    return np.array([_advance_one_zero_byte(1 << bit) for bit in range(32)],
                    dtype=np.uint64)


def _apply(columns, regs):
    regs = np.asarray(regs, dtype=np.uint64)
    out = np.zeros_like(regs)
    for bit in range(32):
        out ^= np.where((regs >> np.uint64(bit)) & np.uint64(1),
                        columns[bit], np.uint64(0))
    return out


def _compose(outer, inner):
    """Operator ``outer . inner`` (apply ``inner`` first)."""
    return _apply(outer, inner)


@functools.lru_cache(maxsize=None)
def _shift_operator(nbytes):
    """Operator advancing a register through ``nbytes`` zero bytes."""
    result = None
    power = _zero_byte_columns()
    while nbytes:
        if nbytes & 1:
            result = power if result is None else _compose(power, result)
        nbytes >>= 1
        if nbytes:
            power = _compose(power, power)
    if result is None:
        return np.array([1 << bit for bit in range(32)], dtype=np.uint64)
    return result


def shift_register(regs, nbytes):
    """Advance CRC registers through ``nbytes`` zero bytes."""
    return _apply(_shift_operator(int(nbytes)), regs)


def raw_crc32(data, reg=0):
    """Zero-initialised register of ``data`` (reference implementation)."""
    table = [0] * 256
    for byte in range(256):
        c = byte
        for _ in range(8):
            c = (c >> 1) ^ (_POLY if c & 1 else 0)
        table[byte] = c
    for byte in bytes(data):
        reg = table[(reg ^ byte) & 0xFF] ^ (reg >> 8)
    return reg


def combine_partials(partials, chunk, total):
    """
    ``binascii.crc32`` of a ``total``-byte plane from its chunk registers.

    :param partials: register of every ``chunk`` bytes (last may be short)
    :param chunk: chunk length in bytes
    :param total: plane length in bytes
    """
    regs = np.asarray(partials, dtype=np.uint64)
    if total == 0:
        return 0
    lengths = np.full(regs.shape, chunk, dtype=np.int64)
    lengths[-1] = total - chunk * (len(regs) - 1)
    while len(regs) > 1:
        carry = None
        if len(regs) % 2:
            carry = (regs[-1:], lengths[-1:])
            regs, lengths = regs[:-1], lengths[:-1]
        left, right = regs[0::2], regs[1::2]
        llen, rlen = lengths[0::2], lengths[1::2]
        shifted = np.empty_like(left)
        for value in np.unique(rlen):
            sel = rlen == value
            shifted[sel] = shift_register(left[sel], int(value))
        regs = shifted ^ right
        lengths = llen + rlen
        if carry is not None:
            regs = np.concatenate([regs, carry[0]])
            lengths = np.concatenate([lengths, carry[1]])
    reg = int(regs[0]) ^ int(shift_register([0xFFFFFFFF], total)[0])
    return reg ^ 0xFFFFFFFF
