"""Host-side combination of chunk CRC registers equals zlib's CRC-32."""
import zlib

import numpy as np
import pytest

from xentica_amd.runtime.crc import combine_partials, raw_crc32


@pytest.mark.parametrize("total,chunk", [(1, 1), (100, 7), (4096, 64),
                                         (5000, 128), (777, 1000)])
def test_combine_matches_zlib(total, chunk):
    rng = np.random.default_rng(total)
    data = rng.integers(0, 256, total, dtype=np.uint8).tobytes()
    parts = [raw_crc32(data[i:i + chunk]) for i in range(0, total, chunk)]
    assert combine_partials(parts, chunk, total) == zlib.crc32(data)


def test_zero_one_planes():
    rng = np.random.default_rng(9)
    data = rng.integers(0, 2, 3000, dtype=np.uint8).tobytes()
    parts = [raw_crc32(data[i:i + 96]) for i in range(0, 3000, 96)]
    assert combine_partials(parts, 96, 3000) == zlib.crc32(data)
