"""
Checkpoints in the reference's format (``xentica/core/base.py:755-771``).

A model's ``save()`` pickles ``{"cells": <the whole cells array>,
"colors": <the int3 colours>, "random": <its LocalRandom>}``: ``cells``
is the reference layout, ``n * (k + 1) + 1`` container words of which
plane 0 is the main state.  The packed fields and the batches write the
same dictionary, so a file moves freely between a model and the fast
engines: a packed field saves its state as plane 0 of a uint8 cells array
with zero buffer planes (``k`` = 8 in 2-D, 26 in 3-D for the Moore
neighbourhoods these fields run), zero colours and a fresh random stream,
and loads plane 0 of any such file; a batch saves the list of its
instances' dictionaries under ``"instances"``.
"""
import pickle

import numpy as np

from xentica_amd.core.exceptions import XenticaException

__all__ = ['field_state', 'write', 'read', 'main_plane', 'MOORE_BUFFERS']

#: Buffer planes of the Moore neighbourhood in 2-D and 3-D (the 3-D list
#: keeps the centre and drops (1, 1, 1): 26 offsets).
MOORE_BUFFERS = {2: 8, 3: 26}


def field_state(plane, buffers, random=None):
    """The reference's checkpoint dictionary of a 1-bit field whose main
    plane (one uint8 per cell) is ``plane``."""
    from xentica_amd.seeds.random import LocalRandom
    plane = np.ascontiguousarray(plane, dtype=np.uint8).reshape(-1)
    cells = np.zeros(plane.size * (buffers + 1) + 1, dtype=np.uint8)
    cells[:plane.size] = plane
    return {"cells": cells,
            "colors": np.zeros(3 * plane.size, dtype=np.int32),
            "random": random if random is not None else LocalRandom()}


def write(filename, state):
    """Pickle a checkpoint dictionary (or a batch's) to ``filename``."""
    with open(filename, "wb") as out:
        pickle.dump(state, out)


def read(filename):
    """A checkpoint written by :func:`write` or a model's ``save()``."""
    with open(filename, "rb") as src:
        return pickle.load(src)


def main_plane(state, count):
    """Plane 0 (``count`` cells) of a checkpoint's cells array."""
    cells = np.asarray(state["cells"]).reshape(-1)
    if cells.size < count:
        raise XenticaException(
            "checkpoint holds %d cells, %d expected" % (cells.size, count))
    return cells[:count]
