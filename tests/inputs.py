"""Seeded inputs shared by the golden-vector writer (tests/golden/make_golden.py) and the tests:
raw-file frames of the chroma formats the reference's reader unpacks (4:2:2, 4:2:0 of odd width,
4:0:0 masks) and pictures holding pels above their bit depth."""
from __future__ import annotations

import numpy as np

from ivpsnr_amd.synth import file_frame, inject_out_of_range, make_mask, make_pair, split_planes

FORMAT_CASES = [  # w, h, bd, file chroma format, seed
    (64, 48, 8, 422, 31), (130, 66, 10, 422, 32), (67, 45, 8, 422, 33),
    (67, 44, 8, 420, 34), (67, 44, 10, 420, 35), (65, 46, 10, 420, 36),
]
MASK400_CASES = [(64, 48, 10, 8, 41), (96, 40, 8, 16, 42)]  # w, h, bd, mask bd, seed (pictures 4:2:0)
RANGE_CASES = [  # w, h, bd, chroma format, seed, inject P0, inject P1, mask bd (0: none)
    (96, 64, 10, 420, 51, True, True, 0), (67, 45, 10, 444, 52, True, False, 0),
    (130, 66, 12, 420, 53, False, True, 0), (96, 64, 10, 420, 54, True, False, 10),
]


def format_pair(w, h, bd, fcf, seed):
    """(file frame of P0, of P1) in file chroma format fcf (uint8 bytes), from a 4:4:4 synthetic pair."""
    pair = make_pair(w, h, bd, seed, chroma_format=444)
    return tuple(file_frame(np.stack(split_planes(a, w, h, 444)), bd, fcf) for a in (pair.ref, pair.tst))


def mask400_frame(w, h, bdm, seed):
    """A 4:0:0 mask file frame (uint8 bytes): the luma plane of a soft object mask."""
    m = make_mask(w, h, bdm, seed, chroma_format=444, kind="soft")
    return file_frame(np.stack(split_planes(m, w, h, 444)), bdm, 400)


def range_pair(w, h, bd, cf, seed, inj0, inj1):
    """A synthetic pair (packed frames) with pels above the bit depth in P0 and / or P1."""
    pair = make_pair(w, h, bd, seed, chroma_format=cf)
    a = inject_out_of_range(pair.ref, bd, seed) if inj0 else pair.ref
    b = inject_out_of_range(pair.tst, bd, seed + 1) if inj1 else pair.tst
    return a, b


def range_mask(w, h, bdm, cf, seed):
    """A soft mask (packed frame) with 50 pels above its bit depth."""
    return inject_out_of_range(make_mask(w, h, bdm, seed, chroma_format=cf, kind="soft"), bdm, seed + 2, count=50)
