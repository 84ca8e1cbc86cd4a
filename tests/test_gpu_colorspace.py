"""RGB <-> YCbCr conversion and plane exchange on the GPU, bit-exact against the reference's outputs
(tests/golden/colorspace.json: xColorSpaceSTD::ConvertRGB2YCbCr_I32 / ConvertYCbCr2RGB_I32 on
seeded planes) and against the oracle on larger and ragged pictures, in place and out of place."""
import hashlib
import json
import os

import numpy as np
import pytest

from ivpsnr_amd.capi import QmivError

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "colorspace.json")
with open(GOLDEN) as f:
    COLOUR = json.load(f)


def as_storage(planes, bd):
    return np.ascontiguousarray(planes, np.uint8 if bd <= 8 else np.uint16)


def to_planes(pic, h, w):
    return pic.download().astype(np.uint16).reshape(3, h, w)


@pytest.mark.parametrize("case", COLOUR["cases"], ids=lambda c: f"b{c['bit_depth']}s{c['space']}r{c['to_rgb']}")
def test_matches_reference_outputs_in_place(engine, case):
    bd, w, h = case["bit_depth"], case["width"], case["height"]
    src = np.random.default_rng(case["seed"]).integers(0, 1 << bd, size=(3, h, w), dtype=np.uint16)
    pic = engine.picture(w, h, bd, 444, data=as_storage(src, bd))
    engine.colorspace(pic, case["space"], bool(case["to_rgb"]))
    out = to_planes(pic, h, w)
    assert out.reshape(-1)[:8].tolist() == case["head"]
    assert hashlib.sha256(out.astype("<u2").tobytes()).hexdigest() == case["sha256_out"]


@pytest.mark.parametrize("w,h,bd", [(1920, 1080, 10), (4096, 2048, 16), (129, 65, 14), (57, 41, 8), (1, 1, 12)])
@pytest.mark.parametrize("to_rgb", [False, True])
def test_matches_oracle_out_of_place(engine, oracle, w, h, bd, to_rgb):
    rng = np.random.default_rng(w * 7 + h + bd)
    src = rng.integers(0, 1 << bd, size=(3, h, w), dtype=np.uint16)
    a = engine.picture(w, h, bd, 444, data=as_storage(src, bd))
    b = engine.picture(w, h, bd, 444)
    for space in range(4):
        engine.colorspace(a, space, to_rgb, dst=b)
        np.testing.assert_array_equal(to_planes(b, h, w), oracle.colorspace(src, bd, space, to_rgb))
    np.testing.assert_array_equal(to_planes(a, h, w), src)


@pytest.mark.parametrize("space", range(4))
def test_round_trip_14bit_gradients(engine, space):
    """RGB -> YCbCr -> RGB of horizontal gradients returns every sample within 1
    (test-core-xColorspace.cpp:107-125)."""
    w, h = 129, 65
    x = np.arange(w, dtype=np.int64)
    src = np.stack([np.broadcast_to((x + off) & 0x3FFF, (h, w)) for off in (0, 100, 200)]).astype(np.uint16)
    pic = engine.picture(w, h, 14, 444, data=src)
    engine.colorspace(pic, space, False)
    engine.colorspace(pic, space, True)
    assert np.abs(to_planes(pic, h, w).astype(np.int32) - src).max() <= 1


@pytest.mark.parametrize("w,h,bd", [(64, 32, 8), (57, 41, 10), (1920, 1080, 16)])
def test_swap_planes(engine, w, h, bd):
    src = np.random.default_rng(w + bd).integers(0, 1 << bd, size=(3, h, w), dtype=np.uint16)
    pic = engine.picture(w, h, bd, 444, data=as_storage(src, bd))
    engine.swap_planes(pic, 0, 2)        # BGR -> RGB
    np.testing.assert_array_equal(to_planes(pic, h, w), src[[2, 1, 0]])
    engine.swap_planes(pic, 1, 1)
    np.testing.assert_array_equal(to_planes(pic, h, w), src[[2, 1, 0]])
    engine.swap_planes(pic, 0, 1)
    np.testing.assert_array_equal(to_planes(pic, h, w), src[[1, 2, 0]])


def test_rejects_bad_arguments(engine):
    p420 = engine.picture(64, 32, 10, 420)
    p444 = engine.picture(64, 32, 10, 444)
    other = engine.picture(64, 32, 8, 444)
    with pytest.raises(QmivError):
        engine.colorspace(p420, 1, False)
    with pytest.raises(QmivError):
        engine.colorspace(p444, 4, False)
    with pytest.raises(QmivError):
        engine.colorspace(p444, -1, True)
    with pytest.raises(QmivError):
        engine.colorspace(p444, 1, False, dst=other)
    with pytest.raises(QmivError):
        engine.swap_planes(p444, 0, 3)
    with pytest.raises(QmivError):
        engine.swap_planes(p420, 0, 1)
