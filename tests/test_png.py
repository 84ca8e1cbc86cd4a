"""PNG-list sequences (xSeqPNG, libs/pmbb-spng/src/xSeqPNG.cpp) without a GPU: the engine's decoder
against the reference's decoder (tests/golden/png.json: libspng as xSeqPNG calls it, on every colour
type at bit depth 8, both interlace methods, each scanline filter, odd sizes and malformed files)
and against the pixels the files were encoded from; frame discovery and file-name patterns."""
import hashlib
import json
import os

import numpy as np
import pytest

from ivpsnr_amd import Sequence
from ivpsnr_amd.capi import QmivError
from tests import png_writer

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "png.json")
with open(GOLDEN) as f:
    PNG = {c["name"]: c for c in json.load(f)["cases"]}
CASES = png_writer.cases()


def decode(tmp_path, name, data, w, h):
    (tmp_path / f"{name}_0.png").write_bytes(data)
    seq = Sequence.png(str(tmp_path / f"{name}_{{}}.png"), w, h)
    try:
        assert len(seq) == 1 and seq.frame_bytes == 3 * w * h
        return seq.read_host(0)
    finally:
        seq.close()


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_decode_matches_reference_and_pixels(tmp_path, case):
    data, (w, h), expected = png_writer.build_case(case)
    want = PNG[case["name"]]
    if not want["ok"]:
        with pytest.raises(QmivError) as e:
            decode(tmp_path, case["name"], data, w, h)
        if want["error"].endswith("does not match"):
            assert want["error"] in str(e.value)
        return
    got = decode(tmp_path, case["name"], data, w, h)
    assert hashlib.sha256(got.tobytes()).hexdigest() == want["sha256"]
    assert expected is not None
    np.testing.assert_array_equal(got.reshape(3, h, w), expected)


def test_golden_covers_every_case():
    assert sorted(PNG) == sorted(c["name"] for c in CASES)
    # the decoder the reference uses rejects exactly these malformed files
    assert sorted(n for n, c in PNG.items() if not c["ok"]) == sorted(
        ["e_truncated", "e_wrong_width", "e_depth16", "e_ihdr_bad_crc", "e_first_idat_bad_crc",
         "e_unknown_critical_before_idat"])


def frame_png(w, h, value):
    pix = np.full((h, w, 3), value, np.uint8)
    pix[0, 0] = (value, 255 - value, 7)
    return png_writer.png_bytes(pix, 2)


def test_discovery_starts_at_the_last_of_0_and_1(tmp_path):
    """xSeqPNG::xBackendOpen (xSeqPNG.cpp:55-78): with files 0..4 the sequence starts at file 1."""
    for i in range(5):
        (tmp_path / f"a{i:03d}.png").write_bytes(frame_png(8, 4, 10 * i))
    seq = Sequence.png(str(tmp_path / "a{:03d}.png"), 8, 4)
    assert len(seq) == 4
    assert [int(seq.read_host(k)[1]) for k in range(4)] == [10, 20, 30, 40]
    seq.close()


def test_discovery_from_1_and_stops_at_a_gap(tmp_path):
    for i in (1, 2, 3, 5):
        (tmp_path / f"b_{i}.png").write_bytes(frame_png(8, 4, i))
    seq = Sequence.png(str(tmp_path / "b_{}.png"), 8, 4)
    assert len(seq) == 3
    assert [int(seq.read_host(k)[1]) for k in range(3)] == [1, 2, 3]
    seq.close()
    with pytest.raises(QmivError, match="First Idx"):
        Sequence.png(str(tmp_path / "c_{}.png"), 8, 4)


@pytest.mark.parametrize("pattern,name", [("p{}.png", "p0.png"), ("p{:d}.png", "p0.png"), ("p{:4d}.png", "p   0.png"),
                                          ("p{0:05d}.png", "p00000.png"), ("{{p}}{:02d}.png", "{p}00.png")])
def test_file_name_fields(tmp_path, pattern, name):
    assert pattern.format(0) == name  # the fields mean what they mean to fmt and to str.format
    (tmp_path / name).write_bytes(frame_png(4, 2, 50))
    seq = Sequence.png(str(tmp_path / pattern), 4, 2)
    assert len(seq) == 1 and int(seq.read_host(0)[8]) == 205  # G of pixel (0, 0)
    seq.close()


def test_unsupported_field_and_read_errors(tmp_path):
    with pytest.raises(QmivError, match="unsupported replacement field"):
        Sequence.png(str(tmp_path / "x{:x}.png"), 4, 2)
    (tmp_path / "y0.png").write_bytes(b"not a png")
    seq = Sequence.png(str(tmp_path / "y{}.png"), 4, 2)
    with pytest.raises(QmivError, match="not a PNG file"):
        seq.read_host(0)
    with pytest.raises(QmivError, match="out of range"):
        seq.read_host(1)
    seq.close()
