"""Raw planar YUV sequences (qmiv_seq_*, the read side of xSeq): frame size and count from the file
size with a partial last frame ignored (libs/pmbb-core/src/xSeq.cpp:408-428,451-470), random access
by frame (:519-525), 8- and 16-bit samples, 4:2:0 and 4:4:4, and the error paths. Host-only: these
run without a GPU."""
import numpy as np
import pytest

from ivpsnr_amd import Sequence
from ivpsnr_amd.capi import QmivError


def write_frames(path, w, h, bd, cf, n, extra=0, seed=0):
    dt = np.uint8 if bd <= 8 else np.uint16
    per = w * h * (3 if cf == 444 else 1) + (0 if cf == 444 else 2 * (w // 2) * (h // 2))
    rng = np.random.default_rng(seed)
    frames = [rng.integers(0, 1 << bd, size=per, dtype=np.int64).astype(dt) for _ in range(n)]
    with open(path, "wb") as f:
        for fr in frames:
            f.write(fr.astype("<u2" if bd > 8 else np.uint8).tobytes())
        f.write(b"\x07" * extra)
    return frames


@pytest.mark.parametrize("w,h,bd,cf", [(64, 32, 8, 420), (130, 66, 10, 420), (67, 45, 16, 444), (2, 2, 12, 420)])
def test_frame_count_and_random_access(tmp_path, w, h, bd, cf):
    path = tmp_path / "s.yuv"
    bps = 1 if bd <= 8 else 2
    per = (w * h * 3 if cf == 444 else w * h * 3 // 2) * bps
    frames = write_frames(path, w, h, bd, cf, 5, extra=per - 1)  # a partial sixth frame
    seq = Sequence(str(path), w, h, bd, cf)
    assert seq.frame_bytes == per
    assert len(seq) == 5  # floor(size / frame bytes)
    for i in (3, 0, 4, 4, 1):
        np.testing.assert_array_equal(seq.read_host(i), frames[i])
    seq.close()


def test_empty_and_short_files(tmp_path):
    (tmp_path / "e.yuv").write_bytes(b"")
    assert len(Sequence(str(tmp_path / "e.yuv"), 64, 32, 8, 420)) == 0
    (tmp_path / "p.yuv").write_bytes(b"\x00" * (64 * 32 * 3 // 2 - 1))
    seq = Sequence(str(tmp_path / "p.yuv"), 64, 32, 8, 420)
    assert len(seq) == 0
    with pytest.raises(QmivError):
        seq.read_host(0)


def test_errors(tmp_path):
    with pytest.raises(QmivError):
        Sequence(str(tmp_path / "missing.yuv"), 64, 32, 8, 420)
    path = tmp_path / "s.yuv"
    write_frames(path, 64, 32, 8, 420, 2)
    with pytest.raises(QmivError):
        Sequence(str(path), 64, 32, 8, 411)  # not a chroma format
    odd = Sequence(str(path), 63, 32, 8, 420)  # odd widths open as the reference's reader does
    assert odd.frame_bytes == 63 * 32 + (63 * 32) // 2
    odd.close()
    with pytest.raises(QmivError):
        Sequence(str(path), 64, 32, 7, 420)
    seq = Sequence(str(path), 64, 32, 8, 420)
    with pytest.raises(QmivError):
        seq.read_host(2)
    with pytest.raises(QmivError):
        seq.read_host(-1)
