"""Sequences read into device pictures (qmiv_seq_read): pinned buffers, one-frame prefetch on a
background thread and asynchronous copies ordered on the engine stream. Frames land bit-exact in
the pictures, in any access order, and metrics over sequence-fed pictures equal those over
uploaded arrays."""
import numpy as np
import pytest

from ivpsnr_amd import IVPSNR, PSNR, WSPSNR, Sequence, default_params
from ivpsnr_amd.synth import make_pair

pytestmark = pytest.mark.gpu


def write_pairs(tmp_path, w, h, bd, n):
    pairs = [make_pair(w, h, bd, seed=100 + i) for i in range(n)]
    f0, f1 = tmp_path / "p0.yuv", tmp_path / "p1.yuv"
    f0.write_bytes(b"".join(p.ref.tobytes() for p in pairs))
    f1.write_bytes(b"".join(p.tst.tobytes() for p in pairs))
    return pairs, str(f0), str(f1)


@pytest.mark.parametrize("w,h,bd", [(256, 128, 10), (1920, 1080, 8)])
def test_frames_land_bit_exact_in_any_order(engine, tmp_path, w, h, bd):
    pairs, f0, _ = write_pairs(tmp_path, w, h, bd, 5)
    seq = Sequence(f0, w, h, bd, 420)
    pics = [engine.picture(w, h, bd, 420) for _ in range(3)]
    order = [0, 1, 2, 3, 4, 2, 0, 4, 4, 1]
    for k, i in enumerate(order):
        print(f"read frame {i} into picture {k % 3}", flush=True)
        seq.read(engine, i, pics[k % 3])
        engine.sync()
        np.testing.assert_array_equal(pics[k % 3].download(), pairs[i].ref)
    seq.close()


def test_sequence_fed_metrics_equal_uploaded(engine, tmp_path):
    w, h, bd, n = 640, 360, 10, 6
    pairs, f0, f1 = write_pairs(tmp_path, w, h, bd, n)
    s0, s1 = Sequence(f0, w, h, bd, 420), Sequence(f1, w, h, bd, 420)
    a, b = engine.picture(w, h, bd, 420), engine.picture(w, h, bd, 420)
    prm = default_params(erp=True)
    m = PSNR | WSPSNR | IVPSNR
    got = []
    for i in range(n):
        s0.read(engine, i, a)   # ordered after the previous frame's kernels on the same stream
        s1.read(engine, i, b)
        got.append(engine.frame(a, b, m, prm))
    for i, p in enumerate(pairs):
        u0 = engine.picture(w, h, bd, 420, data=p.ref)
        u1 = engine.picture(w, h, bd, 420, data=p.tst)
        want = engine.frame(u0, u1, m, prm)
        assert got[i].ivpsnr == want.ivpsnr and list(got[i].psnr) == list(want.psnr)
        assert list(got[i].wspsnr) == list(want.wspsnr) and list(got[i].gcd) == list(want.gcd)
    s0.close()
    s1.close()


def test_geometry_mismatch_rejected(engine, tmp_path):
    _, f0, _ = write_pairs(tmp_path, 64, 32, 8, 2)
    seq = Sequence(f0, 64, 32, 8, 420)
    from ivpsnr_amd.capi import QmivError
    with pytest.raises(QmivError):
        seq.read(engine, 0, engine.picture(64, 32, 10, 420))
    with pytest.raises(QmivError):
        seq.read(engine, 2, engine.picture(64, 32, 8, 420))
    seq.close()


def test_png_frames_land_in_device_pictures(engine, tmp_path):
    """PNG-list sequences (qmiv_seq_open_png) read into 8-bit 4:4:4 device pictures through the same
    pinned buffers and prefetch: every frame equals its host decode and the pixels it was encoded from."""
    from tests.png_writer import expected_rgb8, png_bytes, random_image
    w, h = 96, 40
    want = []
    for i in range(4):
        colour = (0, 2, 3, 6)[i]
        pix, pal = random_image(w, h, colour, 900 + i)
        (tmp_path / f"v{i + 1}.png").write_bytes(png_bytes(pix, colour, palette=pal, interlace=i % 2 == 1))
        want.append(expected_rgb8(pix, colour, pal).reshape(-1))
    seq = Sequence.png(str(tmp_path / "v{}.png"), w, h)
    assert len(seq) == 4
    pic = engine.picture(w, h, 8, 444)
    for i in (0, 1, 2, 3, 1):
        print(f"png frame {i}", flush=True)
        seq.read(engine, i, pic)
        engine.sync()
        np.testing.assert_array_equal(pic.download(), want[i])
        np.testing.assert_array_equal(seq.read_host(i), want[i])
    seq.close()
