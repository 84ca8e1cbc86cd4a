"""Frame sets with their uploads (qmiv_frames_upload): every frame's copies on the copy stream,
overlapping the frames before it, from pinned (qmiv_host_alloc) and from ordinary host memory
(staged through the context's four pinned slots of 8 MiB: a 4096x2048 10-bit frame takes three
chunks per picture, so the slots are reused while earlier transfers may still run).

The results must equal the resident path's (qmiv_frames over pictures uploaded beforehand) field
by field, every double included; one frame is checked against the oracle as well.
"""
import numpy as np
import pytest

from ivpsnr_amd import IVPSNR, IVSSIM, MSSSIM, PSNR, SSIM, WSPSNR, default_params
from ivpsnr_amd.capi import QmivError
from ivpsnr_amd.synth import make_pair
from tests.helpers import assert_frame_equal, oracle_params

pytestmark = pytest.mark.gpu

ALL = PSNR | WSPSNR | IVPSNR | SSIM | MSSSIM | IVSSIM
FRAMES = [(256, 128, 10, 420, 1), (130, 66, 10, 420, 2), (67, 45, 8, 444, 3), (256, 128, 10, 420, 4),
          (1920, 1080, 10, 420, 5), (4096, 2048, 10, 420, 6)]


def _same(a, b):
    da, db = a.as_dict(), b.as_dict()
    assert da.keys() == db.keys()
    for k in da:
        va, vb = np.asarray(da[k], np.float64), np.asarray(db[k], np.float64)
        assert np.array_equal(np.isnan(va), np.isnan(vb)) and np.array_equal(va[~np.isnan(va)], vb[~np.isnan(vb)]), k


def _pics(engine, pairs):
    return [(engine.picture(p.width, p.height, p.bit_depth, p.chroma_format),
             engine.picture(p.width, p.height, p.bit_depth, p.chroma_format)) for p in pairs]


@pytest.mark.parametrize("pinned", [True, False])
@pytest.mark.parametrize("erp", [False, True])
def test_matches_resident_path(engine, pinned, erp):
    print(f"frames_upload pinned={pinned} erp={erp}", flush=True)
    pairs = [make_pair(w, h, bd, s, chroma_format=cf) for w, h, bd, cf, s in FRAMES]
    prm = default_params(erp=erp)
    resident = [(engine.picture(p.width, p.height, p.bit_depth, p.chroma_format, data=p.ref),
                 engine.picture(p.width, p.height, p.bit_depth, p.chroma_format, data=p.tst)) for p in pairs]
    want = engine.frames(resident, ALL, prm)
    srcs = []
    for p in pairs:
        a, b = np.ascontiguousarray(p.ref).view(np.uint8).reshape(-1), np.ascontiguousarray(p.tst).view(np.uint8).reshape(-1)
        if pinned:
            ha, hb = engine.host_array(a.nbytes), engine.host_array(b.nbytes)
            ha[:] = a
            hb[:] = b
            a, b = ha, hb
        srcs.append((a, b))
    pics = _pics(engine, pairs)
    # twice in a row into the same pictures: the second set's copies follow the first set's kernels
    for rep in range(2):
        got = engine.frames_upload(pics, srcs, ALL, prm)
        for i, (g, w_) in enumerate(zip(got, want)):
            print(f"  rep {rep} frame {i} {FRAMES[i]}", flush=True)
            _same(g, w_)


def test_matches_oracle(engine, oracle):
    w, h, bd, cf, s = 130, 66, 10, 420, 7
    pair = make_pair(w, h, bd, s)
    prm = default_params(erp=True)
    pics = _pics(engine, [pair])
    got = engine.frames_upload(pics, [(pair.ref, pair.tst)], PSNR | WSPSNR | IVPSNR, prm)[0]
    orc = oracle.frame(pair.ref, pair.tst, oracle_params(oracle, w, h, bd, cf, PSNR | WSPSNR | IVPSNR, prm))
    assert_frame_equal(got, orc, PSNR | WSPSNR | IVPSNR, "frames_upload")


def test_picture_twice_is_rejected(engine):
    pair = make_pair(64, 48, 8, 1)
    p0, p1 = _pics(engine, [pair])[0]
    with pytest.raises(QmivError, match="twice"):
        engine.frames_upload([(p0, p1), (p0, p1)], [(pair.ref, pair.tst)] * 2, PSNR)


def test_pageable_uploads_through_staging(engine):
    """qmiv_pic_upload from ordinary host memory goes through the pinned staging slots of 8 MiB:
    4000x2000 10-bit pictures (24.0 MB: two full chunks and a partial one, whose last 1 MiB piece
    is partial too) come back unchanged several times in a row, and so does a source starting at an
    odd address."""
    rng = np.random.default_rng(5)
    w, h = 4000, 2000
    pics = [engine.picture(w, h, 10, 420) for _ in range(2)]
    for rep in range(3):
        datas = [rng.integers(0, 1024, w * h * 3 // 2, dtype=np.uint16) for _ in pics]
        for p, d in zip(pics, datas):
            p.upload(d)
        for p, d in zip(pics, datas):
            assert np.array_equal(p.download(), d), rep
    w8, h8 = 1920, 1080
    raw = rng.integers(0, 256, w8 * h8 * 3 // 2 + 1, dtype=np.uint8)
    src = raw[1:]
    assert src.ctypes.data % 2 == 1
    p = engine.picture(w8, h8, 8, 420)
    p.upload(src)
    assert np.array_equal(p.download(), src)


def test_large_unpack_through_staging(engine):
    """qmiv_pic_upload_unpack of a raw 4:2:2 10-bit frame of 32 MB from ordinary memory: the
    copy goes through the pinned staging slots in four chunks (the last partial) into the device
    staging buffer, then the unpack; the 4:4:4 picture holds the luma and the chroma repeated
    horizontally (libs/pmbb-core/src/xSeq.cpp:197-229)."""
    w, h, bd = 4000, 2000, 10
    rng = np.random.default_rng(3)
    y = rng.integers(0, 1 << bd, (h, w), dtype=np.uint16)
    cb = rng.integers(0, 1 << bd, (h, w // 2), dtype=np.uint16)
    cr = rng.integers(0, 1 << bd, (h, w // 2), dtype=np.uint16)
    frame = np.concatenate([y.ravel(), cb.ravel(), cr.ravel()])
    assert frame.nbytes > 3 * (8 << 20)
    p = engine.picture(w, h, bd, 444).upload_unpack(frame, 422)
    got = p.download().reshape(3, h, w)
    assert np.array_equal(got[0], y)
    assert np.array_equal(got[1], np.repeat(cb, 2, axis=1)) and np.array_equal(got[2], np.repeat(cr, 2, axis=1))
