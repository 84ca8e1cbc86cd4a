"""Raw-file formats, invalid pels and the reference's picture layouts on the engine, against golden
vectors of the reference's own code (tests/golden/formats.json) and the oracle.

* 4:2:2, odd-width 4:2:0 and 4:0:0 frames unpacked on the device (qmiv_pic_upload_unpack, and the
  sequence reader) into 4:4:4 pictures: the reference reader's pictures bit for bit
  (libs/pmbb-core/src/xSeq.cpp:179-249), then every metric.
* 4:0:0 masks in masked mode.
* InvalidPelActn (apps/QMIV/src/xAppQMIV.cpp:681-683,753-770): STOP returns QMIV_ERR_PEL_RANGE;
  WARN and SKIP compute over the samples as read (the fast search when the frame is in range, the
  general one when it is not); CNCL clips the failing pictures in place and computes over them;
  in single frames, frame sets and captured batches.
* xPicP / xPicI layouts with margins (qmiv_pic_upload_planar / _interleaved and the downloads),
  xPicP::check / conceal / equalCmp / copy.
PSNR family, GCD and CPS sums bit-exact; SSIM family within 1e-9 relative.
"""
import json
import os

import numpy as np
import pytest

from ivpsnr_amd import ALL_METRICS, IVPSNR, PSNR, SSIM, WSPSNR, Sequence, default_params
from ivpsnr_amd.capi import PEL_CNCL, PEL_SKIP, PEL_STOP, PEL_WARN, QmivError
from ivpsnr_amd.synth import make_pair, split_planes
from tests.inputs import FORMAT_CASES, MASK400_CASES, RANGE_CASES, format_pair, mask400_frame, range_mask, range_pair

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLDEN, "formats.json")) as f:
    FX = json.load(f)
PSNRS = PSNR | WSPSNR | IVPSNR
REL = 1e-9


def close(a, b, rel=REL):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nan = np.isnan(b)
    return bool(np.array_equal(np.isnan(a), nan) and
                np.all(np.abs(a[~nan] - b[~nan]) <= rel * np.maximum(np.abs(b[~nan]), 1e-300)))


def check_frame(r, want, metrics, ctx):
    d = r.as_dict()
    for k in ("in_range", "exact"):
        assert d[k] == want[k], (ctx, k, d[k], want[k])
    if metrics & (IVPSNR | 32):
        assert d["gcd"] == want["gcd"], (ctx, d["gcd"], want["gcd"])
    if metrics & PSNR:
        assert d["psnr"] == want["psnr"] and d["psnr_pic"] == want["psnr_pic"], ctx
    if metrics & WSPSNR:
        assert d["wspsnr"] == want["wspsnr"] and d["wspsnr_pic"] == want["wspsnr_pic"], ctx
    if metrics & IVPSNR:
        assert (d["ivpsnr"], d["ivpsnr_r2t"], d["ivpsnr_t2r"]) == (want["ivpsnr"], want["r2t"], want["t2r"]), ctx
        assert d["cps_sum"] == want["cps_rows_sum"], ctx
    if metrics & SSIM:
        for k in ("ssim", "msssim", "ivssim", "ivssim_r2t", "ivssim_t2r"):
            assert close(d[k], want[k]), (ctx, k, d[k], want[k])


@pytest.mark.parametrize("i", range(len(FORMAT_CASES)), ids=[f"{c[0]}x{c[1]}b{c[2]}cf{c[3]}" for c in FORMAT_CASES])
def test_file_formats_unpack_on_device(engine, oracle, i):
    w, h, bd, fcf, seed = FORMAT_CASES[i]
    print(f"format {w}x{h} bd{bd} cf{fcf}", flush=True)
    a, b = format_pair(w, h, bd, fcf, seed)
    p0 = engine.picture(w, h, bd, 444).upload_unpack(a, fcf)
    p1 = engine.picture(w, h, bd, 444).upload_unpack(b, fcf)
    for pic, raw in ((p0, a), (p1, b)):
        got = pic.download().astype(np.uint16).reshape(3, h, w)
        assert np.array_equal(got, oracle.unpack(raw, w, h, bd, fcf))
    r = engine.frame(p0, p1, ALL_METRICS, default_params())
    check_frame(r, FX["formats"][i]["frame"], ALL_METRICS, f"format {i}")
    p0.close()
    p1.close()


def test_sequence_reader_unpacks_file_formats(engine, oracle, tmp_path):
    """qmiv_seq_read of 4:2:2 and odd-width 4:2:0 files into 4:4:4 pictures, two frames each."""
    for (w, h, bd, fcf, seed) in (FORMAT_CASES[1], FORMAT_CASES[5]):
        frames = [format_pair(w, h, bd, fcf, seed + k)[0] for k in range(2)]
        path = tmp_path / f"s{fcf}_{w}x{h}.yuv"
        path.write_bytes(b"".join(f.tobytes() for f in frames))
        seq = Sequence(str(path), w, h, bd, fcf)
        assert len(seq) == 2
        pic = engine.picture(w, h, bd, 444)
        for k in range(2):
            seq.read(engine, k, pic)
            engine.sync()
            assert np.array_equal(pic.download().astype(np.uint16).reshape(3, h, w), oracle.unpack(frames[k], w, h, bd, fcf))
        seq.close()
        pic.close()


@pytest.mark.parametrize("i", range(len(MASK400_CASES)))
@pytest.mark.parametrize("mode", ["plain", "erp"])
def test_mask400(engine, i, mode):
    w, h, bd, bdm, seed = MASK400_CASES[i]
    pair = make_pair(w, h, bd, seed)
    p0 = engine.picture(w, h, bd, 420, data=pair.ref)
    p1 = engine.picture(w, h, bd, 420, data=pair.tst)
    mf = mask400_frame(w, h, bdm, seed)
    pm = engine.picture(w, h, bdm, 400, data=mf.view(np.uint8 if bdm <= 8 else np.uint16))
    r = engine.frame_masked(p0, p1, pm, PSNRS, default_params(erp=mode == "erp")).as_dict()
    want = FX["mask400"][i][mode]
    assert r["num_non_masked"] == want["num_non_masked"] and r["gcd"] == want["gcd"]
    for k in ("psnr", "wspsnr", "psnr_pic", "wspsnr_pic", "ivpsnr"):
        assert r[k] == want[k], k
    assert (r["ivpsnr_r2t"], r["ivpsnr_t2r"]) == (want["r2t"], want["t2r"])
    # a 4:0:0 picture serves as a mask only
    with pytest.raises(QmivError):
        engine.frame(pm, pm, PSNR, default_params())
    for p in (p0, p1, pm):
        p.close()


def _range_inputs(engine, case):
    w, h, bd, cf, seed, inj0, inj1, bdm = case
    a, b = range_pair(w, h, bd, cf, seed, inj0, inj1)
    pics = [engine.picture(w, h, bd, cf, data=a), engine.picture(w, h, bd, cf, data=b)]
    mk = None
    if bdm:
        mk = range_mask(w, h, bdm, cf, seed)
        pics.append(engine.picture(w, h, bdm, cf, data=mk))
    return (a, b, mk), pics


@pytest.mark.parametrize("i", range(len(RANGE_CASES)))
@pytest.mark.parametrize("action", [PEL_WARN, PEL_SKIP, PEL_CNCL, PEL_STOP])
def test_invalid_pel_actions(engine, i, action):
    case = RANGE_CASES[i]
    w, h, bd, cf, seed, inj0, inj1, bdm = case
    print(f"range case {i} action {action}", flush=True)
    (a, b, mk), pics = _range_inputs(engine, case)
    prm = default_params(invalid_pel_action=action)
    want = FX["range"][i]["CNCL" if action == PEL_CNCL else "WARN"]
    if bdm:
        r = engine.frame_masked(pics[0], pics[1], pics[2], PSNRS, prm, raise_on_range=False)
        d = r.as_dict()
        assert d["in_range"] == want["in_range"] + [want["in_range_mask"]]
        if action == PEL_STOP:
            assert r.status == 5
        else:
            assert r.status == 0
            assert d["num_non_masked"] == want["num_non_masked"] and d["gcd"] == want["gcd"]
            for k in ("psnr", "wspsnr", "psnr_pic", "wspsnr_pic", "ivpsnr", "cps_sum"):
                assert d[k] == (want["cps_rows_sum"] if k == "cps_sum" else want[k]), k
    else:
        r = engine.frame(pics[0], pics[1], ALL_METRICS, prm, raise_on_range=False)
        if action == PEL_STOP:
            assert r.status == 5 and list(r.in_range) == want["in_range"]
        else:
            assert r.status == 0
            check_frame(r, want, ALL_METRICS, f"range {i} action {action}")
    # CNCL conceals in place (xPicP::conceal); the other actions leave the pictures as they were
    m = (1 << bd) - 1
    for pic, raw in zip(pics[:2], (a, b)):
        assert np.array_equal(pic.download(), np.minimum(raw, m) if action == PEL_CNCL else raw)
    for p in pics:
        p.close()


def test_invalid_pel_frame_set_and_batch(engine):
    """Frames with and without out-of-range pels in one frame set (two streams) and in one captured
    batch: each frame takes its own search path and equals its single-frame result."""
    w, h, bd = 96, 64, 10
    clean = make_pair(w, h, bd, 60)
    (a, b, _), bad = _range_inputs(engine, RANGE_CASES[0])
    good = [engine.picture(w, h, bd, 420, data=clean.ref), engine.picture(w, h, bd, 420, data=clean.tst)]
    pairs = [tuple(good), tuple(bad), tuple(good), tuple(bad)]
    for action in (PEL_WARN, PEL_SKIP):
        prm = default_params(invalid_pel_action=action)
        single = [engine.frame(p0, p1, ALL_METRICS, prm).as_dict() for p0, p1 in pairs[:2]]
        out = engine.frames(pairs, ALL_METRICS, prm)
        for k, r in enumerate(out):
            assert r.as_dict() == single[k % 2], (action, k)
        bt = engine.batch(pairs[:2], ALL_METRICS, [prm, prm])
        for _ in range(2):
            got = bt.run()
            assert [g.as_dict() for g in got] == single
        bt.close()
    check_frame(engine.frame(bad[0], bad[1], ALL_METRICS, default_params(invalid_pel_action=PEL_WARN)),
                FX["range"][0]["WARN"], ALL_METRICS, "after the sets")
    for p in good + bad:
        p.close()


def test_concealment_in_frame_set(engine):
    w, h, bd, cf, seed, inj0, inj1, _ = RANGE_CASES[0]
    prm = default_params(invalid_pel_action=PEL_CNCL)
    (a, b, _), bad = _range_inputs(engine, RANGE_CASES[0])
    (a2, b2, _), bad2 = _range_inputs(engine, RANGE_CASES[0])
    out = engine.frames([tuple(bad), tuple(bad2)], ALL_METRICS, prm)
    for r in out:
        check_frame(r, FX["range"][0]["CNCL"], ALL_METRICS, "concealed frame set")
    for p in bad + bad2:
        p.close()


def test_reference_layouts(engine, oracle):
    """xPicP (planar uint16 with a margin) and xPicI (interleaved 4 x uint16 with a margin) uploads
    and downloads, check / conceal / equal / copy; frames computed from them equal the packed path."""
    w, h, m = 70, 42, 4
    for bd in (8, 10):
        pair = make_pair(w, h, bd, 70 + bd, chroma_format=444)
        planes = [np.stack(split_planes(x, w, h, 444)).astype(np.uint16) for x in (pair.ref, pair.tst)]
        rng = np.random.default_rng(bd)
        with_margin = []
        for pl in planes:
            arr = rng.integers(0, 1 << bd, size=(3, h + 2 * m, w + 2 * m)).astype(np.uint16)  # margin: noise
            arr[:, m:m + h, m:m + w] = pl
            with_margin.append(arr)
        p0 = engine.picture(w, h, bd, 444).upload_planar(with_margin[0], m)
        p1 = engine.picture(w, h, bd, 444).upload_planar(with_margin[1], m)
        q0 = engine.picture(w, h, bd, 444, data=pair.ref)
        q1 = engine.picture(w, h, bd, 444, data=pair.tst)
        assert engine.frame(p0, p1, ALL_METRICS).as_dict() == engine.frame(q0, q1, ALL_METRICS).as_dict()
        back = p0.download_planar(m)
        assert np.array_equal(back[:, m:m + h, m:m + w], planes[0]) and (back[:, :m] == 0xFFFF).all()
        inter = np.zeros((h + 2 * m, w + 2 * m, 4), np.uint16)
        inter[m:m + h, m:m + w, :3] = planes[1].transpose(1, 2, 0)
        inter[m:m + h, m:m + w, 3] = 7  # A is ignored
        i1 = engine.picture(w, h, bd, 444).upload_interleaved(inter, m)
        assert i1.equal(q1) == [1, 1, 1]
        got = i1.download_interleaved(m)
        assert np.array_equal(got[m:m + h, m:m + w, :3], inter[m:m + h, m:m + w, :3])
        assert (got[m:m + h, m:m + w, 3] == 0).all() and (got[:m] == 0xFFFF).all()
        c = engine.picture(w, h, bd, 444).copy_from(p0)
        assert c.equal(p0) == [1, 1, 1] and c.equal(p1) != [1, 1, 1]
        for p in (p0, p1, q0, q1, i1, c):
            p.close()
    # check and conceal of a 10-bit picture holding pels above 1023 in its Cb plane only
    pl = np.zeros((3, h, w), np.uint16)
    pl[1, 5, 7], pl[1, 9, 1] = 2000, 1500
    p = engine.picture(w, h, 10, 444).upload_planar(pl)
    assert p.check_range() == [1, 0, 1]
    p.conceal()
    assert p.check_range() == [1, 1, 1] and p.download().reshape(3, h, w)[1, 5, 7] == 1023
    with pytest.raises(QmivError) as ei:  # an 8-bit picture holds samples up to 255
        engine.picture(w, h, 8, 444).upload_planar(pl)
    assert ei.value.status == 5
    p.close()
