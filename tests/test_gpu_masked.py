"""Masked mode on the GPU: the frame flow and the processor-level calls against the reference's
masked processors (tests/golden/masked.json, bit-exact) and against the oracle on larger, ragged
and re-parameterised cases; the full mask against the unmasked flow; error paths."""
import json
import os

import numpy as np
import pytest

from ivpsnr_amd import default_params
from ivpsnr_amd.capi import FrameResult, QmivError, lib
from ivpsnr_amd.synth import make_mask, make_pair, sha256

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "masked.json")
with open(GOLDEN) as f:
    MASKED = json.load(f)
PSNRS = 1 | 2 | 4


def upload(engine, case):
    w, h, bd, cf = case["width"], case["height"], case["bit_depth"], case["chroma_format"]
    pair = make_pair(w, h, bd, case["seed"], chroma_format=cf, **case["knobs"])
    mask = make_mask(w, h, case["mask_bit_depth"], case["seed"], chroma_format=case["mask_chroma_format"],
                     kind=case["mask_kind"])
    assert sha256(pair.ref) == case["sha256_ref"] and sha256(mask) == case["sha256_mask"]
    p0 = engine.picture(w, h, bd, cf, data=pair.ref)
    p1 = engine.picture(w, h, bd, cf, data=pair.tst)
    pm = engine.picture(w, h, case["mask_bit_depth"], case["mask_chroma_format"], data=mask)
    return pair, mask, p0, p1, pm


@pytest.mark.parametrize("case", MASKED, ids=lambda c: f"{c['width']}x{c['height']}b{c['bit_depth']}{c['mask_kind']}")
@pytest.mark.parametrize("mode", ["plain", "erp"])
def test_frame_masked_matches_reference(engine, case, mode):
    print(f"masked {case['width']}x{case['height']} bd{case['bit_depth']} {case['mask_kind']} {mode}", flush=True)
    _, _, p0, p1, pm = upload(engine, case)
    want = case[mode]
    r = engine.frame_masked(p0, p1, pm, PSNRS, default_params(erp=mode == "erp")).as_dict()
    assert r["num_non_masked"] == want["num_non_masked"]
    assert r["gcd"] == want["gcd"]
    assert r["in_range"] == want["in_range"] + [want["in_range_mask"]]
    assert r["exact"] == want["exact"]
    for k in ("psnr", "wspsnr", "psnr_pic", "wspsnr_pic", "ivpsnr"):
        assert r[k] == want[k], k
    assert r["ivpsnr_r2t"] == want["r2t"] and r["ivpsnr_t2r"] == want["t2r"]
    assert r["cps_sum"] == want["cps_rows_sum"]


@pytest.mark.parametrize("case", MASKED[:5], ids=lambda c: f"{c['width']}x{c['height']}b{c['bit_depth']}")
def test_processor_calls_match_reference(engine, case):
    _, mask, p0, p1, pm = upload(engine, case)
    plain, erp = case["plain"], case["erp"]
    nnm = plain["num_non_masked"]
    assert engine.count_nonzero(pm) == nnm
    assert engine.calc_gcd_masked(p0, p1, pm).tolist() == plain["gcd"]
    assert engine.calc_gcd_masked(p0, p1, pm, nnm).tolist() == plain["gcd"]
    assert engine.calc_psnr_masked(p0, p1, pm, nnm).tolist() == plain["psnr"]
    assert engine.calc_psnr_masked(p0, p1, pm).tolist() == plain["psnr"]
    assert engine.calc_wspsnr_masked(p0, p1, pm, nnm).tolist() == plain["wspsnr"]
    assert engine.calc_wspsnr_masked(p0, p1, pm, nnm, default_params(erp=True)).tolist() == erp["wspsnr"]
    auto = engine.calc_ivpsnr_masked(p0, p1, pm)
    assert [auto["ivpsnr"], auto["r2t"], auto["t2r"]] == [plain["ivpsnr"], plain["r2t"], plain["t2r"]]
    given = engine.calc_ivpsnr_masked(p0, p1, pm, plain["gcd"], nnm, default_params(erp=True))
    assert [given["ivpsnr"], given["r2t"], given["t2r"]] == [erp["ivpsnr"], erp["r2t"], erp["t2r"]]


@pytest.mark.parametrize("w,h,bd,cf,bdm,kind,erp", [
    (4096, 2048, 10, 420, 8, "objects", True),
    (1920, 1080, 10, 420, 10, "soft", False),
    (333, 199, 12, 444, 8, "soft", False),
    (2, 2, 8, 420, 8, "full", False),
    (640, 360, 10, 420, 8, "speckle", False),
    (516, 262, 8, 420, 16, "speckle", True),
    (1920, 1080, 10, 420, 10, "speckle", False),
    (130, 66, 10, 420, 8, "zero", False),
])
def test_frame_masked_matches_oracle(engine, oracle, w, h, bd, cf, bdm, kind, erp):
    print(f"masked oracle {w}x{h} bd{bd} cf{cf} {kind}", flush=True)
    pair = make_pair(w, h, bd, 31, chroma_format=cf, luma_offset=4.0)
    mask = make_mask(w, h, bdm, 31, chroma_format=420 if w % 2 == 0 and h % 2 == 0 else 444, kind=kind)
    cfm = 420 if w % 2 == 0 and h % 2 == 0 else 444
    p0 = engine.picture(w, h, bd, cf, data=pair.ref)
    p1 = engine.picture(w, h, bd, cf, data=pair.tst)
    pm = engine.picture(w, h, bdm, cfm, data=mask)
    r = engine.frame_masked(p0, p1, pm, PSNRS, default_params(erp=erp)).as_dict()
    o = oracle.frame_masked(pair.ref, pair.tst, mask, bdm, cfm,
                            oracle.params(w, h, bd, cf, metrics=PSNRS, erp=erp)).as_dict()
    for k in ("num_non_masked", "gcd", "in_range", "exact", "psnr", "wspsnr", "psnr_pic", "wspsnr_pic", "ivpsnr",
              "cps_sum"):
        assert r[k] == o[k], k
    assert r["ivpsnr_r2t"] == o["r2t"] and r["ivpsnr_t2r"] == o["t2r"]


@pytest.mark.parametrize("w,h,bd,bdm,kind,erp", [
    (128, 16, 8, 8, "objects", False), (256, 64, 8, 8, "speckle", True), (384, 40, 10, 8, "soft", False),
    (1024, 34, 10, 10, "speckle", False), (512, 66, 10, 10, "soft", True), (256, 64, 10, 16, "speckle", False),
    (256, 62, 8, 16, "soft", False), (1152, 18, 8, 8, "full", False),
])
def test_sixteen_bit_masked_search_matches_oracle(engine, oracle, w, h, bd, bdm, kind, erp):
    """4:2:0 pictures of 8 and 10 bits whose width is a multiple of 128 run the masked search on
    the 16-bit kernel while every lane's masked sum of a sweep stays below 2^32 (2 pels per row
    and tile, each square times a mask value): 8- and 10-bit masks, heights that leave a partial
    band (40, 34, 66, 18), several segments per row (a 10-bit mask over 10-bit pictures allows one tile per sweep). A
    16-bit mask over these bit depths exceeds the bound and takes the 32-bit kernel: the same
    values."""
    print(f"masked s16 {w}x{h} bd{bd} bdm{bdm} {kind} erp{int(erp)}", flush=True)
    pair = make_pair(w, h, bd, 91, luma_offset=3.0)
    mask = make_mask(w, h, bdm, 91, kind=kind)
    p0 = engine.picture(w, h, bd, 420, data=pair.ref)
    p1 = engine.picture(w, h, bd, 420, data=pair.tst)
    pm = engine.picture(w, h, bdm, 420, data=mask)
    prm = default_params(erp=erp)
    engine.profile(True)
    engine.profile_read(reset=True)
    r = engine.frame_masked(p0, p1, pm, PSNRS, prm).as_dict()
    prof = engine.profile_read(reset=True)
    engine.profile(False)
    # the fast search was planned (a 10-bit frame also enqueues the general search, gated to run
    # only if a pel is out of range, which returns at once here)
    assert prof.get("cps", (0, 0))[1] == 1, prof
    o = oracle.frame_masked(pair.ref, pair.tst, mask, bdm, 420,
                            oracle.params(w, h, bd, 420, metrics=PSNRS, erp=erp)).as_dict()
    for k in ("num_non_masked", "gcd", "in_range", "exact", "psnr", "wspsnr", "psnr_pic", "wspsnr_pic", "ivpsnr",
              "cps_sum"):
        assert r[k] == o[k], k
    assert r["ivpsnr_r2t"] == o["r2t"] and r["ivpsnr_t2r"] == o["t2r"]


@pytest.mark.parametrize("w,h,cf,cfm,bd,bdm", [
    (256, 96, 420, 444, 10, 8), (256, 96, 444, 420, 10, 8), (256, 96, 444, 444, 8, 16), (252, 98, 444, 444, 12, 10),
    (254, 98, 420, 420, 10, 8), (130, 40, 444, 420, 8, 8),
])
def test_chroma_formats_match_oracle(engine, oracle, w, h, cf, cfm, bd, bdm):
    """Every pairing of picture and mask chroma formats, at widths that take the statistics
    kernel's four-pel loads and the search kernels' aligned fills (divisible by 4) and ones that
    do not (254, 130). 8- and 10-bit 4:4:4 pictures take the masked 4:4:4 fast search, 12-bit the
    general one."""
    print(f"masked formats {w}x{h} cf{cf} cfm{cfm} bd{bd} bdm{bdm}", flush=True)
    pair = make_pair(w, h, bd, 61, chroma_format=cf, luma_offset=3.0)
    mask = make_mask(w, h, bdm, 61, chroma_format=cfm, kind="speckle")
    p0 = engine.picture(w, h, bd, cf, data=pair.ref)
    p1 = engine.picture(w, h, bd, cf, data=pair.tst)
    pm = engine.picture(w, h, bdm, cfm, data=mask)
    r = engine.frame_masked(p0, p1, pm, PSNRS, default_params()).as_dict()
    o = oracle.frame_masked(pair.ref, pair.tst, mask, bdm, cfm, oracle.params(w, h, bd, cf, metrics=PSNRS)).as_dict()
    for k in ("num_non_masked", "gcd", "in_range", "exact", "psnr", "wspsnr", "psnr_pic", "wspsnr_pic", "ivpsnr",
              "cps_sum"):
        assert r[k] == o[k], k


@pytest.mark.parametrize("w,bd,cf,g", [(200, 8, 420, [200, -150, 90, 0]), (200, 10, 420, [-300, 250, -280, 0]),
                                       (200, 10, 444, [300, -260, 120, 0]), (256, 8, 420, [200, -150, 90, 0]),
                                       (256, 10, 420, [-300, 250, -280, 0])])
def test_large_given_gcd_matches_oracle(engine, oracle, w, bd, cf, g):
    """A caller-given GCD far beyond the unnoticeable bound: tested values well outside [0, max]
    (no clipping, xc_CLIP_CURR_TST_RANGE = false) still fit the fast kernels' key bounds, so the
    fast masked search runs and must equal the oracle's search. Width 200 takes the 32-bit 4:2:0
    kernel, width 256 (a multiple of its 128-pel tile) the 16-bit one, whose sweep the larger
    differences shorten (10 bits: 4 tiles per lane sum)."""
    h = 72
    pair = make_pair(w, h, bd, 81, chroma_format=cf, luma_offset=1.0)
    mask = make_mask(w, h, 8, 81, chroma_format=cf, kind="speckle")
    P0, P1 = oracle.unpack(pair.ref, w, h, bd, cf), oracle.unpack(pair.tst, w, h, bd, cf)
    M = oracle.unpack(mask, w, h, 8, cf)[0]
    p0 = engine.picture(w, h, bd, cf, data=pair.ref)
    p1 = engine.picture(w, h, bd, cf, data=pair.tst)
    pm = engine.picture(w, h, 8, cf, data=mask)
    engine.profile(True)
    engine.profile_read(reset=True)
    r = engine.calc_ivpsnr_masked(p0, p1, pm, g, -1, default_params())
    prof = engine.profile_read(reset=True)
    engine.profile(False)
    # the fast masked search was planned (a case that drops to the general search enqueues no fast
    # kernel); processor-level calls validate nothing, so the general kernel is enqueued beside it,
    # gated by the range flag, and returns at once for this in-range frame
    assert prof.get("cps", (0, 0))[1] == 1, prof
    nnm = int((M != 0).sum())
    want = []
    for rows in (oracle.cps_rows_masked(P1, P0, M, [-v for v in g]), oracle.cps_rows_masked(P0, P1, M, g)):
        err = [float(rows[c].sum()) for c in range(3)]
        q = [oracle.psnr_from_masked_ssd(e if e > 0 else 1.0, nnm, bd, 8) for e in err]
        want.append((((q[0] * 4 + q[1] * 1) + q[2] * 1) + 0.0 * 0) * (1.0 / 6))
    assert r["r2t"] == want[0] and r["t2r"] == want[1]


def test_frames_masked_equal_single_frames(engine):
    """A masked frame set (qmiv_frames_masked, two streams) gives each frame exactly what the
    single-frame call gives, over frames of different sizes, formats, masks and bit depths; a
    frame with a pel above its bit depth fails alone and the others are still finalised."""
    specs = [(320, 180, 10, 420, 8, "objects"), (256, 96, 8, 444, 16, "speckle"), (130, 66, 10, 420, 10, "soft"),
             (640, 360, 8, 420, 8, "speckle"), (96, 70, 16, 420, 16, "soft")]
    triples = []
    for k, (w, h, bd, cf, bdm, kind) in enumerate(specs):
        pair = make_pair(w, h, bd, 70 + k, chroma_format=cf, luma_offset=2.0)
        mask = make_mask(w, h, bdm, 70 + k, chroma_format=cf, kind=kind)
        triples.append((engine.picture(w, h, bd, cf, data=pair.ref), engine.picture(w, h, bd, cf, data=pair.tst),
                        engine.picture(w, h, bdm, cf, data=mask)))
    for erp in (False, True):
        prm = default_params(erp=erp)
        single = [engine.frame_masked(a, b, m, PSNRS, prm).as_dict() for a, b, m in triples]
        res, st = engine.frames_masked(triples, PSNRS, prm)
        assert st == 0
        for r, want in zip(res, single):
            got = r.as_dict()
            for k in want:
                if k != "status":
                    assert got[k] == want[k], k
    # frame 2 out of range: pel 1023 + 1 in a 10-bit picture
    w, h = 130, 66
    bad = make_pair(w, h, 10, 99, luma_offset=0.0).tst.copy()
    bad[5] = 1024
    triples[2] = (triples[2][0], engine.picture(w, h, 10, 420, data=bad), triples[2][2])
    res, st = engine.frames_masked(triples, PSNRS, default_params(), raise_on_error=False)
    assert st == 5
    assert res[2].as_dict()["in_range"][1] == 0
    for k in (0, 1, 3, 4):
        assert res[k].as_dict()["ivpsnr"] == engine.frame_masked(*triples[k], PSNRS, default_params()).ivpsnr


@pytest.mark.parametrize("B,wts", [(0, (4, 1, 1, 0)), (1, (4, 1, 1, 0)), (2, (4, 1, 1, 0)), (2, (1, 2, 3, 0)),
                                   (2, (4, 1, 1, 1)), (5, (1, 2, 3, 0)), (12, (4, 1, 1, 0))])
def test_search_parameters_match_oracle(engine, oracle, B, wts):
    w, h, bd = 150, 90, 10
    pair = make_pair(w, h, bd, 41, luma_offset=6.0)
    mask = make_mask(w, h, 8, 41, kind="soft")
    P0, P1 = oracle.unpack(pair.ref, w, h, bd), oracle.unpack(pair.tst, w, h, bd)
    M = oracle.unpack(mask, w, h, 8)[0]
    p0 = engine.picture(w, h, bd, 420, data=pair.ref)
    p1 = engine.picture(w, h, bd, 420, data=pair.tst)
    pm = engine.picture(w, h, 8, 420, data=mask)
    prm = default_params()
    prm.search_range = B
    for i in range(4):
        prm.cmp_weights_search[i] = wts[i]
    g = [3, -2, 1, 2]  # a non-zero fourth entry exercises the fourth lane of the error
    r = engine.calc_ivpsnr_masked(p0, p1, pm, g, -1, prm)
    nnm = int((M != 0).sum())
    rows_r2t = oracle.cps_rows_masked(P1, P0, M, [-v for v in g], B, wts)
    rows_t2r = oracle.cps_rows_masked(P0, P1, M, g, B, wts)
    want = []
    for rows in (rows_r2t, rows_t2r):
        err = [float(rows[c].sum()) for c in range(3)]  # exact uint64 sums, then double
        q = [oracle.psnr_from_masked_ssd(e if e > 0 else 1.0, nnm, bd, 8) for e in err]
        want.append((((q[0] * 4 + q[1] * 1) + q[2] * 1) + 0.0 * 0) * (1.0 / 6))
    assert r["r2t"] == want[0] and r["t2r"] == want[1]


@pytest.mark.parametrize("cf", [420, 444])
def test_full_mask_equals_unmasked_flow(engine, cf):
    w, h, bd = 320, 180, 10
    pair = make_pair(w, h, bd, 51, chroma_format=cf, luma_offset=5.0)
    mask = make_mask(w, h, 8, 51, chroma_format=cf, kind="full")
    p0 = engine.picture(w, h, bd, cf, data=pair.ref)
    p1 = engine.picture(w, h, bd, cf, data=pair.tst)
    pm = engine.picture(w, h, 8, cf, data=mask)
    for erp in (False, True):
        prm = default_params(erp=erp)
        prm.cps_int32_rows = 0
        m = engine.frame_masked(p0, p1, pm, PSNRS, prm)
        u = engine.frame(p0, p1, PSNRS, prm)
        # Without WS every value is a quotient of the same integers scaled by 255: identical. With
        # WS the weighted row sums round differently, so only to 1e-12.
        same = (lambda a, b: a == b) if not erp else (lambda a, b: a == pytest.approx(b, rel=1e-12))
        assert list(m.gcd) == list(u.gcd)
        assert list(m.psnr) == list(u.psnr)
        assert all(same(a, b) for a, b in zip(m.wspsnr, u.wspsnr))
        assert same(m.ivpsnr_r2t, u.ivpsnr_t2r) and same(m.ivpsnr_t2r, u.ivpsnr_r2t) and same(m.ivpsnr, u.ivpsnr)
        assert [list(x) for x in m.cps_sum] == [[v * 255 for v in u.cps_sum[1]], [v * 255 for v in u.cps_sum[0]]]


def test_errors_and_range_flags(engine):
    w, h = 64, 32
    pair = make_pair(w, h, 10, 61)
    p0 = engine.picture(w, h, 10, 420, data=pair.ref)
    p1 = engine.picture(w, h, 10, 420, data=pair.tst)
    bad = make_mask(w, h, 8, 61, kind="soft")
    m9 = engine.picture(w, h, 9, 444, data=np.full(3 * w * h, 600, np.uint16))  # 600 > 511
    r = engine.frame_masked(p0, p1, m9, PSNRS, raise_on_range=False)
    assert r.status == 5 and list(r.in_range) == [1, 1, 0]
    pm = engine.picture(w, h, 8, 420, data=bad)
    with pytest.raises(QmivError):
        engine.frame_masked(p0, p1, pm, PSNRS | 8)  # SSIM with a mask
    with pytest.raises(QmivError):
        engine.frame_masked(p0, p1, engine.picture(w + 2, h, 8, 420), PSNRS)
    prm = default_params()
    prm.search_range = 13
    with pytest.raises(QmivError):
        engine.frame_masked(p0, p1, pm, PSNRS, prm)


def test_masked_batch_equals_frame_set(engine):
    """qmiv_batch_create_masked: the captured masked frame set gives each frame exactly what
    qmiv_frames_masked gives, over frames of different sizes, formats, masks and bit depths (the
    16-bit search, the 24-bit one, 4:4:4, the general search of 16-bit pictures), replayed several
    times, tuned and with a fixed branch count, run and launch / collect; the unmasked calls refuse
    it; a frame with a pel above its bit depth fails the replay with its flags reported."""
    specs = [(1920, 1080, 8, 420, 8, "objects"), (256, 96, 8, 444, 16, "speckle"), (130, 66, 10, 420, 10, "soft"),
             (640, 360, 10, 420, 8, "speckle"), (96, 70, 16, 420, 16, "soft"), (384, 40, 10, 420, 8, "soft")]
    triples = []
    for k, (w, h, bd, cf, bdm, kind) in enumerate(specs):
        pair = make_pair(w, h, bd, 90 + k, chroma_format=cf, luma_offset=2.0)
        mask = make_mask(w, h, bdm, 90 + k, chroma_format=cf, kind=kind)
        triples.append((engine.picture(w, h, bd, cf, data=pair.ref), engine.picture(w, h, bd, cf, data=pair.tst),
                        engine.picture(w, h, bdm, cf, data=mask)))
    for erp, branches in ((False, 0), (True, 2)):
        print(f"masked batch erp={erp} branches={branches}", flush=True)
        prm = default_params(erp=erp)
        want, st = engine.frames_masked(triples, PSNRS, prm)
        assert st == 0
        bt = engine.batch_masked(triples, PSNRS, prm, branches=branches)
        info = bt.info()
        assert info["views"] == len(triples) and (branches == 0 or info["branches"] == branches)
        for rep in range(3):
            if rep == 2:
                bt.launch()
                got = bt.collect()
            else:
                got = bt.run()
            for i, (g, w_) in enumerate(zip(got, want)):
                assert g.as_dict() == w_.as_dict(), (erp, rep, i)
        out = (FrameResult * len(triples))()
        assert lib().qmiv_batch_run(bt.handle, out) != 0  # the unmasked run refuses a masked batch
        bt.launch()
        assert lib().qmiv_batch_collect(bt.handle, out) != 0  # ... and so does the unmasked collect
        assert [g.as_dict() for g in bt.collect()] == [w_.as_dict() for w_ in want]
        bt.close()
    # frame 2 out of range: pel 1023 + 1 in a 10-bit picture, reported by the replay
    w, h = 130, 66
    bad = make_pair(w, h, 10, 99, luma_offset=0.0).tst.copy()
    bad[5] = 1024
    triples[2] = (triples[2][0], engine.picture(w, h, 10, 420, data=bad), triples[2][2])
    bt = engine.batch_masked(triples, PSNRS, default_params(), branches=1)
    got = bt.run(raise_on_error=False)
    assert got[2].as_dict()["in_range"][1] == 0
    assert got[0].as_dict()["ivpsnr"] == engine.frame_masked(*triples[0], PSNRS, default_params()).ivpsnr
    with pytest.raises(QmivError):
        bt.run()
    bt.close()


def test_masked_batch_argument_errors(engine):
    """qmiv_batch_create_masked validates as qmiv_frames_masked does: structural-similarity metrics
    and unknown metric bits, mismatched pictures, a mask of another size and an empty set fail with
    QMIV_ERR_INVALID_ARG before anything is captured."""
    w, h = 128, 64
    pair = make_pair(w, h, 8, 5)
    mask = make_mask(w, h, 8, 5, kind="soft")
    p0 = engine.picture(w, h, 8, 420, data=pair.ref)
    p1 = engine.picture(w, h, 8, 420, data=pair.tst)
    pm = engine.picture(w, h, 8, 420, data=mask)
    other = engine.picture(w, h + 2, 8, 420)
    for metrics, triple in ((PSNRS | 8, (p0, p1, pm)), (PSNRS | (1 << 20), (p0, p1, pm)), (PSNRS, (p0, other, pm)),
                            (PSNRS, (p0, p1, other))):
        with pytest.raises(QmivError):
            engine.batch_masked([triple], metrics, default_params())
    with pytest.raises(QmivError):
        engine.batch_masked([], PSNRS, default_params())
    bt = engine.batch_masked([(p0, p1, pm)], PSNRS, default_params(), branches=1)
    assert bt.run()[0].as_dict() == engine.frame_masked(p0, p1, pm, PSNRS, default_params()).as_dict()
    bt.close()
