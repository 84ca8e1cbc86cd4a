"""The engine against the golden vectors the reference's own code wrote (tests/golden/), bit-exact."""
import json
import os

import numpy as np
import pytest

from ivpsnr_amd import IVPSNR, PSNR, WSPSNR, default_params
from ivpsnr_amd.synth import make_pair, sha256

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL_PSNR = PSNR | WSPSNR | IVPSNR


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


def _pics(engine, pair):
    return (engine.picture(pair.width, pair.height, pair.bit_depth, pair.chroma_format, data=pair.ref),
            engine.picture(pair.width, pair.height, pair.bit_depth, pair.chroma_format, data=pair.tst))


@pytest.mark.parametrize("mode", ["plain", "erp"])
@pytest.mark.parametrize("idx", range(len(load("small_cases"))))
def test_small_cases(engine, idx, mode):
    case = load("small_cases")[idx]
    w, h, bd, cf = case["width"], case["height"], case["bit_depth"], case["chroma_format"]
    pair = make_pair(w, h, bd, case["seed"], chroma_format=cf, **case["knobs"])
    assert sha256(pair.ref) == case["sha256_ref"]
    p0, p1 = _pics(engine, pair)
    prm = default_params(erp=mode == "erp")
    r = engine.frame(p0, p1, ALL_PSNR, prm)
    want = case[mode]
    assert list(r.gcd) == want["gcd"]
    assert list(r.psnr) == want["psnr"] and list(r.wspsnr) == want["wspsnr"]
    assert (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r) == (want["ivpsnr"], want["r2t"], want["t2r"])
    rows = engine.calc_ivpsnr(p0, p1, want["gcd"], prm, want_rows=True)["rows"]
    assert rows.tolist() == [want["cps_rows"][0], want["cps_rows"][1]]


def test_plain_1080p8_three_frames(engine, oracle):
    fx = load("plain_1080p8")
    ivs, psnrs = [], []
    for i, seed in enumerate(fx["seeds"]):
        pair = make_pair(1920, 1080, 8, seed=seed)
        assert sha256(pair.tst) == fx["sha256_tst"][i]
        p0, p1 = _pics(engine, pair)
        r = engine.frame(p0, p1, PSNR | IVPSNR, default_params())
        want = fx["frames"][i]
        assert list(r.gcd) == want["gcd"]
        assert list(r.psnr) == want["psnr"] and r.psnr_pic == want["psnr_pic"]
        assert (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r) == (want["ivpsnr"], want["r2t"], want["t2r"])
        assert [list(x) for x in r.cps_sum] == want["cps_rows_sum"]
        ivs.append(r.ivpsnr)
        psnrs.append(r.psnr_pic)
    assert oracle.frame_average(ivs) == fx["avg"]["IVPSNR"]["pic"]
    assert oracle.frame_average(psnrs) == fx["avg"]["PSNR"]["pic"]


def test_erp_4k10_frame(engine):
    fx = load("erp_4k10")
    pair = make_pair(4096, 2048, 10, seed=fx["seed"])
    assert sha256(pair.tst) == fx["sha256_tst"]
    p0, p1 = _pics(engine, pair)
    r = engine.frame(p0, p1, ALL_PSNR, default_params(erp=True))
    want = fx["frame"]
    assert list(r.gcd) == want["gcd"]
    assert list(r.psnr) == want["psnr"] and list(r.wspsnr) == want["wspsnr"]
    assert (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r) == (want["ivpsnr"], want["r2t"], want["t2r"])
    assert [list(x) for x in r.cps_sum] == want["cps_rows_sum"]
    rows = engine.calc_ivpsnr(p0, p1, want["gcd"], default_params(erp=True), want_rows=True)["rows"]
    import hashlib
    got_hash = [[hashlib.sha256(np.asarray(c, dtype="<u8").tobytes()).hexdigest() for c in d] for d in rows]
    assert got_hash == want["cps_rows_sha256"]
