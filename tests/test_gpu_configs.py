"""Frame sets at the measured workloads' real shapes against golden vectors of the reference's own
code (tests/golden/make_golden.py --multiview / --ssimsets, oracle/_ref/qmiv_ref).

* The 24-view batch: 8 x 4096x2048 ERP, 8 x 1920x1080 and 8 x 2048x1088 views, 10-bit 4:2:0,
  IV-PSNR + WS-PSNR + PSNR + IV-SSIM, as one captured graph (qmiv_batch_create / qmiv_batch_run,
  autotuned), its first frame set and then a second one uploaded into the same pictures
  (the reference's per-view frame loop, apps/QMIV/src/xAppQMIV.cpp:661-751). PSNR family, GCD and
  the CPS sums bit-exact; IV-SSIM within 1e-9 relative (libs/pmbb-ivqm/src/xSSIM.cpp:192-212).
* SSIM-family frame sets through qmiv_frames (two streams, one set of shift-compensated pictures
  per stream, a pyramid per frame): 16 frames 1920x1080 10-bit, and five frames of mixed sizes and
  bit depths with every metric (libs/pmbb-ivqm/src/xSSIM.cpp:54-87). Any reuse of a stream's
  buffers before the previous frame on it has finished shows as a wrong value.
"""
import concurrent.futures as cf
import json
import os

import numpy as np
import pytest

from ivpsnr_amd import ALL_METRICS, IVPSNR, IVSSIM, MSSSIM, PSNR, SSIM, WSPSNR, default_params
from ivpsnr_amd.synth import make_pair, sha256

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REL = 1e-9


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


def close(a, b, rel=REL):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return bool(np.all(np.abs(a - b) <= rel * np.maximum(np.abs(b), 1e-300)))


def check_psnr_family(r, want, ctx):
    assert list(r.in_range) == want["in_range"] and list(r.exact) == want["exact"], ctx
    assert list(r.gcd) == want["gcd"], ctx
    assert list(r.psnr) == want["psnr"] and r.psnr_pic == want["psnr_pic"], ctx
    assert list(r.wspsnr) == want["wspsnr"] and r.wspsnr_pic == want["wspsnr_pic"], ctx
    assert (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r) == (want["ivpsnr"], want["r2t"], want["t2r"]), ctx
    assert [list(x) for x in r.cps_sum] == want["cps_rows_sum"], ctx


def check_ssim_family(r, want, metrics, ctx):
    if metrics & SSIM:
        assert close(list(r.ssim) + [r.ssim_pic], want["ssim"] + [want["ssim_pic"]]), (ctx, list(r.ssim), want["ssim"])
    if metrics & MSSSIM:
        assert close(list(r.msssim) + [r.msssim_pic], want["msssim"] + [want["msssim_pic"]]), ctx
        assert close(np.array(r.msssim_scales), want["msssim_scales"]), ctx
    if metrics & IVSSIM:
        got = [r.ivssim, r.ivssim_r2t, r.ivssim_t2r]
        assert close(got, [want["ivssim"], want["ivssim_r2t"], want["ivssim_t2r"]]), (ctx, got)


def test_multiview24_batch_two_frame_sets(engine):
    fx = load("multiview24")
    views = fx["views"]
    assert len(views) == 24
    metrics = PSNR | WSPSNR | IVPSNR | IVSSIM
    with cf.ThreadPoolExecutor(8) as ex:
        base = list(ex.map(lambda v: make_pair(v["width"], v["height"], 10, seed=v["seed"]), views))
    pics = []
    for v, pair in zip(views, base):
        assert [sha256(pair.ref), sha256(pair.tst)] == v["sha256"][0]
        pics.append((engine.picture(v["width"], v["height"], 10, 420, data=pair.ref),
                     engine.picture(v["width"], v["height"], 10, 420, data=pair.tst)))
    prms = [default_params(erp=v["erp"]) for v in views]
    bt = engine.batch(pics, [metrics] * 24, prms)
    info = bt.info()
    print("autotuned:", info)
    for f in range(len(views[0]["frames"])):
        if f:
            # the next frame set: new content uploaded into the same pictures, then one replay
            for (p0, p1), v, pair in zip(pics, views, base):
                r, t = np.roll(pair.ref, 64 * f), np.roll(pair.tst, 64 * f)
                assert [sha256(r), sha256(t)] == v["sha256"][f]
                p0.upload(r)
                p1.upload(t)
        out = bt.run()
        for i, (r, v) in enumerate(zip(out, views)):
            ctx = f"view {i} ({v['width']}x{v['height']}{' ERP' if v['erp'] else ''}) frame {f}"
            want = v["frames"][f]
            check_psnr_family(r, want, ctx)
            check_ssim_family(r, want, IVSSIM, ctx)
    # a second replay of the same frame set gives the same bits
    again = bt.run()
    for r, s in zip(out, again):
        assert r.as_dict() == s.as_dict()
    bt.close()
    for p0, p1 in pics:
        p0.close()
        p1.close()


@pytest.mark.parametrize("name", ["1080p10", "mixed"])
def test_ssim_family_frame_sets(engine, name):
    fx = load("ssim_sets")[name]
    metrics = {"SSIM,MSSSIM,IVSSIM": SSIM | MSSSIM | IVSSIM, "All": ALL_METRICS}[fx["metrics"]]
    frames = fx["frames"]
    with cf.ThreadPoolExecutor(8) as ex:
        pairs = list(ex.map(lambda e: make_pair(e["width"], e["height"], e["bit_depth"], seed=e["seed"]), frames))
    pics = []
    for e, pair in zip(frames, pairs):
        assert sha256(pair.ref) == e["sha256_ref"] and sha256(pair.tst) == e["sha256_tst"]
        pics.append((engine.picture(e["width"], e["height"], e["bit_depth"], 420, data=pair.ref),
                     engine.picture(e["width"], e["height"], e["bit_depth"], 420, data=pair.tst)))
    # twice: the second set reuses the context's buffers laid out by the first
    for rep in range(2):
        out = engine.frames(pics, metrics, default_params())
        for i, (r, e) in enumerate(zip(out, frames)):
            ctx = f"{name} frame {i} ({e['width']}x{e['height']} {e['bit_depth']}b) pass {rep}"
            want = e["frame"]
            assert list(r.gcd) == want["gcd"], ctx
            if metrics & PSNR:
                check_psnr_family(r, want, ctx)
            check_ssim_family(r, want, metrics, ctx)
    for p0, p1 in pics:
        p0.close()
        p1.close()
