"""Batched frame sets (qmiv_batch_*): views of uneven sizes, bit depths, chroma formats,
parameters and metrics captured into one HIP graph. Every replay must equal the per-view eager
results field for field (the same kernels run; only their scheduling differs), including replays
after new frames are uploaded into the bound pictures and with eager work between replays."""
import numpy as np
import pytest

from ivpsnr_amd import ALL_METRICS, IVPSNR, IVSSIM, PSNR, WSPSNR, default_params
from ivpsnr_amd.synth import make_pair

pytestmark = pytest.mark.gpu

VIEWS = [  # (w, h, bd, cf, erp, metrics)
    (256, 128, 10, 420, True, PSNR | WSPSNR | IVPSNR | IVSSIM),
    (192, 108, 10, 420, False, PSNR | WSPSNR | IVPSNR | IVSSIM),
    (208, 112, 10, 420, False, ALL_METRICS),
    (160, 96, 8, 420, False, IVPSNR),
    (96, 64, 12, 444, True, ALL_METRICS),
    (128, 64, 10, 420, True, PSNR | WSPSNR | IVPSNR),
]
FIELDS = ["in_range", "exact", "gcd", "psnr", "wspsnr", "cps_sum", "ssim", "msssim", "msssim_scales"]
SCALARS = ["psnr_pic", "wspsnr_pic", "ivpsnr", "ivpsnr_r2t", "ivpsnr_t2r", "ssim_pic", "msssim_pic", "ivssim",
           "ivssim_r2t", "ivssim_t2r"]


def as_tuple(r):
    out = []
    for f in FIELDS:
        v = getattr(r, f)
        out.append(tuple(tuple(x) if hasattr(x, "__len__") else x for x in v))
    out += [getattr(r, f) for f in SCALARS]
    return out


def same(a, b):
    # NaN (MS-SSIM of tiny pictures) compares equal to NaN here
    return np.array_equal(np.array(as_tuple(a)[-len(SCALARS):], float), np.array(as_tuple(b)[-len(SCALARS):], float),
                          equal_nan=True) and as_tuple(a)[:len(FIELDS)] == as_tuple(b)[:len(FIELDS)]


def make_views(engine, seed0):
    pics, data = [], []
    for i, (w, h, bd, cf, erp, m) in enumerate(VIEWS):
        pair = make_pair(w, h, bd, seed0 + i, chroma_format=cf)
        pics.append((engine.picture(w, h, bd, cf, data=pair.ref), engine.picture(w, h, bd, cf, data=pair.tst)))
        data.append(pair)
    return pics, data


def eager(engine, pics):
    return [engine.frame(p0, p1, v[5], default_params(erp=v[4])) for (p0, p1), v in zip(pics, VIEWS)]


@pytest.mark.parametrize("branches,sweep", [(0, -1), (1, 0), (3, 1), (8, 2)])
def test_batch_equals_eager_over_replays(engine, branches, sweep):
    pics, _ = make_views(engine, 100)
    prms = [default_params(erp=v[4]) for v in VIEWS]
    bt = engine.batch(pics, [v[5] for v in VIEWS], prms, branches=branches, cps_sweep=sweep)
    info = bt.info()
    assert info["views"] == len(VIEWS)
    if branches:
        assert info["branches"] == min(branches, len(VIEWS)) and info["candidates_tried"] == 0
        assert info["instances"] == 0  # a given shape is not timed
    else:
        if not info["from_cache"]:
            assert info["tune_rounds"] == 8 and info["candidates_tried"] >= 8 * 4
        # the kept instance was checked against the shape's fastest
        assert info["tuned_ms"] > 0 and 1 <= info["instances"] <= 6 and info["instance_ms"] > 0
    want = eager(engine, pics)
    for rep in range(3):
        got = bt.run()
        for i, (g, w) in enumerate(zip(got, want)):
            assert same(g, w), (rep, i)
        engine.frame(pics[0][0], pics[0][1], ALL_METRICS, default_params())  # eager work between replays
    # new frames into the same pictures: the next replay computes them
    for i, (w, h, bd, cf, erp, m) in enumerate(VIEWS):
        pair = make_pair(w, h, bd, 500 + i, chroma_format=cf)
        pics[i][0].upload(pair.ref)
        pics[i][1].upload(pair.tst)
    want2 = eager(engine, pics)
    assert as_tuple(want2[0]) != as_tuple(want[0])
    got2 = bt.run()
    for i, (g, w) in enumerate(zip(got2, want2)):
        assert same(g, w), i
    bt.close()


def test_batch_matches_oracle_bits(engine, oracle):
    """One view of the batch against the oracle directly (bit-exact IV-PSNR through the graph)."""
    pics, data = make_views(engine, 7)
    bt = engine.batch(pics, [v[5] for v in VIEWS], [default_params(erp=v[4]) for v in VIEWS], branches=2, cps_sweep=0)
    got = bt.run()
    for i in (0, 3, 5):
        w, h, bd, cf, erp, m = VIEWS[i]
        o = oracle.frame(data[i].ref, data[i].tst, oracle.params(w, h, bd, cf, metrics=m & 7, erp=erp))
        assert list(got[i].gcd) == list(o.gcd)
        assert got[i].ivpsnr == o.ivpsnr and list(got[i].psnr) == list(o.psnr) and list(got[i].wspsnr) == list(o.wspsnr)
    bt.close()


def test_batch_rejects_bad_arguments(engine):
    a = engine.picture(64, 32, 10)
    b = engine.picture(64, 34, 10)
    with pytest.raises(Exception):
        engine.batch([(a, b)], [IVPSNR])
    with pytest.raises(Exception):
        engine.batch([(a, a)], [IVPSNR], branches=17)


def test_tuned_choice_is_remembered(engine):
    """A second autotuned batch of the same signature takes the first one's choice without timing."""
    pics, _ = make_views(engine, 300)
    prms = [default_params(erp=v[4]) for v in VIEWS]
    a = engine.batch(pics, [v[5] for v in VIEWS], prms).info()
    b = engine.batch(pics, [v[5] for v in VIEWS], prms).info()
    assert b["from_cache"] == 1 and b["candidates_tried"] == 0
    assert (a["branches"], a["cps_sweep"], a["tuned_ms"]) == (b["branches"], b["cps_sweep"], b["tuned_ms"])
    assert 1 <= b["instances"] <= 6
    other = [default_params(erp=not v[4]) for v in VIEWS]  # another signature is tuned anew
    c = engine.batch(pics, [v[5] for v in VIEWS], other).info()
    assert c["from_cache"] == 0 and c["candidates_tried"] > 0


def test_launch_collect_pipelines_frame_sets(engine):
    """qmiv_batch_launch / qmiv_batch_collect: several batches launched back to back, then
    collected, give the eager results; a second launch before the collect and a collect with
    nothing launched are refused."""
    sets = [make_views(engine, 700 + 10 * k)[0] for k in range(3)]
    prms = [default_params(erp=v[4]) for v in VIEWS]
    bts = [engine.batch(pics, [v[5] for v in VIEWS], prms, branches=4, cps_sweep=0) for pics in sets]
    want = [eager(engine, pics) for pics in sets]
    for rep in range(2):
        for bt in bts:
            bt.launch()
        with pytest.raises(Exception):
            bts[0].launch()  # one replay in flight per batch
        for k in (1, 0, 2):  # any collect order
            got = bts[k].collect()
            for i, (g, w) in enumerate(zip(got, want[k])):
                assert same(g, w), (rep, k, i)
    with pytest.raises(Exception):
        bts[0].collect()  # nothing in flight
    got = bts[2].run()  # run still works after launch / collect
    assert all(same(g, w) for g, w in zip(got, want[2]))
    for bt in bts:
        bt.close()
