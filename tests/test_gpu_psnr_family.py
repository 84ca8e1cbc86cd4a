"""Parity of the engine's PSNR family (GCD, PSNR, WS-PSNR, IV-PSNR) with the oracle.

All comparisons are bit-exact: integer sums, GCD and the final doubles. Inputs are the seeded
synthetic pairs of ivpsnr_amd.synth plus hand-made edge cases.
"""
import math

import numpy as np
import pytest

from ivpsnr_amd import IVPSNR, PSNR, WSPSNR, default_params
from ivpsnr_amd.synth import make_pair
from tests.helpers import assert_frame_equal, oracle_params

pytestmark = pytest.mark.gpu

ALL_PSNR = PSNR | WSPSNR | IVPSNR

CASES = [
    # w, h, bd, chroma, seed, generator knobs
    (64, 48, 8, 420, 1, {}),
    (130, 66, 10, 420, 2, {}),                      # partial tiles on both axes
    (67, 45, 8, 444, 3, {}),                        # odd sizes (4:4:4)
    (256, 128, 10, 420, 4, {"luma_offset": 12.0}),  # GCD clipped to +-10
    (2, 2, 8, 420, 5, {}),                          # smaller than the search window
    (1, 1, 10, 444, 6, {}),
    (200, 8, 10, 420, 7, {"max_shift": 2}),         # fewer rows than a tile band
    (8, 200, 8, 420, 8, {}),                        # fewer columns than a tile
    (1920, 1080, 8, 420, 1, {}),                    # 1080p 8-bit frame of the plain workload
    # more than 4096 statistics rows (h + 2 ch): the summary kernel's row loop takes a second pass
    (2048, 1400, 10, 444, 9, {}),                   # 4200 rows
    (1024, 2304, 10, 420, 10, {}),                  # 4608 rows
]


def _upload_pair(engine, pair):
    p0 = engine.picture(pair.width, pair.height, pair.bit_depth, pair.chroma_format, data=pair.ref)
    p1 = engine.picture(pair.width, pair.height, pair.bit_depth, pair.chroma_format, data=pair.tst)
    return p0, p1


@pytest.mark.parametrize("erp", [False, True])
@pytest.mark.parametrize("w,h,bd,cf,seed,kw", CASES)
def test_frame_matches_oracle(engine, oracle, w, h, bd, cf, seed, kw, erp):
    pair = make_pair(w, h, bd, seed, chroma_format=cf, **kw)
    p0, p1 = _upload_pair(engine, pair)
    prm = default_params(erp=erp)
    got = engine.frame(p0, p1, ALL_PSNR, prm)
    want = oracle.frame(pair.ref, pair.tst, oracle_params(oracle, w, h, bd, cf, ALL_PSNR, prm))
    assert_frame_equal(got, want, ALL_PSNR, f"{w}x{h} bd{bd} cf{cf} seed{seed} erp{erp}")


@pytest.mark.parametrize("int32_rows", [True, False])
@pytest.mark.parametrize("w,h,bd,cf,seed", [(96, 70, 10, 420, 11), (67, 45, 8, 444, 12), (320, 180, 8, 420, 13)])
def test_cps_rows_match_oracle(engine, oracle, w, h, bd, cf, seed, int32_rows):
    pair = make_pair(w, h, bd, seed, chroma_format=cf)
    p0, p1 = _upload_pair(engine, pair)
    prm = default_params(cps_int32_rows=int(int32_rows))
    gcd = engine.calc_gcd(p0, p1, prm)
    P0 = oracle.unpack(pair.ref, w, h, bd, cf)
    P1 = oracle.unpack(pair.tst, w, h, bd, cf)
    assert list(gcd) == list(oracle.gcd(P0, P1, bd))
    got = engine.calc_ivpsnr(p0, p1, gcd, prm, want_rows=True)
    r2t = oracle.cps_rows(P0, P1, gcd, int32_rows=int32_rows)
    t2r = oracle.cps_rows(P1, P0, -gcd, int32_rows=int32_rows)
    np.testing.assert_array_equal(got["rows"][0], r2t)
    np.testing.assert_array_equal(got["rows"][1], t2r)
    area = w * h
    assert got["r2t"] == oracle.qual_from_rows(r2t, area, bd)
    assert got["t2r"] == oracle.qual_from_rows(t2r, area, bd)


GENERIC_PARAMS = [
    dict(search_range=1),
    dict(search_range=3),
    dict(cmp_weights_search=(200, 1, 1, 0)),   # beyond the fast kernel's 24-bit operand range
    dict(cmp_weights_search=(-1, 2, 2, 0)),    # negative weight: wrapping int32 errors
    dict(cmp_weights_search=(1, 1, 1, 0), cmp_weights_average=(1, 1, 1, 0)),  # fast path, other weights
    dict(cmp_weights_search=(6, 1, 1, 0), unnoticeable_coef=(0.02, 0.005, 0.0, 0.0)),
]


@pytest.mark.parametrize("kw", GENERIC_PARAMS, ids=lambda k: "-".join(f"{a}{b}" for a, b in k.items()))
def test_parameter_variants(engine, oracle, kw):
    w, h, bd, cf = 150, 90, 10, 420
    pair = make_pair(w, h, bd, 21, chroma_format=cf, luma_offset=5.0)
    p0, p1 = _upload_pair(engine, pair)
    prm = default_params(**kw)
    got = engine.frame(p0, p1, ALL_PSNR, prm)
    want = oracle.frame(pair.ref, pair.tst, oracle_params(oracle, w, h, bd, cf, ALL_PSNR, prm))
    assert_frame_equal(got, want, ALL_PSNR, str(kw))


@pytest.mark.parametrize("bd", [12, 14, 16])
def test_high_bit_depths(engine, oracle, bd):
    w, h = 80, 40
    pair = make_pair(w, h, bd, 31, chroma_format=420)
    p0, p1 = _upload_pair(engine, pair)
    got = engine.frame(p0, p1, ALL_PSNR, default_params())
    want = oracle.frame(pair.ref, pair.tst, oracle.params(w, h, bd, 420, metrics=ALL_PSNR))
    assert_frame_equal(got, want, ALL_PSNR, f"bd{bd}")


def test_ties_resolve_in_raster_order(engine, oracle):
    """Pictures drawn from {0, 1, 2} make most windows hold several equal errors."""
    rng = np.random.default_rng(5)
    w, h = 96, 64
    n = w * h + 2 * (w // 2) * (h // 2)
    a = rng.integers(0, 3, n).astype(np.uint8)
    b = rng.integers(0, 3, n).astype(np.uint8)
    p0 = engine.picture(w, h, 8, 420, data=a)
    p1 = engine.picture(w, h, 8, 420, data=b)
    prm = default_params()
    gcd = engine.calc_gcd(p0, p1, prm)
    got = engine.calc_ivpsnr(p0, p1, gcd, prm, want_rows=True)
    P0, P1 = oracle.unpack(a, w, h, 8), oracle.unpack(b, w, h, 8)
    np.testing.assert_array_equal(got["rows"][0], oracle.cps_rows(P0, P1, gcd))
    np.testing.assert_array_equal(got["rows"][1], oracle.cps_rows(P1, P0, -gcd))


def test_identical_pictures_closed_form(engine):
    """Identical pictures: every component exact; PSNR = IV-PSNR = 10 log10(area * max^2)."""
    w, h, bd = 1920, 1080, 8
    pair = make_pair(w, h, bd, 3)
    p0 = engine.picture(w, h, bd, data=pair.ref)
    p1 = engine.picture(w, h, bd, data=pair.ref)
    r = engine.frame(p0, p1, ALL_PSNR, default_params())
    closed = 10.0 * math.log10(w * h * 255.0 ** 2)
    assert list(r.exact) == [1, 1, 1]
    assert abs(r.psnr_pic - closed) < 1e-12 and abs(r.ivpsnr - closed) < 1e-12
    assert list(r.gcd) == [0, 0, 0, 0]
    assert [list(x) for x in r.cps_sum] == [[0, 0, 0], [0, 0, 0]]


def test_constant_offset_gcd(engine):
    """P1 = P0 + k on every pel: SD = area * k, so GCD = clip(k, +-rint(0.01 * max))."""
    w, h, bd = 128, 64, 10
    pair = make_pair(w, h, bd, 9)
    base = np.clip(pair.ref.astype(np.int32), 30, 990)
    for k, want in [(3, 3), (-7, -7), (25, 10), (-25, -10)]:
        p0 = engine.picture(w, h, bd, data=base.astype(np.uint16))
        p1 = engine.picture(w, h, bd, data=(base + k).astype(np.uint16))
        g = engine.calc_gcd(p0, p1)
        assert list(g) == [want, want, want, 0], (k, g)


def test_pel_range_error(engine):
    w, h, bd = 64, 32, 10
    pair = make_pair(w, h, bd, 4)
    bad = pair.tst.copy()
    bad[100] = 1024
    p0 = engine.picture(w, h, bd, data=pair.ref)
    p1 = engine.picture(w, h, bd, data=bad)
    r = engine.frame(p0, p1, ALL_PSNR, default_params(), raise_on_range=False)
    assert r.status == 5
    assert list(r.in_range) == [1, 0]


def test_erp_full_size_frame(engine, oracle):
    """One 4096x2048 10-bit ERP frame (the IV-PSNR + WS-PSNR workload) against the oracle."""
    w, h, bd = 4096, 2048, 10
    pair = make_pair(w, h, bd, 1)
    p0, p1 = _upload_pair(engine, pair)
    prm = default_params(erp=True)
    got = engine.frame(p0, p1, ALL_PSNR, prm)
    want = oracle.frame(pair.ref, pair.tst, oracle_params(oracle, w, h, bd, 420, ALL_PSNR, prm))
    assert_frame_equal(got, want, ALL_PSNR, "4096x2048 erp")
    assert list(got.gcd)[:3] == [10, 0, 0]  # +12 luma offset clipped to rint(0.01 * 1023)


def test_frames_batch_matches_single_calls(engine, oracle):
    """qmiv_frames on mixed sizes and bit depths equals per-frame qmiv_frame and the oracle."""
    specs = [(256, 128, 10, 420, 1), (130, 66, 8, 420, 2), (67, 45, 10, 444, 3), (512, 256, 10, 420, 4)]
    pairs, pics = [], []
    for (w, h, bd, cf, seed) in specs:
        pr = make_pair(w, h, bd, seed, chroma_format=cf)
        pairs.append(pr)
        pics.append(_upload_pair(engine, pr))
    prm = default_params(erp=True)
    batch = engine.frames(pics, ALL_PSNR, prm)
    for (w, h, bd, cf, seed), pr, (p0, p1), got in zip(specs, pairs, pics, batch):
        single = engine.frame(p0, p1, ALL_PSNR, prm)
        assert got.as_dict() == single.as_dict()
        want = oracle.frame(pr.ref, pr.tst, oracle_params(oracle, w, h, bd, cf, ALL_PSNR, prm))
        assert_frame_equal(got, want, ALL_PSNR, f"batch {w}x{h}")
