"""Pins the C oracle (oracle/qmiv_oracle.c) before anything is checked against it.

* golden vectors written by the reference's own code (tests/golden/make_golden.py, the reference
  sources built without FMA) on seeded synthetic inputs: bit-exact;
* the published example values (example/README.md:46-47), when the reference tree is present;
* closed forms and invariants (identical pictures, constant offsets, SSIM interior fraction,
  WS weights, KBNS exactness, the SSIM window's regeneration);
* an independent numpy brute force of the CPS argmin on tiny inputs (raster tie order).
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from ivpsnr_amd.synth import make_pair, sha256

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EXAMPLES = "/root/reference/example"


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


def h64(rows):
    return hashlib.sha256(np.asarray(rows, dtype="<u8").tobytes()).hexdigest()


SMALL = load("small_cases")


@pytest.mark.parametrize("case", SMALL, ids=lambda c: f"{c['width']}x{c['height']}b{c['bit_depth']}c{c['chroma_format']}")
@pytest.mark.parametrize("mode", ["plain", "erp"])
def test_small_cases_bit_exact(oracle, case, mode):
    w, h, bd, cf = case["width"], case["height"], case["bit_depth"], case["chroma_format"]
    pair = make_pair(w, h, bd, case["seed"], chroma_format=cf, **case["knobs"])
    assert sha256(pair.ref) == case["sha256_ref"] and sha256(pair.tst) == case["sha256_tst"], "generator drifted"
    want = case[mode]
    prm = oracle.params(w, h, bd, cf, metrics=oracle.ALL, erp=mode == "erp")
    r, scp0, scp1 = oracle.frame(pair.ref, pair.tst, prm, want_scp=True)
    d = r.as_dict()
    assert d["gcd"] == want["gcd"]
    assert d["exact"] == want["exact"] and d["in_range"] == want["in_range"]
    for k in ("psnr", "wspsnr", "ssim", "msssim"):
        assert d[k] == want[k], k
    for k in ("ivpsnr", "r2t", "t2r", "ivssim", "ivssim_r2t", "ivssim_t2r", "psnr_pic", "wspsnr_pic", "ssim_pic",
              "msssim_pic"):
        assert d[k] == want[k], k
    assert np.array(d["msssim_scales"]).tolist() == want["msssim_scales"]
    # per-row CPS values of both directions
    P0 = oracle.unpack(pair.ref, w, h, bd, cf)
    P1 = oracle.unpack(pair.tst, w, h, bd, cf)
    g = np.array(want["gcd"], np.int32)
    assert oracle.cps_rows(P0, P1, g).tolist() == want["cps_rows"][0]
    assert oracle.cps_rows(P1, P0, -g).tolist() == want["cps_rows"][1]
    # shift-compensated pictures
    for c in range(3):
        assert hashlib.sha256(scp0[c].astype("<u2").tobytes()).hexdigest() == want["scp0_sha256"][c]
        assert hashlib.sha256(scp1[c].astype("<u2").tobytes()).hexdigest() == want["scp1_sha256"][c]
    if mode == "erp":
        ws, corr = oracle.ws_init(h)
        assert ws.tolist() == want["ws_weights"] and corr == want["ws_correction"]
        for c in range(3):
            assert oracle.ssd_rows(P1[c], P0[c]).tolist() == want["ws_rows"][c]


def test_plain_1080p8_three_frames(oracle):
    fx = load("plain_1080p8")
    ivs, psnrs = [], []
    for i, seed in enumerate(fx["seeds"]):
        pair = make_pair(1920, 1080, 8, seed=seed)
        assert sha256(pair.ref) == fx["sha256_ref"][i] and sha256(pair.tst) == fx["sha256_tst"][i]
        r = oracle.frame(pair.ref, pair.tst, oracle.params(1920, 1080, 8, metrics=oracle.PSNR | oracle.IVPSNR))
        want = fx["frames"][i]
        assert list(r.gcd) == want["gcd"]
        assert list(r.psnr) == want["psnr"] and r.psnr_pic == want["psnr_pic"]
        assert (r.ivpsnr, r.r2t, r.t2r) == (want["ivpsnr"], want["r2t"], want["t2r"])
        assert [list(x) for x in r.cps_sum] == want["cps_rows_sum"]
        ivs.append(r.ivpsnr)
        psnrs.append(r.psnr_pic)
    assert oracle.frame_average(ivs) == fx["avg"]["IVPSNR"]["pic"]
    assert oracle.frame_average(psnrs) == fx["avg"]["PSNR"]["pic"]


def test_erp_4k10_frame(oracle):
    fx = load("erp_4k10")
    pair = make_pair(4096, 2048, 10, seed=fx["seed"])
    assert sha256(pair.ref) == fx["sha256_ref"] and sha256(pair.tst) == fx["sha256_tst"]
    prm = oracle.params(4096, 2048, 10, metrics=oracle.PSNR | oracle.WSPSNR | oracle.IVPSNR, erp=True)
    r = oracle.frame(pair.ref, pair.tst, prm)
    want = fx["frame"]
    assert list(r.gcd) == want["gcd"] == [10, 0, 0, 0]
    assert list(r.psnr) == want["psnr"] and list(r.wspsnr) == want["wspsnr"]
    assert (r.ivpsnr, r.r2t, r.t2r) == (want["ivpsnr"], want["r2t"], want["t2r"])
    assert [list(x) for x in r.cps_sum] == want["cps_rows_sum"]


def test_ssim_family_1080p10(oracle):
    fx = load("ssim_1080p10")
    pair = make_pair(1920, 1080, 10, seed=fx["seed"])
    assert sha256(pair.ref) == fx["sha256_ref"]
    prm = oracle.params(1920, 1080, 10, metrics=oracle.SSIM | oracle.MSSSIM | oracle.IVSSIM)
    r = oracle.frame(pair.ref, pair.tst, prm)
    want = fx["frame"]
    assert list(r.ssim) == want["ssim"] and list(r.msssim) == want["msssim"]
    assert np.array(r.msssim_scales).tolist() == want["msssim_scales"]
    assert (r.ivssim, r.ivssim_r2t, r.ivssim_t2r) == (want["ivssim"], want["ivssim_r2t"], want["ivssim_t2r"])


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="reference example frames not present")
def test_published_example_values(oracle):
    fx = load("examples")
    orig = np.fromfile(os.path.join(EXAMPLES, "Fencing_v8_original_1920x1080_yuv420p.yuv"), np.uint8)
    for name in ("origJpgCompressed", "synthesized", "original"):
        other = np.fromfile(os.path.join(EXAMPLES, f"Fencing_v8_{name}_1920x1080_yuv420p.yuv"), np.uint8)
        r = oracle.frame(orig, other, oracle.params(1920, 1080, 8, metrics=oracle.ALL)).as_dict()
        want = fx[name]["frame"]
        for k in ("gcd", "psnr", "ivpsnr", "r2t", "t2r", "ssim", "msssim", "ivssim"):
            assert r[k] == want[k], (name, k)
        if name in fx["published"]:
            pub = fx["published"][name]
            assert round(r["psnr"][0], 4) == pub["psnr_y"] and round(r["ivpsnr"], 4) == pub["ivpsnr"]


def test_ssim_window_regenerates_reference_table(oracle):
    bits = np.array(load("ssim_filter_bits"), dtype=np.uint32)
    ours = oracle.ssim_filter().ravel().view(np.uint32)
    np.testing.assert_array_equal(ours, bits)
    assert abs(float(oracle.ssim_filter()[:10, :10].astype(np.float64).sum()) - 0.99794429) < 1e-8


# ------------------------------------------------------------------------------------------------
# Closed forms and invariants
# ------------------------------------------------------------------------------------------------

def test_identical_pictures(oracle):
    w, h = 1920, 1080
    pair = make_pair(w, h, 8, 2)
    r = oracle.frame(pair.ref, pair.ref, oracle.params(w, h, 8, metrics=oracle.PSNR | oracle.IVPSNR | oracle.SSIM))
    closed = 10 * math.log10(w * h * 255.0 ** 2)
    assert abs(r.psnr[0] - 111.2980534506) < 1e-9 and abs(r.psnr[0] - closed) < 1e-12
    assert abs(r.ivpsnr - closed) < 1e-12
    assert abs(r.ssim[0] - (w - 10) * (h - 10) / (w * h)) < 1e-15  # 0.985580632716


def test_msssim_identical_closed_form(oracle):
    w, h = 1920, 1080
    pair = make_pair(w, h, 8, 2)
    y = oracle.unpack(pair.ref, w, h, 8)[0]
    v, _ = oracle.msssim_plane(y, y, 8)
    wts = [0.0448, 0.2856, 0.3001, 0.2363, 0.1333]
    closed = 1.0
    for i in range(5):
        wi, hi = w >> i, h >> i
        closed *= ((wi - 10) * (hi - 10) / (wi * hi)) ** wts[i]
    assert abs(v - closed) < 1e-12 and abs(v - 0.915749972949) < 1e-11


@pytest.mark.parametrize("k", [-30, -5, 0, 2, 9, 40])
def test_constant_offset_sd_and_gcd(oracle, k):
    w, h, bd = 64, 32, 10
    rng = np.random.default_rng(k + 100)
    P0 = rng.integers(100, 900, size=(3, h, w)).astype(np.uint16)
    P1 = (P0.astype(np.int32) + k).astype(np.uint16)
    g = oracle.gcd(P0, P1, bd)
    assert list(g) == [max(-10, min(k, 10))] * 3 + [0]
    for c in range(3):
        assert int(oracle.ssd_rows(P0[c], P1[c]).sum()) == w * h * k * k


def test_gcd_clip_bounds_per_bit_depth(oracle):
    # rint_fp32(0.01 * max): 3 / 10 / 41 / 164 for 8 / 10 / 12 / 14 bits
    for bd, bound in [(8, 3), (10, 10), (12, 41), (14, 164)]:
        P0 = np.zeros((3, 4, 4), np.uint16)
        P1 = np.full((3, 4, 4), (1 << bd) - 1, np.uint16)
        assert list(oracle.gcd(P0, P1, bd))[:3] == [bound] * 3


def test_gcd_sum_wraps_in_int32(oracle):
    """The reference sums (P1 - P0) in int32; a sum beyond 2^31 wraps before the division."""
    w, h, bd = 4096, 4096, 16
    P0 = np.zeros((3, h, w), np.uint16)
    P1 = np.zeros((3, h, w), np.uint16)
    P1[0, :, :] = 200  # 16.7M pels x 200 = 3.36e9 > 2^31
    g = oracle.gcd(P0, P1, bd, unnoticeable=(1.0, 1.0, 1.0, 0.0))
    wrapped = np.int32(np.uint32((w * h * 200) % (1 << 32)))
    assert g[0] == int(np.rint(float(wrapped) / (w * h)))


def test_ws_weights_properties(oracle):
    for h in (256, 2048):
        w, corr = oracle.ws_init(h)
        assert np.all(w > 0) and np.all(w <= 1)
        np.testing.assert_allclose(w, w[::-1], rtol=0, atol=1e-15)
        assert abs(w.sum() / h - 2 / math.pi) < 2.0 / h
        assert abs(corr - h / oracle.kbns(w)) == 0


@pytest.mark.parametrize("v", [0.0, 0.01, 0.1, 0.9, 0.99, 1.0])
def test_kbns_exact_for_repeated_values(oracle, v):
    n = 1 << 16
    assert oracle.kbns(np.full(n, v)) == n * v


def test_psnr_from_ssd_textbook(oracle):
    for ssd, area, bd in [(1.0, 1920 * 1080, 8), (12345.0, 4096 * 2048, 10), (7.5, 100, 12)]:
        want = 10 * math.log10(area * ((1 << bd) - 1) ** 2 / ssd)
        assert abs(oracle.psnr_from_ssd(ssd, area, bd) - want) < 1e-12


def test_downsample_is_rounded_mean(oracle):
    rng = np.random.default_rng(1)
    src = rng.integers(0, 1024, size=(34, 50)).astype(np.uint16)
    got = oracle.downsample(src).astype(np.int64)
    s = src.astype(np.int64)
    want = (s[0::2, 0::2] + s[0::2, 1::2] + s[1::2, 0::2] + s[1::2, 1::2] + 2) >> 2
    np.testing.assert_array_equal(got, want)


def test_unpack_is_nearest_neighbour(oracle):
    pair = make_pair(32, 16, 10, 3)
    P = oracle.unpack(pair.ref, 32, 16, 10)
    y, cb, cr = pair.planes("ref")
    np.testing.assert_array_equal(P[0], y)
    np.testing.assert_array_equal(P[1], np.repeat(np.repeat(cb, 2, axis=0), 2, axis=1))
    np.testing.assert_array_equal(P[2], np.repeat(np.repeat(cr, 2, axis=0), 2, axis=1))


def _brute_cps(X, Y, g, B=2, wts=(4, 1, 1)):
    """Independent numpy restatement: all window errors, first argmin in raster order."""
    _, h, w = X.shape
    Xg = X.astype(np.int64) + np.array(g[:3], np.int64)[:, None, None]
    errs, cands = [], []
    for dy in range(-B, B + 1):
        for dx in range(-B, B + 1):
            ys = np.clip(np.arange(h) + dy, 0, h - 1)
            xs = np.clip(np.arange(w) + dx, 0, w - 1)
            Yq = Y[:, ys][:, :, xs].astype(np.int64)
            d2 = (Xg - Yq) ** 2
            errs.append(sum(wts[c] * d2[c] for c in range(3)))
            cands.append(d2)
    errs = np.stack(errs)            # (25, h, w)
    best = np.argmin(errs, axis=0)   # first occurrence = raster order
    d2 = np.stack(cands)             # (25, 3, h, w)
    sel = np.take_along_axis(d2, best[None, None], axis=0)[0]
    return sel.sum(axis=2).astype(np.uint64), best


@pytest.mark.parametrize("seed", range(6))
def test_cps_rows_against_brute_force(oracle, seed):
    rng = np.random.default_rng(seed)
    h, w = int(rng.integers(1, 20)), int(rng.integers(1, 24))
    hi = [3, 5, 1024][seed % 3]  # tiny value ranges force many equal errors
    X = rng.integers(0, hi, size=(3, h, w)).astype(np.uint16)
    Y = rng.integers(0, hi, size=(3, h, w)).astype(np.uint16)
    g = np.array([int(rng.integers(-2, 3)), 0, int(rng.integers(-1, 2)), 0], np.int32)
    rows, best = oracle.cps_rows(X, Y, g, want_best=True)
    want_rows, want_best = _brute_cps(X, Y, g)
    np.testing.assert_array_equal(best, want_best)
    np.testing.assert_array_equal(rows, want_rows)


def test_int32_row_wrap_semantics(oracle):
    """A row whose squared errors exceed 2^31 wraps in the SSE path and is sign-extended."""
    w, h = 2100, 1
    X = np.zeros((3, h, w), np.uint16)
    Y = np.full((3, h, w), 1023, np.uint16)
    rows32 = oracle.cps_rows(X, Y, [0, 0, 0, 0], int32_rows=True)
    rows64 = oracle.cps_rows(X, Y, [0, 0, 0, 0], int32_rows=False)
    exact = w * 1023 * 1023
    assert int(rows64[0, 0]) == exact
    assert int(rows32[0, 0]) == int(np.uint64(np.int64(np.int32(np.uint32(exact % (1 << 32))))))


def test_window_table_helper_matches_reference_bits():
    """The test helper's SSIM window equals the reference's fp32 table bit for bit."""
    import json, os
    import numpy as np
    from .helpers import ssim_window_table
    bits = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ssim_filter_bits.json")))
    table = np.array(bits, np.uint32).view(np.float32).reshape(11, 11)
    np.testing.assert_array_equal(ssim_window_table(), table[:10, :10].astype(np.float64))


COLOUR = load("colorspace")


@pytest.mark.parametrize("space", range(4))
@pytest.mark.parametrize("to_rgb", [0, 1])
def test_colour_tables_equal_reference(oracle, space, to_rgb):
    """The Q16 matrices equal those printed by the reference build (xColorSpaceCoeff.cpp)."""
    want = np.array(COLOUR["tables"][space][to_rgb], np.int32).reshape(3, 3)
    np.testing.assert_array_equal(oracle.colorspace_coeffs(space, bool(to_rgb)), want)


def test_colour_tables_rows_follow_from_luma_weights(oracle):
    """Invariants of any RGB<->YCbCr matrix: the luma row sums to 1.0 (within rounding), the chroma
    rows sum to zero, the inverse maps grey to grey and its product with the forward matrix is the
    identity to Q16 rounding."""
    for space in range(4):
        fwd = oracle.colorspace_coeffs(space, False).astype(np.float64) / 65536
        inv = oracle.colorspace_coeffs(space, True).astype(np.float64) / 65536
        assert abs(fwd[0].sum() - 1.0) < 2e-5
        assert abs(fwd[1].sum()) < 2e-5 and abs(fwd[2].sum()) < 2e-5
        assert fwd[1, 2] == 0.5 and fwd[2, 0] == 0.5
        np.testing.assert_allclose(inv[:, 0], 1.0)
        np.testing.assert_allclose(inv @ fwd, np.eye(3), atol=1e-4)


@pytest.mark.parametrize("case", COLOUR["cases"], ids=lambda c: f"b{c['bit_depth']}s{c['space']}r{c['to_rgb']}")
def test_colour_conversion_matches_reference_outputs(oracle, case):
    """Seeded planes converted by the reference's ConvertRGB2YCbCr_I32 / ConvertYCbCr2RGB_I32."""
    src = np.random.default_rng(case["seed"]).integers(0, 1 << case["bit_depth"],
                                                        size=(3, case["height"], case["width"]), dtype=np.uint16)
    out = oracle.colorspace(src, case["bit_depth"], case["space"], bool(case["to_rgb"]))
    assert out.reshape(-1)[:8].tolist() == case["head"]
    assert hashlib.sha256(out.astype("<u2").tobytes()).hexdigest() == case["sha256_out"]


@pytest.mark.parametrize("space", range(4))
@pytest.mark.parametrize("w", [63, 64, 65, 127, 128, 129])
def test_colour_round_trip_14bit_gradients(oracle, space, w):
    """RGB -> YCbCr -> RGB of horizontal gradients (offsets 0, 100, 200) at 14 bits returns every
    sample within 1 (test-core-xColorspace.cpp:107-125)."""
    h = 65
    x = np.arange(w, dtype=np.int64)
    src = np.stack([np.broadcast_to((x + off) & 0x3FFF, (h, w)) for off in (0, 100, 200)]).astype(np.uint16)
    back = oracle.colorspace(oracle.colorspace(src, 14, space, False), 14, space, True)
    assert np.abs(back.astype(np.int32) - src.astype(np.int32)).max() <= 1


def test_colour_grey_maps_to_mid_chroma(oracle):
    """R = G = B = v converts to Y = v and Cb = Cr = 2^(bd-1) (chroma rows sum to zero). Up to 14
    bits: at 16 bits the reference's int32 products wrap (65536 * v > 2^31), which the 16-bit golden
    cases above pin."""
    for bd in (8, 10, 14):
        v = np.arange(0, 1 << bd, max(1, (1 << bd) // 97), dtype=np.uint16)
        out = oracle.colorspace(np.stack([v, v, v])[:, None, :], bd, 1, False)
        assert np.abs(out[0, 0].astype(np.int64) - v).max() <= 1
        assert np.abs(out[1:].astype(np.int64) - (1 << (bd - 1))).max() <= 1


# ---------------------------------------------------------------------------------------------
# Masked mode
# ---------------------------------------------------------------------------------------------
from ivpsnr_amd.synth import make_mask  # noqa: E402

MASKED = load("masked")
MASKED_METRICS = 1 | 2 | 4  # PSNR, WS-PSNR, IV-PSNR


def masked_inputs(case):
    w, h, bd, cf = case["width"], case["height"], case["bit_depth"], case["chroma_format"]
    pair = make_pair(w, h, bd, case["seed"], chroma_format=cf, **case["knobs"])
    mask = make_mask(w, h, case["mask_bit_depth"], case["seed"], chroma_format=case["mask_chroma_format"],
                     kind=case["mask_kind"])
    assert sha256(pair.ref) == case["sha256_ref"] and sha256(pair.tst) == case["sha256_tst"], "generator drifted"
    assert sha256(mask) == case["sha256_mask"], "mask generator drifted"
    return pair, mask


@pytest.mark.parametrize("case", MASKED, ids=lambda c: f"{c['width']}x{c['height']}b{c['bit_depth']}{c['mask_kind']}")
@pytest.mark.parametrize("mode", ["plain", "erp"])
def test_masked_cases_bit_exact(oracle, case, mode):
    """The masked flow against the reference's calcPicPSNRM / calcPicWSPSNRM / calcPicIVPSNRM and
    CalcGlobalColorDiffM on seeded pictures and masks (binary, soft, empty, full; 8 to 16 bits)."""
    w, h, bd, cf = case["width"], case["height"], case["bit_depth"], case["chroma_format"]
    pair, mask = masked_inputs(case)
    want = case[mode]
    prm = oracle.params(w, h, bd, cf, metrics=MASKED_METRICS, erp=mode == "erp")
    d = oracle.frame_masked(pair.ref, pair.tst, mask, case["mask_bit_depth"], case["mask_chroma_format"], prm).as_dict()
    assert d["num_non_masked"] == want["num_non_masked"]
    assert d["gcd"] == want["gcd"]
    assert d["exact"] == want["exact"] and d["in_range"][:2] == want["in_range"]
    assert d["in_range"][2] == want["in_range_mask"]
    for k in ("psnr", "wspsnr"):
        assert d[k] == want[k], k
    for k in ("ivpsnr", "r2t", "t2r", "psnr_pic", "wspsnr_pic"):
        assert d[k] == want[k], k
    assert [list(x) for x in d["cps_sum"]] == want["cps_rows_sum"]
    P0 = oracle.unpack(pair.ref, w, h, bd, cf)
    P1 = oracle.unpack(pair.tst, w, h, bd, cf)
    M = oracle.unpack(mask, w, h, case["mask_bit_depth"], case["mask_chroma_format"])[0]
    g = np.array(want["gcd"], np.int32)
    rows = [oracle.cps_rows_masked(P1, P0, M, -g), oracle.cps_rows_masked(P0, P1, M, g)]
    assert [[h64(r) for r in d_] for d_ in rows] == want["cps_rows_sha256"]


def test_published_masked_example(oracle):
    """example/README.md:64-69: Fencing synthesized vs original, objects mask: PSNR-Y (M) 39.5414 dB,
    IV-PSNR (M) 50.7716 dB."""
    base = os.path.join(EXAMPLES, "Fencing_v8_")
    if not os.path.exists(base + "objectsMask_1920x1080_yuv420p.yuv"):
        pytest.skip("reference example frames not present")
    n = 1920 * 1080 * 3 // 2
    f0 = np.fromfile(base + "original_1920x1080_yuv420p.yuv", np.uint8)[:n]
    f1 = np.fromfile(base + "synthesized_1920x1080_yuv420p.yuv", np.uint8)[:n]
    m = np.fromfile(base + "objectsMask_1920x1080_yuv420p.yuv", np.uint8)[:n]
    r = oracle.frame_masked(f0, f1, m, 8, 420, oracle.params(1920, 1080, 8, 420, metrics=MASKED_METRICS))
    assert f"{r.psnr[0]:.4f}" == "39.5414"
    assert f"{r.ivpsnr:.4f}" == "50.7716"


@pytest.mark.parametrize("bd,bdm,cf", [(8, 8, 420), (10, 8, 444), (12, 12, 420)])
@pytest.mark.parametrize("erp", [False, True])
def test_full_mask_equals_unmasked(oracle, bd, bdm, cf, erp):
    """A mask of 2^bdm - 1 everywhere scales every SSD, SD and row sum by the same integer as the
    denominators: without WS every value is a quotient of the same exact integers, identical to the
    unmasked flow's (with exact uint64 CPS rows); with WS the weighted sums round differently, so
    to 1e-12. The masked processor's R2T is the unmasked T2R and vice versa."""
    w, h = 70, 42
    pair = make_pair(w, h, bd, 21, chroma_format=cf, luma_offset=5.0)
    mask = make_mask(w, h, bdm, 21, chroma_format=420, kind="full")
    prm = oracle.params(w, h, bd, cf, metrics=MASKED_METRICS, erp=erp, int32_rows=False)
    mr = oracle.frame_masked(pair.ref, pair.tst, mask, bdm, 420, prm)
    ur = oracle.frame(pair.ref, pair.tst, prm)
    same = (lambda a, b: a == b) if not erp else (lambda a, b: a == pytest.approx(b, rel=1e-12))
    assert mr.num_non_masked == w * h
    assert list(mr.gcd) == list(ur.gcd)
    assert list(mr.psnr) == list(ur.psnr)
    assert all(same(a, b) for a, b in zip(mr.wspsnr, ur.wspsnr))
    assert same(mr.r2t, ur.t2r) and same(mr.t2r, ur.r2t) and same(mr.ivpsnr, ur.ivpsnr)
    scale = (1 << bdm) - 1
    assert [list(x) for x in mr.cps_sum] == [[v * scale for v in ur.cps_sum[1]], [v * scale for v in ur.cps_sum[0]]]


def _brute_cps_masked(X, Y, M, g, B=2, wts=(4, 1, 1)):
    """Independent numpy restatement of the masked search: masked pels contribute nothing, masked
    candidates (mask read at the clamped position) are never chosen, first argmin in raster order,
    squared differences weighted by the pel's mask value."""
    _, h, w = X.shape
    Xg = X.astype(np.int64) + np.array(g[:3], np.int64)[:, None, None]
    errs, cands = [], []
    big = np.int64(1) << 62
    for dy in range(-B, B + 1):
        for dx in range(-B, B + 1):
            ys = np.clip(np.arange(h) + dy, 0, h - 1)
            xs = np.clip(np.arange(w) + dx, 0, w - 1)
            Yq = Y[:, ys][:, :, xs].astype(np.int64)
            Mq = M[ys][:, xs]
            d2 = (Xg - Yq) ** 2
            e = sum(wts[c] * d2[c] for c in range(3))
            errs.append(np.where(Mq != 0, e, big))
            cands.append(d2)
    best = np.argmin(np.stack(errs), axis=0)
    sel = np.take_along_axis(np.stack(cands), best[None, None], axis=0)[0]
    sel = sel * M.astype(np.int64)[None]
    return sel.sum(axis=2).astype(np.uint64)


@pytest.mark.parametrize("seed", range(6))
def test_masked_cps_rows_against_brute_force(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    h, w = int(rng.integers(1, 20)), int(rng.integers(1, 24))
    hi = [3, 5, 1024][seed % 3]
    X = rng.integers(0, hi, size=(3, h, w)).astype(np.uint16)
    Y = rng.integers(0, hi, size=(3, h, w)).astype(np.uint16)
    M = (rng.integers(0, 3, size=(h, w)) * rng.integers(1, 256, size=(h, w))).astype(np.uint16)
    g = np.array([int(rng.integers(-2, 3)), 0, int(rng.integers(-1, 2)), 0], np.int32)
    np.testing.assert_array_equal(oracle.cps_rows_masked(X, Y, M, g), _brute_cps_masked(X, Y, M, g))


def test_masked_gcd_and_psnr_closed_forms(oracle):
    """A constant offset k gives GCD = clip(rint(k sum(m) / (nnm MaxM))): k itself for a binary
    mask; masked PSNR of a constant error e is 10 log10(Max^2 MaxM nnm / (e^2 sum m))."""
    w, h, bd, bdm = 40, 24, 10, 8
    rng = np.random.default_rng(5)
    P0 = rng.integers(100, 900, size=(3, h, w)).astype(np.uint16)
    M = (rng.integers(0, 2, size=(h, w)) * rng.integers(1, 256, size=(h, w))).astype(np.uint16)
    nnm = int((M != 0).sum())
    binary = np.where(M != 0, 255, 0).astype(np.uint16)
    for k in (-7, 0, 4, 30):
        P1 = (P0.astype(np.int64) + k).astype(np.uint16)
        assert oracle.gcd_masked(P0, P1, binary, bd, bdm, nnm) == [max(-10, min(k, 10))] * 3 + [0]
        r = int(np.rint(k * int(M.astype(np.int64).sum()) / (nnm * 255)))
        assert oracle.gcd_masked(P0, P1, M, bd, bdm, nnm) == [max(-10, min(r, 10))] * 3 + [0]
    e = 3
    ssd = float(e * e * int(M.astype(np.int64).sum()))
    want = 10.0 * math.log10(float(nnm * 1023 * 1023 * 255) / ssd)
    assert oracle.psnr_from_masked_ssd(ssd, nnm, bd, bdm) == pytest.approx(want, rel=1e-15)
    assert oracle.gcd_masked(P0, P0, np.zeros_like(M), bd, bdm, 0) == [-10, -10, -10, 0]  # 0/0: integer indefinite
