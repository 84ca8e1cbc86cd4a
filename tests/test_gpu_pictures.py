"""Picture operations on the GPU: 4:2:0 -> 4:4:4 upsampling against the oracle's unpack, SCP into
device pictures against the host-array form, IV-SSIM from caller-provided SCP pictures against
the two-argument form, and metrics on upsampled 4:4:4 pictures equal to those on 4:2:0."""
import numpy as np
import pytest

from ivpsnr_amd import ALL_METRICS, default_params
from ivpsnr_amd.synth import make_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,bd", [(64, 32, 8), (130, 66, 10), (1920, 1080, 16)])
def test_upsample_matches_unpack(engine, oracle, w, h, bd):
    pair = make_pair(w, h, bd, 3)
    p = engine.picture(w, h, bd, 420, data=pair.ref)
    u = engine.upsample(p)
    got = u.download().astype(np.uint16).reshape(3, h, w)
    np.testing.assert_array_equal(got, oracle.unpack(pair.ref, w, h, bd))


@pytest.mark.parametrize("bd", [8, 10])
def test_scp_pictures_and_ivssim_scp(engine, bd):
    w, h = 200, 120
    pair = make_pair(w, h, bd, 9)
    p0 = engine.upsample(engine.picture(w, h, bd, 420, data=pair.ref))
    p1 = engine.upsample(engine.picture(w, h, bd, 420, data=pair.tst))
    prm = default_params()
    g = engine.calc_gcd(p0, p1, prm)
    ref_arr, tst_arr = engine.gen_scp(p1, p0, g, prm)
    scp1 = engine.picture(w, h, bd, 444)
    scp0 = engine.picture(w, h, bd, 444)
    engine.gen_scp_pics(p1, p0, g, scp1, scp0, prm)
    np.testing.assert_array_equal(scp1.download().astype(np.uint16).reshape(3, h, w), ref_arr)
    np.testing.assert_array_equal(scp0.download().astype(np.uint16).reshape(3, h, w), tst_arr)
    four = engine.calc_ivssim_scp(p0, p1, scp0, scp1, prm)
    two = engine.calc_ivssim(p0, p1, prm)
    assert four == two


def test_metrics_on_upsampled_pictures_equal_420(engine):
    w, h, bd = 256, 144, 10
    pair = make_pair(w, h, bd, 4)
    a = engine.picture(w, h, bd, 420, data=pair.ref)
    b = engine.picture(w, h, bd, 420, data=pair.tst)
    prm = default_params(erp=True)
    r420 = engine.frame(a, b, ALL_METRICS, prm)
    r444 = engine.frame(engine.upsample(a), engine.upsample(b), ALL_METRICS, prm)
    for f in ("gcd", "psnr", "wspsnr", "ssim", "msssim"):
        assert list(getattr(r420, f)) == list(getattr(r444, f)), f
    for f in ("ivpsnr", "ivpsnr_r2t", "ivpsnr_t2r", "ivssim", "ivssim_r2t", "ivssim_t2r"):
        assert getattr(r420, f) == getattr(r444, f), f
