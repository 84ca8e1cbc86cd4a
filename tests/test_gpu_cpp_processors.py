"""The C++ processor classes (include/qmiv.hpp) on the GPU, in the reference application's call
order, against the oracle: PSNR family and GCD bit-exact, SSIM family to 1e-9 relative, the
shift-compensated pictures and the RGB -> YCbCr conversion bit-exact. The program also asserts the processor API's own
identities (auto-GCD IV-PSNR, two- vs four-argument IV-SSIM, error on incompatible pictures)."""
import json
import os
import subprocess

import numpy as np
import pytest

from ivpsnr_amd.synth import make_mask, make_pair

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_processors")


def close(a, b, rel=1e-9):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return bool(np.all(np.abs(a - b) <= rel * np.abs(b)))


@pytest.mark.parametrize("w,h,bd,cf,erp,seed", [(320, 192, 10, 420, True, 1), (256, 128, 8, 420, False, 2),
                                                (96, 64, 12, 444, False, 3), (1920, 1080, 10, 420, False, 4)])
def test_processor_classes(oracle, tmp_path, w, h, bd, cf, erp, seed):
    run_processors(oracle, tmp_path, w, h, bd, cf, erp, seed, masked=seed % 2 == 1)


def run_processors(oracle, tmp_path, w, h, bd, cf, erp, seed, masked):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (make -C tests/cpp)")
    pair = make_pair(w, h, bd, seed, chroma_format=cf)
    f0, f1 = tmp_path / "p0.yuv", tmp_path / "p1.yuv"
    s0, s1, sc = tmp_path / "scp0.bin", tmp_path / "scp1.bin", tmp_path / "clr.bin"
    f0.write_bytes(pair.ref.tobytes())
    f1.write_bytes(pair.tst.tobytes())
    print(f"case {w}x{h} bd{bd} cf{cf} erp{int(erp)} seed{seed}", flush=True)
    extra = []
    if masked:
        fm = tmp_path / "m.yuv"
        mask = make_mask(w, h, 8, seed, chroma_format=420, kind="soft")
        fm.write_bytes(mask.tobytes())
        extra = [str(fm), "8", "420"]
    r = subprocess.run([BIN, str(w), str(h), str(bd), str(cf), str(f0), str(f1), str(int(erp)), str(s0), str(s1),
                        str(sc)] + extra,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    assert got["failures"] == 0

    m = oracle.PSNR | oracle.WSPSNR | oracle.IVPSNR | oracle.SSIM | oracle.MSSSIM | oracle.IVSSIM
    o, scp0, scp1 = oracle.frame(pair.ref, pair.tst, oracle.params(w, h, bd, cf, metrics=m, erp=erp), want_scp=True)
    assert not any(o.exact)  # the processor classes return DBL_MAX for exact components, the app a fake value
    assert got["gcd"] == list(o.gcd)
    assert got["psnr"] == list(o.psnr) and got["wspsnr"] == list(o.wspsnr)
    assert [got["ivpsnr"], got["ivpsnr_r2t"], got["ivpsnr_t2r"]] == [o.ivpsnr, o.r2t, o.t2r]
    assert got["ivpsnr_interleaved"] == o.ivpsnr  # calcPicIVPSNR(const xPicI*, const xPicI*, GCD)
    assert close(got["ssim"], list(o.ssim)) and close(got["msssim"], list(o.msssim))
    assert close([got["ivssim"], got["ivssim_r2t"], got["ivssim_t2r"]], [o.ivssim, o.ivssim_r2t, o.ivssim_t2r])
    P0, P1 = oracle.unpack(pair.ref, w, h, bd, cf), oracle.unpack(pair.tst, w, h, bd, cf)
    assert close(got["ssim_cs"], [oracle.ssim_plane(P0[c], P1[c], bd, False) for c in range(3)])
    n = 3 * w * h
    np.testing.assert_array_equal(np.fromfile(s0, np.uint16, n).reshape(3, h, w), scp0)
    np.testing.assert_array_equal(np.fromfile(s1, np.uint16, n).reshape(3, h, w), scp1)
    np.testing.assert_array_equal(np.fromfile(sc, np.uint16, n).reshape(3, h, w), oracle.colorspace(P0, bd, 1, False))
    if masked:
        # The processors start without fake values (xPSNR's default), the application flow with them
        # for identical components: none here.
        mo = oracle.frame_masked(pair.ref, pair.tst, mask, 8, 420, oracle.params(w, h, bd, cf, metrics=7, erp=erp))
        gm = got["masked"]
        assert gm["nnm"] == mo.num_non_masked and gm["gcd"] == list(mo.gcd)
        assert gm["psnr"] == list(mo.psnr) and gm["wspsnr"] == list(mo.wspsnr)
        assert [gm["ivpsnr"], gm["r2t"], gm["t2r"]] == [mo.ivpsnr, mo.r2t, mo.t2r]
