"""The allocation-free frame entry point (qmiv_frame_noalloc) against qmiv_frame.

qmiv_frame_noalloc runs the same per-frame flow out of the device's fixed scratch block and the
context's fixed pinned results block. Two calls in a row, the second on a smaller frame, must give
exactly qmiv_frame's results (the same kernels in the same order: every field bit-identical), so
leftovers of a larger frame in the shared block cannot leak into a smaller one. A frame larger than
the block is refused with QMIV_ERR_UNSUPPORTED before any launch.
"""
import pytest

from ivpsnr_amd import ALL_METRICS, IVPSNR, IVSSIM, PSNR, WSPSNR, default_params
from ivpsnr_amd.capi import QmivError
from ivpsnr_amd.synth import make_pair

pytestmark = pytest.mark.gpu


def _pictures(engine, w, h, bd, seed):
    pair = make_pair(w, h, bd, seed=seed)
    return (engine.picture(w, h, bd, 420, data=pair.ref), engine.picture(w, h, bd, 420, data=pair.tst))


def _close(pics):
    for p in pics:
        p.close()


def test_two_calls_second_smaller_equal_qmiv_frame(engine):
    cases = [
        # (w, h, bd, metrics, ERP)
        (1920, 1080, 10, ALL_METRICS, False),
        (640, 360, 8, ALL_METRICS, True),
    ]
    pics = [_pictures(engine, w, h, bd, seed=40 + i) for i, (w, h, bd, _, _) in enumerate(cases)]
    prms = [default_params(erp=erp) for (_, _, _, _, erp) in cases]
    want = [engine.frame(p0, p1, m, prm).as_dict() for (p0, p1), (_, _, _, m, _), prm in zip(pics, cases, prms)]
    # Back to back through the fixed block: large frame, then the smaller one.
    got = [engine.frame_noalloc(p0, p1, m, prm).as_dict() for (p0, p1), (_, _, _, m, _), prm in zip(pics, cases, prms)]
    for i, (g, w) in enumerate(zip(got, want)):
        assert g == w, f"case {cases[i][:3]}: qmiv_frame_noalloc differs from qmiv_frame"
        assert g["ivpsnr"] > 0 and 0 < g["ssim_pic"] < 1 and 0 < g["ivssim"] < 1
    # And again in the other order, the small frame first, then the large one.
    got = [engine.frame_noalloc(p0, p1, m, prm).as_dict()
           for (p0, p1), (_, _, _, m, _), prm in reversed(list(zip(pics, cases, prms)))]
    assert got[::-1] == want
    for pr in pics:
        _close(pr)


def test_4k_psnr_family_fits_and_4k_ivssim_is_refused(engine):
    w, h, bd = 4096, 2048, 10
    p0, p1 = _pictures(engine, w, h, bd, seed=7)
    prm = default_params(erp=True)
    metrics = PSNR | WSPSNR | IVPSNR
    want = engine.frame(p0, p1, metrics, prm).as_dict()
    assert engine.frame_noalloc(p0, p1, metrics, prm).as_dict() == want
    with pytest.raises(QmivError) as ei:
        engine.frame_noalloc(p0, p1, metrics | IVSSIM, prm)
    assert ei.value.status == 2 and "bytes of device scratch" in str(ei.value)
    # The refused call launched nothing: the block still serves the next frame.
    s0, s1 = _pictures(engine, 256, 128, 10, seed=8)
    assert engine.frame_noalloc(s0, s1, ALL_METRICS).as_dict() == engine.frame(s0, s1, ALL_METRICS).as_dict()
    _close((p0, p1, s0, s1))
