"""Shared helpers for the parity tests (comparison of engine results with the oracle's)."""
from __future__ import annotations

import numpy as np

FIELDS_PSNR = ("in_range", "exact", "gcd", "psnr", "psnr_pic", "wspsnr", "wspsnr_pic")
FIELDS_IV = ("ivpsnr", "cps_sum")


def oracle_params(oracle, w, h, bd, cf, metrics, params=None, erp=False):
    kw = {}
    if params is not None:
        kw = dict(search_range=params.search_range,
                  weights_search=tuple(params.cmp_weights_search),
                  weights_average=tuple(params.cmp_weights_average),
                  unnoticeable=tuple(params.unnoticeable_coef),
                  int32_rows=bool(params.cps_int32_rows),
                  erp=bool(params.erp), lat_range=params.lat_range_deg)
    else:
        kw["erp"] = erp
    return oracle.params(w, h, bd, cf, metrics=metrics, **kw)


def assert_frame_equal(gpu, orc, metrics, ctx=""):
    """Bit-exact comparison of the integer and PSNR-family fields of two frame results."""
    g = gpu.as_dict()
    o = orc.as_dict()
    g["ivpsnr_r2t"], g["ivpsnr_t2r"] = g["ivpsnr_r2t"], g["ivpsnr_t2r"]
    pairs = [("in_range", "in_range"), ("exact", "exact")]
    if metrics & (4 | 32):
        pairs.append(("gcd", "gcd"))
    if metrics & 1:
        pairs += [("psnr", "psnr"), ("psnr_pic", "psnr_pic")]
    if metrics & 2:
        pairs += [("wspsnr", "wspsnr"), ("wspsnr_pic", "wspsnr_pic")]
    if metrics & 4:
        pairs += [("cps_sum", "cps_sum"), ("ivpsnr", "ivpsnr"), ("ivpsnr_r2t", "r2t"), ("ivpsnr_t2r", "t2r")]
    for gk, ok in pairs:
        gv, ov = g[gk], o[ok]
        assert gv == ov, f"{ctx}: {gk} differs: engine {gv!r} oracle {ov!r}"


def planes444(oracle, packed, w, h, bd, cf):
    return oracle.unpack(np.ascontiguousarray(packed), w, h, bd, cf)


def ssim_window_table():
    """The 10 x 10 SSIM window (rows dy, columns dx) as doubles: the normalised 11 x 11 Gaussian
    (sigma 1.5) with each tap rounded to ten decimals and stored in fp32
    (libs/pmbb-ivqm/src/xStructSimConsts.h:30-43), first 10 rows and columns
    (libs/pmbb-ivqm/src/xStructSim.cpp:19-21). Pinned by tests/golden/ssim_filter_bits.json."""
    import numpy as np
    g = np.exp(-((np.arange(11) - 5.0) ** 2) / (2 * 1.5 * 1.5))
    m = np.outer(g, g)
    m = m / m.sum()
    taps = np.array([np.float32(float("%.10f" % v)) for v in m.reshape(-1)], np.float32).reshape(11, 11)
    return taps[:10, :10].astype(np.float64)
