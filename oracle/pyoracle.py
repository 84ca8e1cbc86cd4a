"""ctypes binding of the plain-C oracle (oracle/_build/libqmiv_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker and never as the thing
measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libqmiv_oracle.so")

PSNR, WSPSNR, IVPSNR, SSIM, MSSSIM, IVSSIM = 1, 2, 4, 8, 16, 32
ALL = PSNR | WSPSNR | IVPSNR | SSIM | MSSSIM | IVSSIM


class Params(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("bit_depth", C.c_int32), ("chroma_format", C.c_int32),
        ("search_range", C.c_int32),
        ("cmp_weights_search", C.c_int32 * 4),
        ("cmp_weights_average", C.c_int32 * 4),
        ("unnoticeable_coef", C.c_float * 4),
        ("erp", C.c_int32), ("lon_range_deg", C.c_int32), ("lat_range_deg", C.c_int32),
        ("cps_int32_rows", C.c_int32),
        ("metrics", C.c_int32),
        ("conceal", C.c_int32),
    ]


class FrameResult(C.Structure):
    _fields_ = [
        ("in_range", C.c_int32 * 2), ("exact", C.c_int32 * 3), ("gcd", C.c_int32 * 4),
        ("psnr", C.c_double * 3), ("psnr_pic", C.c_double),
        ("wspsnr", C.c_double * 3), ("wspsnr_pic", C.c_double),
        ("ivpsnr", C.c_double), ("r2t", C.c_double), ("t2r", C.c_double),
        ("cps_sum", (C.c_uint64 * 3) * 2),
        ("ssim", C.c_double * 3), ("ssim_pic", C.c_double),
        ("msssim", C.c_double * 3), ("msssim_pic", C.c_double),
        ("msssim_scales", (C.c_double * 3) * 5),
        ("ivssim", C.c_double), ("ivssim_r2t", C.c_double), ("ivssim_t2r", C.c_double),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if hasattr(v, "__len__"):
                v = [list(x) if hasattr(x, "__len__") else x for x in v]
            out[name] = v
        return out


class MaskedResult(C.Structure):
    _fields_ = [
        ("in_range", C.c_int32 * 3), ("exact", C.c_int32 * 3), ("num_non_masked", C.c_int32),
        ("gcd", C.c_int32 * 4),
        ("psnr", C.c_double * 3), ("psnr_pic", C.c_double),
        ("wspsnr", C.c_double * 3), ("wspsnr_pic", C.c_double),
        ("ivpsnr", C.c_double), ("r2t", C.c_double), ("t2r", C.c_double),
        ("cps_sum", (C.c_uint64 * 3) * 2),
    ]
    as_dict = FrameResult.as_dict


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        u16p = np.ctypeslib.ndpointer(np.uint16, flags="C")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        P3 = C.c_void_p * 3
        L.orc_default_params.argtypes = [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.orc_unpack.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, u16p, u16p, u16p]
        L.orc_unpack.restype = C.c_int
        L.orc_gcd.argtypes = [P3, P3, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_float * 4), i32p]
        L.orc_cps_rows.argtypes = [P3, P3, C.c_int32, C.c_int32, i32p, C.c_int32, i32p, C.c_int32, u64p, C.c_void_p]
        L.orc_scp.argtypes = [P3, P3, C.c_int32, C.c_int32, C.c_int32, i32p, C.c_int32, i32p, P3]
        L.orc_ssd_rows.argtypes = [u16p, u16p, C.c_int32, C.c_int32, u64p]
        L.orc_psnr_from_ssd.argtypes = [C.c_double, C.c_int32, C.c_int32]
        L.orc_psnr_from_ssd.restype = C.c_double
        L.orc_ws_init.argtypes = [C.c_int32, C.c_int32, f64p, C.POINTER(C.c_double)]
        L.orc_kbns.argtypes = [f64p, C.c_size_t]
        L.orc_kbns.restype = C.c_double
        L.orc_qual_from_rows.argtypes = [u64p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, i32p]
        L.orc_qual_from_rows.restype = C.c_double
        L.orc_ssim_filter.argtypes = [np.ctypeslib.ndpointer(np.float32, flags="C")]
        L.orc_ssim_plane.argtypes = [u16p, u16p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.orc_ssim_plane.restype = C.c_double
        L.orc_downsample.argtypes = [u16p, u16p, C.c_int32, C.c_int32, C.c_int32]
        L.orc_msssim_plane.argtypes = [u16p, u16p, C.c_int32, C.c_int32, C.c_int32, f64p]
        L.orc_msssim_plane.restype = C.c_double
        L.orc_frame.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Params), C.POINTER(FrameResult), C.c_void_p,
                                C.c_void_p]
        L.orc_frame.restype = C.c_int
        L.orc_frame_masked.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(Params),
                                       C.POINTER(MaskedResult)]
        L.orc_frame_masked.restype = C.c_int
        L.orc_gcd_masked.argtypes = [P3, P3, u16p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                     C.POINTER(C.c_float * 4), i32p]
        L.orc_cps_rows_masked.argtypes = [P3, P3, u16p, C.c_int32, C.c_int32, i32p, C.c_int32, i32p, u64p]
        L.orc_psnr_from_masked_ssd.argtypes = [C.c_double, C.c_int32, C.c_int32, C.c_int32]
        L.orc_psnr_from_masked_ssd.restype = C.c_double
        L.orc_frame_average.argtypes = [f64p, C.c_size_t]
        L.orc_frame_average.restype = C.c_double
        L.orc_colorspace_coeffs.argtypes = [C.c_int32, C.c_int32, i32p]
        L.orc_colorspace_coeffs.restype = C.c_int
        L.orc_colorspace.argtypes = [u16p, u16p, C.c_size_t, C.c_int32, C.c_int32, C.c_int32]
        L.orc_colorspace.restype = C.c_int
        _lib = L
    return _lib


def params(width, height, bit_depth, chroma_format=420, *, metrics=PSNR | WSPSNR | IVPSNR, erp=False,
           lat_range=180, lon_range=360, search_range=2, weights_search=(4, 1, 1, 0),
           weights_average=(4, 1, 1, 0), unnoticeable=(0.01, 0.01, 0.01, 0.0), int32_rows=True,
           conceal=False) -> Params:
    p = Params()
    lib().orc_default_params(C.byref(p), width, height, bit_depth, chroma_format)
    p.metrics = metrics
    p.erp = int(erp)
    p.lat_range_deg = lat_range
    p.lon_range_deg = lon_range
    p.search_range = search_range
    for i in range(4):
        p.cmp_weights_search[i] = weights_search[i]
        p.cmp_weights_average[i] = weights_average[i]
        p.unnoticeable_coef[i] = unnoticeable[i]
    p.cps_int32_rows = int(int32_rows)
    p.conceal = int(conceal)
    return p


def frame(p0: np.ndarray, p1: np.ndarray, prm: Params, want_scp: bool = False):
    """Run one frame pair (packed arrays) through the oracle's frame flow."""
    r = FrameResult()
    n = prm.width * prm.height
    p0 = np.ascontiguousarray(p0)
    p1 = np.ascontiguousarray(p1)
    scp0 = scp1 = None
    if want_scp:
        scp0 = np.zeros(3 * n, np.uint16)
        scp1 = np.zeros(3 * n, np.uint16)
    rc = lib().orc_frame(p0.ctypes.data, p1.ctypes.data, C.byref(prm), C.byref(r),
                         scp0.ctypes.data if want_scp else None, scp1.ctypes.data if want_scp else None)
    if rc != 0:
        raise ValueError(f"orc_frame failed ({rc})")
    if want_scp:
        return r, scp0.reshape(3, prm.height, prm.width), scp1.reshape(3, prm.height, prm.width)
    return r


def frame_masked(p0: np.ndarray, p1: np.ndarray, mask: np.ndarray, bdm: int, cfm: int, prm: Params) -> MaskedResult:
    """Run one frame pair with a mask (packed arrays) through the oracle's masked flow."""
    r = MaskedResult()
    p0, p1, mask = (np.ascontiguousarray(a) for a in (p0, p1, mask))
    rc = lib().orc_frame_masked(p0.ctypes.data, p1.ctypes.data, mask.ctypes.data, bdm, cfm, C.byref(prm), C.byref(r))
    if rc != 0:
        raise ValueError(f"orc_frame_masked failed ({rc})")
    return r


def cps_rows_masked(x: np.ndarray, y: np.ndarray, mask: np.ndarray, g, B: int = 2, wts=(4, 1, 1, 0)) -> np.ndarray:
    """Masked CPS rows (3, h) of 4:4:4 planes x, y (3, h, w) with a mask plane (h, w)."""
    _, h, w = x.shape
    x = np.ascontiguousarray(x, np.uint16)
    y = np.ascontiguousarray(y, np.uint16)
    mask = np.ascontiguousarray(mask, np.uint16)
    rows = np.zeros(3 * h, np.uint64)
    lib().orc_cps_rows_masked(_planes3(x), _planes3(y), mask.reshape(-1), w, h, np.asarray(g, np.int32),
                              B, np.asarray(wts, np.int32), rows)
    return rows.reshape(3, h)


def gcd_masked(p0: np.ndarray, p1: np.ndarray, mask: np.ndarray, bd: int, bdm: int, nnm: int,
               ucoef=(0.01, 0.01, 0.01, 0.0)) -> list:
    _, h, w = p0.shape
    g = np.zeros(4, np.int32)
    uc = (C.c_float * 4)(*ucoef)
    lib().orc_gcd_masked(_planes3(np.ascontiguousarray(p0, np.uint16)), _planes3(np.ascontiguousarray(p1, np.uint16)),
                         np.ascontiguousarray(mask, np.uint16).reshape(-1), w, h, bd, bdm, nnm, C.byref(uc), g)
    return g.tolist()


def psnr_from_masked_ssd(ssd: float, nnm: int, bd: int, bdm: int) -> float:
    return lib().orc_psnr_from_masked_ssd(ssd, nnm, bd, bdm)


def unpack(packed: np.ndarray, w: int, h: int, bd: int, cf: int = 420) -> np.ndarray:
    out = np.zeros((3, h, w), np.uint16)
    packed = np.ascontiguousarray(packed)
    rc = lib().orc_unpack(packed.ctypes.data, w, h, bd, cf, out[0], out[1], out[2])
    if rc:
        raise ValueError("unpack failed")
    return out


def _planes3(p: np.ndarray):
    arr = (C.c_void_p * 3)()
    for i in range(3):
        arr[i] = p[i].ctypes.data
    return arr


def gcd(P0: np.ndarray, P1: np.ndarray, bd: int, unnoticeable=(0.01, 0.01, 0.01, 0.0)) -> np.ndarray:
    P0 = np.ascontiguousarray(P0)
    P1 = np.ascontiguousarray(P1)
    g = np.zeros(4, np.int32)
    uc = (C.c_float * 4)(*unnoticeable)
    lib().orc_gcd(_planes3(P0), _planes3(P1), P0.shape[2], P0.shape[1], bd, C.byref(uc), g)
    return g


def cps_rows(X: np.ndarray, Y: np.ndarray, g, B=2, weights=(4, 1, 1, 0), int32_rows=True, want_best=False):
    X = np.ascontiguousarray(X)
    Y = np.ascontiguousarray(Y)
    h, w = X.shape[1:]
    rows = np.zeros(3 * h, np.uint64)
    best = np.zeros(h * w, np.int32) if want_best else None
    lib().orc_cps_rows(_planes3(X), _planes3(Y), w, h, np.asarray(g, np.int32), B, np.asarray(weights, np.int32),
                       int(int32_rows), rows, best.ctypes.data if want_best else None)
    rows = rows.reshape(3, h)
    return (rows, best.reshape(h, w)) if want_best else rows


def scp(Y: np.ndarray, X: np.ndarray, bd: int, g, B=2, weights=(4, 1, 1, 0)) -> np.ndarray:
    Y = np.ascontiguousarray(Y)
    X = np.ascontiguousarray(X)
    h, w = X.shape[1:]
    dst = np.zeros((3, h, w), np.uint16)
    lib().orc_scp(_planes3(Y), _planes3(X), w, h, bd, np.asarray(g, np.int32), B, np.asarray(weights, np.int32),
                  _planes3(dst))
    return dst


def ssd_rows(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint16)
    b = np.ascontiguousarray(b, np.uint16)
    rows = np.zeros(a.shape[0], np.uint64)
    lib().orc_ssd_rows(a, b, a.shape[1], a.shape[0], rows)
    return rows


def psnr_from_ssd(ssd: float, area: int, bd: int) -> float:
    return lib().orc_psnr_from_ssd(float(ssd), area, bd)


def ws_init(h: int, lat_range: int = 180):
    w = np.zeros(h, np.float64)
    corr = C.c_double()
    lib().orc_ws_init(h, lat_range, w, C.byref(corr))
    return w, corr.value


def kbns(x) -> float:
    x = np.ascontiguousarray(x, np.float64)
    return lib().orc_kbns(x, x.size)


def qual_from_rows(rows: np.ndarray, area: int, bd: int, ws=None, wavg=(4, 1, 1, 0)) -> float:
    rows = np.ascontiguousarray(rows, np.uint64)
    h = rows.shape[1]
    wsa = None if ws is None else np.ascontiguousarray(ws, np.float64)
    return lib().orc_qual_from_rows(rows.ravel(), h, None if wsa is None else wsa.ctypes.data, area, bd,
                                    np.asarray(wavg, np.int32))


def ssim_filter() -> np.ndarray:
    f = np.zeros(121, np.float32)
    lib().orc_ssim_filter(f)
    return f.reshape(11, 11)


def ssim_plane(tst: np.ndarray, ref: np.ndarray, bd: int, calc_l: bool = True, want_rows: bool = False):
    tst = np.ascontiguousarray(tst, np.uint16)
    ref = np.ascontiguousarray(ref, np.uint16)
    h, w = tst.shape
    rows = np.zeros(h, np.float64) if want_rows else None
    v = lib().orc_ssim_plane(tst, ref, w, h, bd, int(calc_l), rows.ctypes.data if want_rows else None)
    return (v, rows) if want_rows else v


def downsample(src: np.ndarray) -> np.ndarray:
    src = np.ascontiguousarray(src, np.uint16)
    h, w = src.shape
    dst = np.zeros((h >> 1, w >> 1), np.uint16)
    lib().orc_downsample(dst, src, w >> 1, h >> 1, w)
    return dst


def msssim_plane(tst: np.ndarray, ref: np.ndarray, bd: int):
    tst = np.ascontiguousarray(tst, np.uint16)
    ref = np.ascontiguousarray(ref, np.uint16)
    h, w = tst.shape
    scales = np.zeros(5, np.float64)
    v = lib().orc_msssim_plane(tst, ref, w, h, bd, scales)
    return v, scales


def frame_average(values) -> float:
    v = np.ascontiguousarray(values, np.float64)
    return lib().orc_frame_average(v, v.size)


def colorspace_coeffs(space: int, to_rgb: bool) -> np.ndarray:
    """Q16 matrix (3 x 3) of colour space 0..3 (BT.601, BT.709, SMPTE 240M, BT.2020)."""
    m = np.zeros(9, np.int32)
    if lib().orc_colorspace_coeffs(space, int(to_rgb), m):
        raise ValueError("unknown colour space")
    return m.reshape(3, 3)


def colorspace(planes: np.ndarray, bd: int, space: int, to_rgb: bool) -> np.ndarray:
    """Three planar uint16 planes (3, h, w) converted RGB -> YCbCr (to_rgb False) or back."""
    src = np.ascontiguousarray(planes, np.uint16)
    dst = np.zeros_like(src)
    if lib().orc_colorspace(src.reshape(-1), dst.reshape(-1), src[0].size, bd, space, int(to_rgb)):
        raise ValueError("unknown colour space")
    return dst
