"""ctypes declarations of the C ABI in include/qmiv.h (libqmiv.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# QMIV_LIB selects another build of the library (e.g. an experimental variant).
LIB_PATH = os.environ.get("QMIV_LIB", os.path.join(HERE, "libqmiv.so"))

QMIV_OK = 0
STATUS_NAMES = {0: "QMIV_OK", 1: "QMIV_ERR_INVALID_ARG", 2: "QMIV_ERR_UNSUPPORTED", 3: "QMIV_ERR_HIP",
                4: "QMIV_ERR_OUT_OF_MEMORY", 5: "QMIV_ERR_PEL_RANGE"}
PSNR, WSPSNR, IVPSNR, SSIM, MSSSIM, IVSSIM = 1, 2, 4, 8, 16, 32
PEL_STOP, PEL_WARN, PEL_SKIP, PEL_CNCL = 0, 1, 2, 3

# Every entry point declared in include/qmiv.h; the CPU tests check the library exports them all.
EXPORTS = []


class Params(C.Structure):
    _fields_ = [
        ("search_range", C.c_int32),
        ("cmp_weights_search", C.c_int32 * 4),
        ("cmp_weights_average", C.c_int32 * 4),
        ("unnoticeable_coef", C.c_float * 4),
        ("erp", C.c_int32), ("lon_range_deg", C.c_int32), ("lat_range_deg", C.c_int32),
        ("cps_int32_rows", C.c_int32),
        ("fake_psnr_for_exact", C.c_int32),
        ("invalid_pel_action", C.c_int32),
    ]


class FrameResult(C.Structure):
    _fields_ = [
        ("in_range", C.c_int32 * 2), ("exact", C.c_int32 * 3), ("gcd", C.c_int32 * 4),
        ("psnr", C.c_double * 3), ("psnr_pic", C.c_double),
        ("wspsnr", C.c_double * 3), ("wspsnr_pic", C.c_double),
        ("ivpsnr", C.c_double), ("ivpsnr_r2t", C.c_double), ("ivpsnr_t2r", C.c_double),
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
        ("ivpsnr", C.c_double), ("ivpsnr_r2t", C.c_double), ("ivpsnr_t2r", C.c_double),
        ("cps_sum", (C.c_uint64 * 3) * 2),
    ]
    as_dict = FrameResult.as_dict


K_STATS, K_SUMMARY, K_CPS, K_CPS_ROWS, K_SSIM, K_PYRAMID, K_SSIM_ROWS, K_CPS_GENERAL, K_COUNT = 0, 1, 2, 3, 4, 5, 6, 7, 8
KERNEL_NAMES = {K_STATS: "stats", K_SUMMARY: "summary", K_CPS: "cps", K_CPS_ROWS: "cps_rows", K_SSIM: "ssim",
                K_PYRAMID: "pyramid", K_SSIM_ROWS: "ssim_rows", K_CPS_GENERAL: "cps_general"}


class Profile(C.Structure):
    _fields_ = [("kernel_ms", C.c_double * 8), ("launches", C.c_int64 * 8)]


class BatchOptions(C.Structure):
    _fields_ = [("branches", C.c_int32), ("cps_sweep", C.c_int32), ("tune_reps", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [("views", C.c_int32), ("branches", C.c_int32), ("cps_sweep", C.c_int32),
                ("candidates_tried", C.c_int32), ("tuned_ms", C.c_double), ("tune_rounds", C.c_int32),
                ("from_cache", C.c_int32), ("instances", C.c_int32), ("instance_ms", C.c_double)]


_ctx_p = C.c_void_p
_pic_p = C.c_void_p
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)

_SIGNATURES = {
    "qmiv_create": (C.c_int, [C.c_int32, C.POINTER(_ctx_p)]),
    "qmiv_destroy": (None, [_ctx_p]),
    "qmiv_last_error": (C.c_char_p, []),
    "qmiv_abi_version": (C.c_int32, []),
    "qmiv_stream": (C.c_void_p, [_ctx_p]),
    "qmiv_sync": (C.c_int, [_ctx_p]),
    "qmiv_profile_enable": (C.c_int, [_ctx_p, C.c_int32]),
    "qmiv_profile_read": (C.c_int, [_ctx_p, C.POINTER(Profile), C.c_int32]),
    "qmiv_params_default": (None, [C.POINTER(Params)]),
    "qmiv_frame_bytes": (C.c_size_t, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "qmiv_pic_create": (C.c_int, [_ctx_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_pic_p)]),
    "qmiv_pic_wrap": (C.c_int, [_ctx_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                C.POINTER(_pic_p)]),
    "qmiv_pic_destroy": (None, [_pic_p]),
    "qmiv_pic_device_ptr": (C.c_void_p, [_pic_p]),
    "qmiv_pic_upload": (C.c_int, [_ctx_p, _pic_p, C.c_void_p, C.c_size_t]),
    "qmiv_pic_download": (C.c_int, [_ctx_p, _pic_p, C.c_void_p, C.c_size_t]),
    "qmiv_pic_upload_unpack": (C.c_int, [_ctx_p, _pic_p, C.c_void_p, C.c_size_t, C.c_int32]),
    "qmiv_pic_upload_planar": (C.c_int, [_ctx_p, _pic_p, C.POINTER(C.c_void_p), C.c_int32]),
    "qmiv_pic_download_planar": (C.c_int, [_ctx_p, _pic_p, C.POINTER(C.c_void_p), C.c_int32]),
    "qmiv_pic_upload_interleaved": (C.c_int, [_ctx_p, _pic_p, C.c_void_p, C.c_int32]),
    "qmiv_pic_download_interleaved": (C.c_int, [_ctx_p, _pic_p, C.c_void_p, C.c_int32]),
    "qmiv_pic_copy": (C.c_int, [_ctx_p, _pic_p, _pic_p]),
    "qmiv_pic_check": (C.c_int, [_ctx_p, _pic_p, _i32p]),
    "qmiv_pic_conceal": (C.c_int, [_ctx_p, _pic_p]),
    "qmiv_pic_equal": (C.c_int, [_ctx_p, _pic_p, _pic_p, _i32p]),
    "qmiv_frame": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.POINTER(Params), C.c_uint32, C.POINTER(FrameResult)]),
    "qmiv_frames": (C.c_int, [_ctx_p, C.c_int32, C.POINTER(_pic_p), C.POINTER(_pic_p), C.POINTER(Params),
                              C.c_uint32, C.POINTER(FrameResult)]),
    "qmiv_frame_noalloc": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.POINTER(Params), C.c_uint32,
                                     C.POINTER(FrameResult)]),
    "qmiv_frames_upload": (C.c_int, [_ctx_p, C.c_int32, C.POINTER(_pic_p), C.POINTER(_pic_p), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_void_p), C.POINTER(Params), C.c_uint32, C.POINTER(FrameResult)]),
    "qmiv_host_alloc": (C.c_int, [_ctx_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "qmiv_host_free": (C.c_int, [_ctx_p, C.c_void_p]),
    "qmiv_calc_gcd": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.POINTER(Params), _i32p]),
    "qmiv_calc_psnr": (C.c_int, [_ctx_p, _pic_p, _pic_p, _f64p]),
    "qmiv_calc_wspsnr": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.POINTER(Params), _f64p]),
    "qmiv_calc_ivpsnr": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.POINTER(Params), _i32p, _f64p, _f64p, _f64p,
                                   _u64p]),
    "qmiv_calc_ssim": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.c_int32, _f64p]),
    "qmiv_calc_msssim": (C.c_int, [_ctx_p, _pic_p, _pic_p, _f64p, _f64p]),
    "qmiv_calc_ivssim": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.POINTER(Params), _f64p, _f64p, _f64p]),
    "qmiv_gen_scp": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.POINTER(Params), _i32p, C.c_void_p, C.c_void_p]),
    "qmiv_gen_scp_pics": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.POINTER(Params), _i32p, _pic_p, _pic_p]),
    "qmiv_calc_ivssim_scp": (C.c_int, [_ctx_p, _pic_p, _pic_p, _pic_p, _pic_p, C.POINTER(Params), _f64p, _f64p,
                                       _f64p]),
    "qmiv_pic_upsample": (C.c_int, [_ctx_p, _pic_p, _pic_p]),
    "qmiv_seq_open": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "qmiv_seq_open_png": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "qmiv_seq_num_frames": (C.c_int32, [C.c_void_p]),
    "qmiv_seq_frame_bytes": (C.c_size_t, [C.c_void_p]),
    "qmiv_seq_read_host": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "qmiv_seq_read": (C.c_int, [_ctx_p, C.c_void_p, C.c_int32, _pic_p]),
    "qmiv_seq_close": (None, [C.c_void_p]),
    "qmiv_frame_masked": (C.c_int, [_ctx_p, _pic_p, _pic_p, _pic_p, C.POINTER(Params), C.c_uint32,
                                    C.POINTER(MaskedResult)]),
    "qmiv_frames_masked": (C.c_int, [_ctx_p, C.c_int32, C.POINTER(_pic_p), C.POINTER(_pic_p), C.POINTER(_pic_p),
                                     C.POINTER(Params), C.c_uint32, C.POINTER(MaskedResult)]),
    "qmiv_count_nonzero": (C.c_int, [_ctx_p, _pic_p, _i32p]),
    "qmiv_calc_gcd_masked": (C.c_int, [_ctx_p, _pic_p, _pic_p, _pic_p, C.POINTER(Params), C.c_int32, _i32p]),
    "qmiv_calc_psnr_masked": (C.c_int, [_ctx_p, _pic_p, _pic_p, _pic_p, C.POINTER(Params), C.c_int32, _f64p]),
    "qmiv_calc_wspsnr_masked": (C.c_int, [_ctx_p, _pic_p, _pic_p, _pic_p, C.POINTER(Params), C.c_int32, _f64p]),
    "qmiv_calc_ivpsnr_masked": (C.c_int, [_ctx_p, _pic_p, _pic_p, _pic_p, C.POINTER(Params), C.c_int32, _i32p,
                                          _f64p, _f64p, _f64p]),
    "qmiv_pic_colorspace": (C.c_int, [_ctx_p, _pic_p, _pic_p, C.c_int32, C.c_int32]),
    "qmiv_pic_swap_planes": (C.c_int, [_ctx_p, _pic_p, C.c_int32, C.c_int32]),
    "qmiv_batch_options_default": (None, [C.POINTER(BatchOptions)]),
    "qmiv_batch_create": (C.c_int, [_ctx_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                    C.POINTER(Params), C.POINTER(C.c_uint32), C.POINTER(BatchOptions),
                                    C.POINTER(C.c_void_p)]),
    "qmiv_batch_run": (C.c_int, [C.c_void_p, C.POINTER(FrameResult)]),
    "qmiv_batch_launch": (C.c_int, [C.c_void_p]),
    "qmiv_batch_collect": (C.c_int, [C.c_void_p, C.POINTER(FrameResult)]),
    "qmiv_batch_create_masked": (C.c_int, [_ctx_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_void_p), C.POINTER(Params), C.c_uint32,
                                           C.POINTER(BatchOptions), C.POINTER(C.c_void_p)]),
    "qmiv_batch_run_masked": (C.c_int, [C.c_void_p, C.POINTER(MaskedResult)]),
    "qmiv_batch_collect_masked": (C.c_int, [C.c_void_p, C.POINTER(MaskedResult)]),
    "qmiv_batch_get_info": (C.c_int, [C.c_void_p, C.POINTER(BatchInfo)]),
    "qmiv_batch_destroy": (None, [C.c_void_p]),
    "qmiv_ssim_backward": (C.c_int, [_ctx_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_int32, _f64p, C.c_void_p, C.c_void_p]),
    "qmiv_msssim_backward": (C.c_int, [_ctx_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_double,
                                       _f64p, _f64p, C.c_void_p, C.c_void_p]),
    "qmiv_ivssim_backward": (C.c_int, [_ctx_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_double,
                                       C.POINTER(Params), _f64p, _f64p, _f64p, C.c_void_p, C.c_void_p]),
    "qmiv_ivssim_backward_batch": (C.c_int, [_ctx_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                                             C.POINTER(C.c_void_p), C.c_double, C.POINTER(Params), _f64p, _f64p,
                                             _f64p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
}
EXPORTS.extend(_SIGNATURES)

_lib = None


def lib():
    """The loaded engine library. Raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"engine library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class QmivError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = lib().qmiv_last_error().decode(errors="replace")
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def check(status: int, where: str):
    if status != QMIV_OK:
        raise QmivError(status, where)


def default_params(**kw) -> Params:
    p = Params()
    lib().qmiv_params_default(C.byref(p))
    for k, v in kw.items():
        if k in ("cmp_weights_search", "cmp_weights_average"):
            for i in range(4):
                getattr(p, k)[i] = int(v[i])
        elif k == "unnoticeable_coef":
            for i in range(4):
                p.unnoticeable_coef[i] = float(v[i])
        else:
            setattr(p, k, int(v))
    return p
