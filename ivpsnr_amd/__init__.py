"""MI355X (gfx950) per-frame metric engine: IV-PSNR, PSNR, WS-PSNR, SSIM, MS-SSIM, IV-SSIM.

The computation lives in libqmiv.so (hand-written HIP kernels behind the C ABI of
include/qmiv.h); this package is the Python view of that ABI plus the synthetic input
generator used by the tests and the benchmark.
"""
from .capi import IVPSNR, IVSSIM, MSSSIM, PSNR, SSIM, WSPSNR  # noqa: F401
from .engine import Engine, Picture, Sequence, default_params  # noqa: F401

ALL_METRICS = PSNR | WSPSNR | IVPSNR | SSIM | MSSSIM | IVSSIM

__all__ = ["Engine", "Picture", "Sequence", "default_params", "ALL_METRICS", "PSNR", "WSPSNR", "IVPSNR", "SSIM", "MSSSIM", "IVSSIM"]
