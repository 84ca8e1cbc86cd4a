"""The C-ABI library loads and exports every entry point include/qmiv.h declares (CPU only)."""
import ctypes
import os
import re

from ivpsnr_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "qmiv.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qmiv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_entry_points()
    assert "qmiv_frame" in names and "qmiv_create" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_entry_points() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    assert sorted(capi.EXPORTS) == declared_entry_points()


def test_cpu_only_calls():
    lib = capi.lib()
    assert lib.qmiv_abi_version() == 2
    assert lib.qmiv_frame_bytes(1920, 1080, 8, 420) == 1920 * 1080 * 3 // 2
    assert lib.qmiv_frame_bytes(4096, 2048, 10, 420) == 4096 * 2048 * 3
    assert lib.qmiv_frame_bytes(67, 45, 8, 444) == 67 * 45 * 3
    # raw-file frames of every chroma format, odd sizes too (xSeq::create: luma + chroma bytes)
    assert lib.qmiv_frame_bytes(67, 45, 8, 420) == 67 * 45 + (67 * 45) // 2
    assert lib.qmiv_frame_bytes(67, 45, 10, 422) == 2 * (2 * 67 * 45)
    assert lib.qmiv_frame_bytes(67, 45, 10, 400) == 2 * 67 * 45
    assert lib.qmiv_frame_bytes(64, 48, 8, 411) == 0  # not a chroma format
    p = capi.default_params()
    assert p.search_range == 2 and list(p.cmp_weights_search) == [4, 1, 1, 0]
    assert abs(p.unnoticeable_coef[0] - 0.01) < 1e-9 and p.cps_int32_rows == 1
