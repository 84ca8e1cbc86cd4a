"""Python mirror of the engine's C ABI (include/qmiv.h), used by the tests, the smoke check and the
benchmark. Every call goes through libqmiv.so; nothing here computes a metric.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

from . import capi
from .capi import FrameResult, Params, check, default_params, lib

__all__ = ["Engine", "Picture", "default_params", "FrameResult", "Params"]


class Picture:
    """One packed frame in device memory (owned by the engine, or wrapping caller memory)."""

    def __init__(self, engine: "Engine", handle: int, width: int, height: int, bit_depth: int,
                 chroma_format: int, keepalive=None):
        self.engine = engine
        self.handle = C.c_void_p(handle)
        self.width, self.height = width, height
        self.bit_depth, self.chroma_format = bit_depth, chroma_format
        self.nbytes = lib().qmiv_frame_bytes(width, height, bit_depth, chroma_format)
        self._keepalive = keepalive

    def upload(self, packed: np.ndarray) -> "Picture":
        a = np.ascontiguousarray(packed)
        if a.nbytes != self.nbytes:
            raise ValueError(f"frame has {a.nbytes} bytes, expected {self.nbytes}")
        check(lib().qmiv_pic_upload(self.engine.ctx, self.handle, a.ctypes.data, a.nbytes), "qmiv_pic_upload")
        # The copy is ordered on the engine stream; keep the source alive until it completes.
        self.engine.sync()
        return self

    def upload_unpack(self, frame: np.ndarray, file_chroma_format: int) -> "Picture":
        """One raw-file frame (bytes) of another chroma format unpacked into this 4:4:4 picture on
        the device (qmiv_pic_upload_unpack)."""
        a = np.ascontiguousarray(frame)
        check(lib().qmiv_pic_upload_unpack(self.engine.ctx, self.handle, a.ctypes.data, a.nbytes, file_chroma_format),
              "qmiv_pic_upload_unpack")
        self.engine.sync()
        return self

    def upload_planar(self, planes: np.ndarray, margin: int = 0) -> "Picture":
        """The reference's planar layout: planes (3, H + 2 margin, W + 2 margin) uint16, margin not read."""
        a = np.ascontiguousarray(planes, np.uint16)
        stride = a.shape[2]
        ptrs = (C.c_void_p * 3)(*[a[c, margin:, margin:].ctypes.data for c in range(3)])
        check(lib().qmiv_pic_upload_planar(self.engine.ctx, self.handle, ptrs, stride), "qmiv_pic_upload_planar")
        self.engine.sync()
        return self

    def download_planar(self, margin: int = 0, fill: int = 0xFFFF) -> np.ndarray:
        out = np.full((3, self.height + 2 * margin, self.width + 2 * margin), fill, np.uint16)
        ptrs = (C.c_void_p * 3)(*[out[c, margin:, margin:].ctypes.data for c in range(3)])
        check(lib().qmiv_pic_download_planar(self.engine.ctx, self.handle, ptrs, out.shape[2]), "qmiv_pic_download_planar")
        return out

    def upload_interleaved(self, pels: np.ndarray, margin: int = 0) -> "Picture":
        """The reference's interleaved layout: (H + 2 margin, W + 2 margin, 4) uint16, margin not read."""
        a = np.ascontiguousarray(pels, np.uint16)
        check(lib().qmiv_pic_upload_interleaved(self.engine.ctx, self.handle, a[margin:, margin:].ctypes.data,
                                                a.shape[1]), "qmiv_pic_upload_interleaved")
        self.engine.sync()
        return self

    def download_interleaved(self, margin: int = 0, fill: int = 0xFFFF) -> np.ndarray:
        out = np.full((self.height + 2 * margin, self.width + 2 * margin, 4), fill, np.uint16)
        check(lib().qmiv_pic_download_interleaved(self.engine.ctx, self.handle, out[margin:, margin:].ctypes.data,
                                                  out.shape[1]), "qmiv_pic_download_interleaved")
        return out

    def check_range(self) -> list:
        v = (C.c_int32 * 3)()
        check(lib().qmiv_pic_check(self.engine.ctx, self.handle, v), "qmiv_pic_check")
        return list(v)

    def conceal(self) -> "Picture":
        check(lib().qmiv_pic_conceal(self.engine.ctx, self.handle), "qmiv_pic_conceal")
        self.engine.sync()
        return self

    def equal(self, other: "Picture") -> list:
        v = (C.c_int32 * 3)()
        check(lib().qmiv_pic_equal(self.engine.ctx, self.handle, other.handle, v), "qmiv_pic_equal")
        return list(v)

    def copy_from(self, other: "Picture") -> "Picture":
        check(lib().qmiv_pic_copy(self.engine.ctx, self.handle, other.handle), "qmiv_pic_copy")
        self.engine.sync()
        return self

    def download(self) -> np.ndarray:
        dt = np.uint8 if self.bit_depth <= 8 else np.uint16
        out = np.empty(self.nbytes // np.dtype(dt).itemsize, dt)
        check(lib().qmiv_pic_download(self.engine.ctx, self.handle, out.ctypes.data, out.nbytes), "qmiv_pic_download")
        return out

    @property
    def device_ptr(self) -> int:
        return lib().qmiv_pic_device_ptr(self.handle)

    def close(self):
        if self.handle:
            lib().qmiv_pic_destroy(self.handle)
            self.handle = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sequence:
    """A raw planar YUV file of packed frames, or a list of PNG files (Sequence.png), read through
    qmiv_seq_*: frame count, host reads and reads into device pictures through pinned buffers with
    one-frame prefetch."""

    def __init__(self, path: str, width: int, height: int, bit_depth: int, chroma_format: int = 420):
        h = C.c_void_p()
        check(lib().qmiv_seq_open(os.fsencode(path), width, height, bit_depth, chroma_format, C.byref(h)),
              "qmiv_seq_open")
        self.handle = h
        self.width, self.height, self.bit_depth, self.chroma_format = width, height, bit_depth, chroma_format
        self.num_frames = lib().qmiv_seq_num_frames(h)
        self.frame_bytes = lib().qmiv_seq_frame_bytes(h)

    @classmethod
    def png(cls, pattern: str, width: int, height: int) -> "Sequence":
        """A list of PNG files named by `pattern` ({} / {:0Nd} fields take the file index), read as
        8-bit 4:4:4 R, G, B frames (qmiv_seq_open_png)."""
        h = C.c_void_p()
        check(lib().qmiv_seq_open_png(os.fsencode(pattern), width, height, C.byref(h)), "qmiv_seq_open_png")
        self = cls.__new__(cls)
        self.handle = h
        self.width, self.height, self.bit_depth, self.chroma_format = width, height, 8, 444
        self.num_frames = lib().qmiv_seq_num_frames(h)
        self.frame_bytes = lib().qmiv_seq_frame_bytes(h)
        return self

    def __len__(self) -> int:
        return self.num_frames

    def read_host(self, index: int) -> np.ndarray:
        dt = np.uint8 if self.bit_depth <= 8 else np.uint16
        out = np.empty(self.frame_bytes // np.dtype(dt).itemsize, dt)
        check(lib().qmiv_seq_read_host(self.handle, index, out.ctypes.data, out.nbytes), "qmiv_seq_read_host")
        return out

    def read(self, engine: "Engine", index: int, pic: "Picture") -> "Picture":
        """Frame `index` into `pic`, ordered on the engine's stream (no wait)."""
        check(lib().qmiv_seq_read(engine.ctx, self.handle, index, pic.handle), "qmiv_seq_read")
        return pic

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().qmiv_seq_close(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """A context on one HIP device with its own stream."""

    def __init__(self, device: int = 0):
        ctx = C.c_void_p()
        check(lib().qmiv_create(device, C.byref(ctx)), "qmiv_create")
        self.ctx = ctx
        self._pics = []
        self._batches = []

    @property
    def stream(self) -> int:
        return lib().qmiv_stream(self.ctx) or 0

    def sync(self):
        check(lib().qmiv_sync(self.ctx), "qmiv_sync")

    def profile(self, on: bool = True):
        """Bracket every kernel launch with HIP events on the engine stream."""
        check(lib().qmiv_profile_enable(self.ctx, int(on)), "qmiv_profile_enable")

    def profile_read(self, reset: bool = True) -> dict:
        """{kernel name: (total ms, launches)} accumulated since the last reset."""
        p = capi.Profile()
        check(lib().qmiv_profile_read(self.ctx, C.byref(p), int(reset)), "qmiv_profile_read")
        return {name: (p.kernel_ms[k], p.launches[k]) for k, name in capi.KERNEL_NAMES.items() if p.launches[k]}

    def picture(self, width: int, height: int, bit_depth: int, chroma_format: int = 420,
                data: np.ndarray | None = None) -> Picture:
        h = C.c_void_p()
        check(lib().qmiv_pic_create(self.ctx, width, height, bit_depth, chroma_format, C.byref(h)), "qmiv_pic_create")
        p = Picture(self, h.value, width, height, bit_depth, chroma_format)
        self._pics.append(weakref.ref(p))
        if data is not None:
            p.upload(data)
        return p

    def wrap(self, device_ptr: int, width: int, height: int, bit_depth: int, chroma_format: int = 420,
             keepalive=None) -> Picture:
        h = C.c_void_p()
        check(lib().qmiv_pic_wrap(self.ctx, width, height, bit_depth, chroma_format, C.c_void_p(device_ptr),
                                  C.byref(h)), "qmiv_pic_wrap")
        return Picture(self, h.value, width, height, bit_depth, chroma_format, keepalive)

    def frame(self, p0: Picture, p1: Picture, metrics: int, params: Params | None = None,
              raise_on_range: bool = True) -> FrameResult:
        r = FrameResult()
        st = lib().qmiv_frame(self.ctx, p0.handle, p1.handle, C.byref(params) if params else None, metrics,
                              C.byref(r))
        if st != capi.QMIV_OK and (raise_on_range or st != 5):
            check(st, "qmiv_frame")
        r.status = st
        return r

    def frame_noalloc(self, p0: Picture, p1: Picture, metrics: int, params: Params | None = None) -> FrameResult:
        """qmiv_frame through the allocation-free entry point (the device's fixed scratch block)."""
        r = FrameResult()
        st = lib().qmiv_frame_noalloc(self.ctx, p0.handle, p1.handle, C.byref(params) if params else None, metrics,
                                      C.byref(r))
        check(st, "qmiv_frame_noalloc")
        r.status = st
        return r

    def frames(self, pairs, metrics: int, params: Params | None = None, raise_on_error: bool = True):
        """All (p0, p1) pairs enqueued back to back on the stream, one wait (qmiv_frames)."""
        n = len(pairs)
        a = (C.c_void_p * n)(*[p[0].handle.value for p in pairs])
        b = (C.c_void_p * n)(*[p[1].handle.value for p in pairs])
        out = (FrameResult * n)()
        st = lib().qmiv_frames(self.ctx, n, a, b, C.byref(params) if params else None, metrics, out)
        if raise_on_error:
            check(st, "qmiv_frames")
        return list(out)

    def frames_upload(self, pairs, sources, metrics: int, params: Params | None = None, raise_on_error: bool = True):
        """The (p0, p1) pairs computed as in frames(), each frame first uploaded from its (a, b) packed
        host frames in `sources` on the copy stream, overlapping the frames before it
        (qmiv_frames_upload). Sources from host_array() are copied at the link's rate."""
        n = len(pairs)
        if len(sources) != n:
            raise ValueError("one (a, b) source pair per frame")
        srcs = [(np.ascontiguousarray(a), np.ascontiguousarray(b)) for a, b in sources]
        for (p0, p1), (a, b) in zip(pairs, srcs):
            if a.nbytes != p0.nbytes or b.nbytes != p1.nbytes:
                raise ValueError("source frame size differs from its picture's")
        pa = (C.c_void_p * n)(*[p[0].handle.value for p in pairs])
        pb = (C.c_void_p * n)(*[p[1].handle.value for p in pairs])
        ha = (C.c_void_p * n)(*[a.ctypes.data for a, _ in srcs])
        hb = (C.c_void_p * n)(*[b.ctypes.data for _, b in srcs])
        out = (FrameResult * n)()
        st = lib().qmiv_frames_upload(self.ctx, n, pa, pb, ha, hb, C.byref(params) if params else None, metrics, out)
        if raise_on_error:
            check(st, "qmiv_frames_upload")
        return list(out)

    def host_array(self, nbytes: int) -> np.ndarray:
        """A uint8 array of nbytes in pinned host memory of this engine's device (qmiv_host_alloc),
        freed when the array (and every view of it) is gone."""
        ptr = C.c_void_p()
        check(lib().qmiv_host_alloc(self.ctx, nbytes, C.byref(ptr)), "qmiv_host_alloc")
        buf = (C.c_uint8 * nbytes).from_address(ptr.value)
        arr = np.frombuffer(buf, dtype=np.uint8)
        ctx = self.ctx
        weakref.finalize(buf, lambda: lib().qmiv_host_free(ctx, ptr))
        return arr

    def batch(self, pairs, metrics, params=None, branches: int = 0, cps_sweep: int = -1,
              tune_reps: int = 12) -> "Batch":
        """A captured frame set over the (p0, p1) pairs (qmiv_batch_create). `metrics` and
        `params` are one value for every view or one per view."""
        bt = Batch(self, pairs, metrics, params, branches, cps_sweep, tune_reps)
        self._batches.append(weakref.ref(bt))
        return bt

    def batch_masked(self, triples, metrics: int, params: Params | None = None, branches: int = 0,
                     tune_reps: int = 12) -> "MaskedBatch":
        """A captured masked frame set over the (p0, p1, mask) triples (qmiv_batch_create_masked):
        one metric mask and parameter set for every view, as qmiv_frames_masked."""
        bt = MaskedBatch(self, triples, metrics, params, branches, tune_reps)
        self._batches.append(weakref.ref(bt))
        return bt

    # ---- differentiable SSIM family on float planes in device memory (pointers) ----
    def ssim_backward(self, w, h, planes, tst_ptr, ref_ptr, max_value, calc_l=True, grad_tst_ptr=0, grad_ref_ptr=0):
        v = (C.c_double * planes)()
        check(lib().qmiv_ssim_backward(self.ctx, w, h, planes, C.c_void_p(tst_ptr), C.c_void_p(ref_ptr), max_value,
                                       1 if calc_l else 0, v, C.c_void_p(grad_tst_ptr or None),
                                       C.c_void_p(grad_ref_ptr or None)), "qmiv_ssim_backward")
        return list(v)

    def msssim_backward(self, w, h, planes, tst_ptr, ref_ptr, max_value, grad_tst_ptr=0, grad_ref_ptr=0):
        v = (C.c_double * planes)()
        sc = (C.c_double * (5 * planes))()
        check(lib().qmiv_msssim_backward(self.ctx, w, h, planes, C.c_void_p(tst_ptr), C.c_void_p(ref_ptr), max_value,
                                         v, sc, C.c_void_p(grad_tst_ptr or None), C.c_void_p(grad_ref_ptr or None)),
              "qmiv_msssim_backward")
        return list(v), np.array(list(sc)).reshape(5, planes)

    def ivssim_backward(self, w, h, tst_ptr, ref_ptr, max_value, params: Params | None = None, grad_tst_ptr=0,
                        grad_ref_ptr=0):
        iv, r2t, t2r = C.c_double(), C.c_double(), C.c_double()
        check(lib().qmiv_ivssim_backward(self.ctx, w, h, C.c_void_p(tst_ptr), C.c_void_p(ref_ptr), max_value,
                                         C.byref(params) if params else None, C.byref(iv), C.byref(r2t),
                                         C.byref(t2r), C.c_void_p(grad_tst_ptr or None),
                                         C.c_void_p(grad_ref_ptr or None)), "qmiv_ivssim_backward")
        return {"ivssim": iv.value, "r2t": r2t.value, "t2r": t2r.value}

    def ivssim_backward_batch(self, w, h, tst_ptrs, ref_ptrs, max_value, params: Params | None = None,
                              grad_tst_ptrs=None, grad_ref_ptrs=None):
        """qmiv_ivssim_backward_batch over the pairs (tst_ptrs[i], ref_ptrs[i]); one dict per pair."""
        n = len(tst_ptrs)
        ptrs = lambda v: (C.c_void_p * n)(*[C.c_void_p(p or None) for p in v]) if v is not None else None
        iv, r2t, t2r = (C.c_double * n)(), (C.c_double * n)(), (C.c_double * n)()
        check(lib().qmiv_ivssim_backward_batch(self.ctx, n, w, h, ptrs(tst_ptrs), ptrs(ref_ptrs), max_value,
                                               C.byref(params) if params else None, iv, r2t, t2r,
                                               ptrs(grad_tst_ptrs), ptrs(grad_ref_ptrs)),
              "qmiv_ivssim_backward_batch")
        return [{"ivssim": iv[i], "r2t": r2t[i], "t2r": t2r[i]} for i in range(n)]

    def calc_gcd(self, tst: Picture, ref: Picture, params: Params | None = None) -> np.ndarray:
        g = (C.c_int32 * 4)()
        check(lib().qmiv_calc_gcd(self.ctx, tst.handle, ref.handle, C.byref(params) if params else None, g),
              "qmiv_calc_gcd")
        return np.array(list(g), np.int32)

    def calc_psnr(self, tst: Picture, ref: Picture) -> np.ndarray:
        v = (C.c_double * 3)()
        check(lib().qmiv_calc_psnr(self.ctx, tst.handle, ref.handle, v), "qmiv_calc_psnr")
        return np.array(list(v))

    def calc_wspsnr(self, tst: Picture, ref: Picture, params: Params | None = None) -> np.ndarray:
        v = (C.c_double * 3)()
        check(lib().qmiv_calc_wspsnr(self.ctx, tst.handle, ref.handle, C.byref(params) if params else None, v),
              "qmiv_calc_wspsnr")
        return np.array(list(v))

    def calc_ivpsnr(self, tst: Picture, ref: Picture, gcd, params: Params | None = None, want_rows: bool = False):
        g = (C.c_int32 * 4)(*[int(x) for x in gcd])
        iv, r2t, t2r = C.c_double(), C.c_double(), C.c_double()
        rows = np.zeros(6 * tst.height, np.uint64) if want_rows else None
        check(lib().qmiv_calc_ivpsnr(self.ctx, tst.handle, ref.handle, C.byref(params) if params else None, g,
                                     C.byref(iv), C.byref(r2t), C.byref(t2r),
                                     rows.ctypes.data_as(C.POINTER(C.c_uint64)) if want_rows else None),
              "qmiv_calc_ivpsnr")
        out = {"ivpsnr": iv.value, "r2t": r2t.value, "t2r": t2r.value}
        if want_rows:
            out["rows"] = rows.reshape(2, 3, tst.height)
        return out

    def calc_ssim(self, tst: Picture, ref: Picture, calc_l: bool = True) -> np.ndarray:
        v = (C.c_double * 3)()
        check(lib().qmiv_calc_ssim(self.ctx, tst.handle, ref.handle, int(calc_l), v), "qmiv_calc_ssim")
        return np.array(list(v))

    def calc_msssim(self, tst: Picture, ref: Picture):
        v = (C.c_double * 3)()
        sc = (C.c_double * 15)()
        check(lib().qmiv_calc_msssim(self.ctx, tst.handle, ref.handle, v, sc), "qmiv_calc_msssim")
        return np.array(list(v)), np.array(list(sc)).reshape(5, 3)

    def calc_ivssim(self, tst: Picture, ref: Picture, params: Params | None = None) -> dict:
        iv, r2t, t2r = C.c_double(), C.c_double(), C.c_double()
        check(lib().qmiv_calc_ivssim(self.ctx, tst.handle, ref.handle, C.byref(params) if params else None,
                                     C.byref(iv), C.byref(r2t), C.byref(t2r)), "qmiv_calc_ivssim")
        return {"ivssim": iv.value, "r2t": r2t.value, "t2r": t2r.value}

    def gen_scp(self, src_ref: Picture, src_tst: Picture, gcd, params: Params | None = None):
        """(DstRef, DstTst) as [3, h, w] uint16 arrays (xShftCompPicProc::GenShftCompPics)."""
        n = 3 * src_ref.width * src_ref.height
        dref = np.zeros(n, np.uint16)
        dtst = np.zeros(n, np.uint16)
        g = (C.c_int32 * 4)(*[int(x) for x in gcd])
        check(lib().qmiv_gen_scp(self.ctx, src_ref.handle, src_tst.handle, C.byref(params) if params else None, g,
                                 dref.ctypes.data, dtst.ctypes.data), "qmiv_gen_scp")
        shape = (3, src_ref.height, src_ref.width)
        return dref.reshape(shape), dtst.reshape(shape)

    def gen_scp_pics(self, src_ref: Picture, src_tst: Picture, gcd, dst_ref: Picture, dst_tst: Picture,
                     params: Params | None = None):
        """GenShftCompPics into two 4:4:4 device pictures (DstRef, DstTst)."""
        g = (C.c_int32 * 4)(*[int(x) for x in gcd])
        check(lib().qmiv_gen_scp_pics(self.ctx, src_ref.handle, src_tst.handle, C.byref(params) if params else None,
                                      g, dst_ref.handle, dst_tst.handle), "qmiv_gen_scp_pics")

    def calc_ivssim_scp(self, tst: Picture, ref: Picture, tst_scp: Picture, ref_scp: Picture,
                        params: Params | None = None) -> dict:
        """xIVSSIM::calcPicIVSSIM(Tst, Ref, TstSCP, RefSCP)."""
        iv, r2t, t2r = C.c_double(), C.c_double(), C.c_double()
        check(lib().qmiv_calc_ivssim_scp(self.ctx, tst.handle, ref.handle, tst_scp.handle, ref_scp.handle,
                                         C.byref(params) if params else None, C.byref(iv), C.byref(r2t),
                                         C.byref(t2r)), "qmiv_calc_ivssim_scp")
        return {"ivssim": iv.value, "r2t": r2t.value, "t2r": t2r.value}

    # --- masked mode -----------------------------------------------------------------------------

    def frame_masked(self, p0: Picture, p1: Picture, mask: Picture, metrics: int, params: Params | None = None,
                     raise_on_range: bool = True) -> capi.MaskedResult:
        """The application's masked flow (PSNR / WS-PSNR / IV-PSNR over the mask's non-zero pels)."""
        r = capi.MaskedResult()
        st = lib().qmiv_frame_masked(self.ctx, p0.handle, p1.handle, mask.handle, C.byref(params) if params else None,
                                     metrics, C.byref(r))
        if st != capi.QMIV_OK and (raise_on_range or st != 5):
            check(st, "qmiv_frame_masked")
        r.status = st
        return r

    def frames_masked(self, triples, metrics: int, params: Params | None = None, raise_on_error: bool = True):
        """All (p0, p1, mask) frames enqueued back to back, then each finalised as its results
        arrive (qmiv_frames_masked). Returns the results and the call's status."""
        n = len(triples)
        a = (C.c_void_p * n)(*[t[0].handle.value for t in triples])
        b = (C.c_void_p * n)(*[t[1].handle.value for t in triples])
        m = (C.c_void_p * n)(*[t[2].handle.value for t in triples])
        out = (capi.MaskedResult * n)()
        st = lib().qmiv_frames_masked(self.ctx, n, a, b, m, C.byref(params) if params else None, metrics, out)
        if raise_on_error:
            check(st, "qmiv_frames_masked")
        return list(out), st

    def count_nonzero(self, mask: Picture) -> int:
        n = C.c_int32()
        check(lib().qmiv_count_nonzero(self.ctx, mask.handle, C.byref(n)), "qmiv_count_nonzero")
        return n.value

    def calc_gcd_masked(self, tst: Picture, ref: Picture, mask: Picture, nnm: int = -1,
                        params: Params | None = None) -> np.ndarray:
        g = (C.c_int32 * 4)()
        check(lib().qmiv_calc_gcd_masked(self.ctx, tst.handle, ref.handle, mask.handle,
                                         C.byref(params) if params else None, nnm, g), "qmiv_calc_gcd_masked")
        return np.array(list(g), np.int32)

    def calc_psnr_masked(self, tst: Picture, ref: Picture, mask: Picture, nnm: int = -1,
                         params: Params | None = None) -> np.ndarray:
        v = (C.c_double * 3)()
        check(lib().qmiv_calc_psnr_masked(self.ctx, tst.handle, ref.handle, mask.handle,
                                          C.byref(params) if params else None, nnm, v), "qmiv_calc_psnr_masked")
        return np.array(list(v))

    def calc_wspsnr_masked(self, tst: Picture, ref: Picture, mask: Picture, nnm: int = -1,
                           params: Params | None = None) -> np.ndarray:
        v = (C.c_double * 3)()
        check(lib().qmiv_calc_wspsnr_masked(self.ctx, tst.handle, ref.handle, mask.handle,
                                            C.byref(params) if params else None, nnm, v), "qmiv_calc_wspsnr_masked")
        return np.array(list(v))

    def calc_ivpsnr_masked(self, tst: Picture, ref: Picture, mask: Picture, gcd=None, nnm: int = -1,
                           params: Params | None = None) -> dict:
        g = (C.c_int32 * 4)(*[int(x) for x in gcd]) if gcd is not None else None
        iv, r2t, t2r = C.c_double(), C.c_double(), C.c_double()
        check(lib().qmiv_calc_ivpsnr_masked(self.ctx, tst.handle, ref.handle, mask.handle,
                                            C.byref(params) if params else None, nnm, g, C.byref(iv), C.byref(r2t),
                                            C.byref(t2r)), "qmiv_calc_ivpsnr_masked")
        return {"ivpsnr": iv.value, "r2t": r2t.value, "t2r": t2r.value}

    def upsample(self, src: Picture) -> Picture:
        """A new 4:4:4 picture holding the nearest-neighbour upsampling of `src`."""
        dst = self.picture(src.width, src.height, src.bit_depth, 444)
        check(lib().qmiv_pic_upsample(self.ctx, src.handle, dst.handle), "qmiv_pic_upsample")
        self.sync()
        return dst

    def colorspace(self, src: Picture, space: int, to_rgb: bool, dst: Picture | None = None) -> Picture:
        """RGB -> YCbCr (to_rgb False) or YCbCr -> RGB of a 4:4:4 picture, into `dst` (default: in
        place), in the reference's Q16 arithmetic (xColorSpace::ConvertRGB2YCbCr / ConvertYCbCr2RGB)."""
        dst = src if dst is None else dst
        check(lib().qmiv_pic_colorspace(self.ctx, src.handle, dst.handle, int(space), int(bool(to_rgb))),
              "qmiv_pic_colorspace")
        self.sync()
        return dst

    def swap_planes(self, pic: Picture, a: int, b: int) -> Picture:
        """Exchanges planes a and b of a 4:4:4 picture in place (xPicP::swapComponents)."""
        check(lib().qmiv_pic_swap_planes(self.ctx, pic.handle, int(a), int(b)), "qmiv_pic_swap_planes")
        self.sync()
        return pic

    def close(self):
        for ref in getattr(self, "_batches", []):
            bt = ref()
            if bt is not None:
                bt.close()
        for ref in self._pics:
            p = ref()
            if p is not None:
                p.close()
        self._pics.clear()
        if self.ctx:
            lib().qmiv_destroy(self.ctx)
            self.ctx = C.c_void_p(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Views bound to one captured HIP graph; run() replays it for the frames the pictures hold."""

    def __init__(self, engine: Engine, pairs, metrics, params, branches, cps_sweep, tune_reps):
        n = len(pairs)
        self.engine = engine
        self.n = n
        self._pairs = list(pairs)  # keep the pictures alive as long as the batch
        a = (C.c_void_p * n)(*[p[0].handle.value for p in pairs])
        b = (C.c_void_p * n)(*[p[1].handle.value for p in pairs])
        ms = list(metrics) if hasattr(metrics, "__len__") else [metrics] * n
        m = (C.c_uint32 * n)(*ms)
        if params is None:
            prm = None
        else:
            ps = list(params) if isinstance(params, (list, tuple)) else [params] * n
            prm = (Params * n)(*ps)
        opt = capi.BatchOptions(branches, cps_sweep, tune_reps)
        h = C.c_void_p()
        check(lib().qmiv_batch_create(engine.ctx, n, a, b, prm, m, C.byref(opt), C.byref(h)), "qmiv_batch_create")
        self.handle = h

    def run(self, raise_on_error: bool = True):
        out = (FrameResult * self.n)()
        st = lib().qmiv_batch_run(self.handle, out)
        if raise_on_error:
            check(st, "qmiv_batch_run")
        return list(out)

    def launch(self):
        """Enqueues one replay and returns (qmiv_batch_launch); collect() waits for it."""
        check(lib().qmiv_batch_launch(self.handle), "qmiv_batch_launch")

    def collect(self, raise_on_error: bool = True):
        """Waits for the launched replay and finalises its views (qmiv_batch_collect)."""
        out = (FrameResult * self.n)()
        st = lib().qmiv_batch_collect(self.handle, out)
        if raise_on_error:
            check(st, "qmiv_batch_collect")
        return list(out)

    def info(self) -> dict:
        i = capi.BatchInfo()
        check(lib().qmiv_batch_get_info(self.handle, C.byref(i)), "qmiv_batch_get_info")
        return {k: getattr(i, k) for k, _ in capi.BatchInfo._fields_}

    def close(self):
        if self.handle:
            lib().qmiv_batch_destroy(self.handle)
            self.handle = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MaskedBatch(Batch):
    """Masked views bound to one captured HIP graph; run() / collect() return masked results."""

    def __init__(self, engine: Engine, triples, metrics: int, params, branches, tune_reps):
        n = len(triples)
        self.engine = engine
        self.n = n
        self._pairs = list(triples)  # keep the pictures alive as long as the batch
        a = (C.c_void_p * n)(*[t[0].handle.value for t in triples])
        b = (C.c_void_p * n)(*[t[1].handle.value for t in triples])
        m = (C.c_void_p * n)(*[t[2].handle.value for t in triples])
        opt = capi.BatchOptions(branches, -1, tune_reps)
        h = C.c_void_p()
        check(lib().qmiv_batch_create_masked(engine.ctx, n, a, b, m, C.byref(params) if params else None, metrics,
                                             C.byref(opt), C.byref(h)), "qmiv_batch_create_masked")
        self.handle = h

    def run(self, raise_on_error: bool = True):
        out = (capi.MaskedResult * self.n)()
        st = lib().qmiv_batch_run_masked(self.handle, out)
        if raise_on_error:
            check(st, "qmiv_batch_run_masked")
        return list(out)

    def collect(self, raise_on_error: bool = True):
        """Waits for the launched replay and finalises its views (qmiv_batch_collect_masked)."""
        out = (capi.MaskedResult * self.n)()
        st = lib().qmiv_batch_collect_masked(self.handle, out)
        if raise_on_error:
            check(st, "qmiv_batch_collect_masked")
        return list(out)
