"""Deterministic synthetic YUV frame pairs for tests and benchmarks.

A reference picture is a sum of eight low-frequency 2-D sinusoids plus uniform
noise around mid-grey; the tested picture is the reference with a per-block
integer shift (clamp-to-edge), a luma offset and Gaussian noise. Chroma planes
use the same generator at 4:2:0 resolution with independent seeds, blocks of
half the size and half the shift.

Determinism: the 1-D sinusoid vectors come from ``math.sin``/``math.cos``
(the C library), the 2-D field from separate numpy multiplies and adds (no
fused operations) and the noise from numpy's PCG64 generator, so the same
seed gives the same bytes on every host running this image. Callers that pin
results to stored fixtures also record the SHA-256 of the generated bytes.

Frames are returned in the raw planar layout of a ``yuv420p``/``yuv420p10le``
file (Y plane, then Cb, then Cr; one byte per sample at 8 bits, little-endian
uint16 above), which is the layout the device engine consumes directly.
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass

import numpy as np

__all__ = ["FramePair", "make_pair", "make_mask", "file_frame", "inject_out_of_range", "grad_batch", "config_pair",
           "sha256", "CONFIGS"]


@dataclass
class FramePair:
    width: int
    height: int
    bit_depth: int
    chroma_format: int
    ref: np.ndarray  # packed frame, uint8 or uint16, 1-D
    tst: np.ndarray

    def planes(self, which: str):
        """Return (Y, Cb, Cr) 2-D views of the packed frame ``which`` ('ref' or 'tst')."""
        return split_planes(getattr(self, which), self.width, self.height, self.chroma_format)


def split_planes(packed: np.ndarray, w: int, h: int, chroma_format: int):
    cw, ch = (w // 2, h // 2) if chroma_format == 420 else (w, h)
    y = packed[: w * h].reshape(h, w)
    cb = packed[w * h : w * h + cw * ch].reshape(ch, cw)
    cr = packed[w * h + cw * ch : w * h + 2 * cw * ch].reshape(ch, cw)
    return y, cb, cr


def sha256(arr: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def _field(w: int, h: int, rng: np.random.Generator, scale: float, max_value: int) -> np.ndarray:
    """Sum of 8 sinusoids (amplitude 20-60, 0.5-4 cycles per dimension) + U(-8, 8), centred at mid-grey."""
    acc = np.zeros((h, w), dtype=np.float64)
    xs = range(w)
    ys = range(h)
    for _ in range(8):
        amp = rng.uniform(20.0, 60.0)
        fx = rng.uniform(0.5, 4.0) * (1.0 if rng.uniform() < 0.5 else -1.0)
        fy = rng.uniform(0.5, 4.0)
        phi = rng.uniform(0.0, 2.0 * math.pi)
        # sin(a + b) = sin(a)cos(b) + cos(a)sin(b), a over x (with phase), b over y
        sa = np.array([math.sin(2.0 * math.pi * fx * x / w + phi) for x in xs], dtype=np.float64)
        ca = np.array([math.cos(2.0 * math.pi * fx * x / w + phi) for x in xs], dtype=np.float64)
        sb = np.array([math.sin(2.0 * math.pi * fy * y / h) for y in ys], dtype=np.float64)
        cb = np.array([math.cos(2.0 * math.pi * fy * y / h) for y in ys], dtype=np.float64)
        t1 = np.multiply(cb[:, None], sa[None, :])
        t2 = np.multiply(sb[:, None], ca[None, :])
        acc = np.add(acc, np.multiply(amp, np.add(t1, t2)))
    noise = rng.uniform(-8.0, 8.0, size=(h, w))
    acc = np.add(acc, noise)
    mid = 128.0
    val = np.add(np.multiply(acc, scale), mid * scale)
    return np.clip(np.rint(val), 0, max_value)


def _shift_blocks(plane: np.ndarray, block: int, shifts: np.ndarray, halve: bool) -> np.ndarray:
    """Per-block integer shift with clamp-to-edge addressing: out(y, x) = in(clamp(y+dy), clamp(x+dx))."""
    h, w = plane.shape
    sh = shifts >> 1 if halve else shifts
    by = np.arange(h) // block
    bx = np.arange(w) // block
    dy = sh[by[:, None], bx[None, :], 0]
    dx = sh[by[:, None], bx[None, :], 1]
    ys = np.clip(np.arange(h)[:, None] + dy, 0, h - 1)
    xs = np.clip(np.arange(w)[None, :] + dx, 0, w - 1)
    return plane[ys, xs]


def make_pair(width: int, height: int, bit_depth: int, seed: int, *, chroma_format: int = 420,
              luma_offset: float = 3.0, noise_sigma: float = 2.0, max_shift: int = 2,
              block: int = 32) -> FramePair:
    """Generate one (reference, tested) frame pair.

    Amplitudes, offsets and noise are given in 8-bit units and scaled by
    2**(bit_depth-8) before rounding, so a 10-bit pair is the 8-bit field x4.
    """
    if chroma_format not in (420, 444):
        raise ValueError("chroma_format must be 420 or 444")
    if chroma_format == 420 and (width % 2 or height % 2):
        raise ValueError("4:2:0 needs even width and height")
    scale = float(1 << (bit_depth - 8))
    max_value = (1 << bit_depth) - 1
    rng = np.random.default_rng(np.random.SeedSequence([int(seed), width, height, bit_depth]))
    cw, ch = (width // 2, height // 2) if chroma_format == 420 else (width, height)

    ref_y = _field(width, height, rng, scale, max_value)
    ref_cb = _field(cw, ch, rng, scale, max_value)
    ref_cr = _field(cw, ch, rng, scale, max_value)

    nby = (height + block - 1) // block
    nbx = (width + block - 1) // block
    shifts = rng.integers(-max_shift, max_shift + 1, size=(nby, nbx, 2))
    cblock = block // 2 if chroma_format == 420 else block
    halve = chroma_format == 420

    def distort(plane, blk, hv, offset):
        shifted = _shift_blocks(plane, blk, shifts, hv)
        noise = rng.normal(0.0, noise_sigma * scale, size=plane.shape)
        return np.clip(np.rint(np.add(np.add(shifted, offset * scale), noise)), 0, max_value)

    tst_y = distort(ref_y, block, False, luma_offset)
    tst_cb = distort(ref_cb, cblock, halve, 0.0)
    tst_cr = distort(ref_cr, cblock, halve, 0.0)

    dt = np.uint8 if bit_depth <= 8 else np.uint16
    ref = np.concatenate([p.astype(dt).ravel() for p in (ref_y, ref_cb, ref_cr)])
    tst = np.concatenate([p.astype(dt).ravel() for p in (tst_y, tst_cb, tst_cr)])
    return FramePair(width, height, bit_depth, chroma_format, ref, tst)


def make_mask(width: int, height: int, bit_depth: int, seed: int, *, chroma_format: int = 420,
              kind: str = "objects", objects: int = 4) -> np.ndarray:
    """One packed mask frame whose luma plane selects the pels of masked mode.

    kind "objects": a few random ellipses at 2^bd - 1 on zero (the binary object masks of masked
    mode); "soft": the same ellipses holding random values in [1, 2^bd - 1]; "speckle": every pel
    independently 0 or a random value in [1, 2^bd - 1], half and half (most search windows mix
    masked and unmasked candidates); "zero" / "full": every pel 0 / 2^bd - 1. Chroma planes hold
    the luma plane's top-left samples.
    """
    if chroma_format == 420 and (width % 2 or height % 2):
        raise ValueError("4:2:0 needs even width and height")
    max_value = (1 << bit_depth) - 1
    rng = np.random.default_rng(np.random.SeedSequence([int(seed), width, height, bit_depth, 77]))
    if kind == "zero":
        y = np.zeros((height, width), np.int64)
    elif kind == "full":
        y = np.full((height, width), max_value, np.int64)
    elif kind in ("objects", "soft"):
        yy, xx = np.mgrid[0:height, 0:width]
        inside = np.zeros((height, width), bool)
        for _ in range(objects):
            cy, cx = rng.uniform(0, height), rng.uniform(0, width)
            ry, rx = rng.uniform(0.08, 0.3) * height, rng.uniform(0.05, 0.2) * width
            inside |= ((yy - cy) / ry) ** 2 + ((xx - cx) / rx) ** 2 <= 1.0
        if kind == "objects":
            y = np.where(inside, max_value, 0)
        else:
            y = np.where(inside, rng.integers(1, max_value + 1, size=(height, width)), 0)
    elif kind == "speckle":
        y = np.where(rng.random((height, width)) < 0.5, 0, rng.integers(1, max_value + 1, size=(height, width)))
    else:
        raise ValueError(f"unknown mask kind {kind!r}")
    c = y[::2, ::2] if chroma_format == 420 else y
    dt = np.uint8 if bit_depth <= 8 else np.uint16
    return np.concatenate([y.astype(dt).ravel(), c.astype(dt).ravel(), c.astype(dt).ravel()])


def file_frame(planes: np.ndarray, bit_depth: int, chroma_format: int) -> np.ndarray:
    """One raw-file frame (bytes, uint8) of chroma format 400, 420, 422 or 444 from 4:4:4 planes
    (3, h, w), in the layout the reference's reader takes (libs/pmbb-core/src/xSeq.cpp:179-249,
    408-428): luma, then each chroma plane as (luma bytes >> 2) bytes for 4:2:0 and (luma bytes >> 1)
    for 4:2:2, holding rows of w >> 1 samples taken at even columns (and even rows for 4:2:0),
    little-endian above 8 bits; the frame is xSeq::create's size. Odd sizes keep the reference's row
    stride of w >> 1; bytes left over at a plane's or the frame's end are 0."""
    _, h, w = planes.shape
    bps = 1 if bit_depth <= 8 else 2
    dt = np.dtype("u1") if bps == 1 else np.dtype("<u2")
    luma = np.ascontiguousarray(planes[0].astype(dt)).tobytes()
    if chroma_format == 400:
        return np.frombuffer(luma, np.uint8).copy()
    if chroma_format == 444:
        return np.frombuffer(b"".join(np.ascontiguousarray(p.astype(dt)).tobytes() for p in planes), np.uint8).copy()
    cmp = w * h * bps
    nbytes = cmp >> 2 if chroma_format == 420 else cmp >> 1
    k = w >> 1
    out = [luma]
    for c in (1, 2):
        sub = planes[c][::2, 0:2 * k:2] if chroma_format == 420 else planes[c][:, 0:2 * k:2]
        raw = np.ascontiguousarray(sub.astype(dt)).tobytes()[:nbytes]
        out.append(raw + bytes(nbytes - len(raw)))
    # The frame is xSeq::create's size: luma + (luma >> 1) bytes for 4:2:0, twice the luma for 4:2:2
    # (one byte more than the two planes when the luma byte count is odd).
    total = cmp + (cmp >> 1) if chroma_format == 420 else 2 * cmp
    data = b"".join(out)
    return np.frombuffer(data + bytes(total - len(data)), np.uint8).copy()


def inject_out_of_range(packed: np.ndarray, bit_depth: int, seed: int, count: int = 200,
                        top: int | None = None) -> np.ndarray:
    """A copy of a packed 16-bit frame with `count` samples (anywhere in its planes) set to values
    in (2^bd - 1, top]: pels above the bit depth, as a broken file holds them. top (default
    max(4095, 2^(bd + 1) - 1)) stays below 2^14 so that every difference fits the reference's
    16-bit SIMD lanes."""
    if top is None:
        top = max(4095, (2 << bit_depth) - 1)
    top = min(top, (1 << 14) - 1)
    rng = np.random.default_rng(np.random.SeedSequence([int(seed), 991]))
    out = packed.copy()
    idx = rng.choice(out.size, size=count, replace=False)
    out[idx] = rng.integers((1 << bit_depth), top + 1, size=count).astype(out.dtype)
    return out


def grad_batch(n: int = 8, planes: int = 3, height: int = 1024, width: int = 1024, seed: int = 5):
    """The differentiable workload's float pictures: (tst, ref), each an (n, planes, h, w) float32
    CUDA tensor in [0, 1]. ref is a smooth random field (six sinusoids per plane around 0.5); tst is
    ref shifted by up to 2 pels per 32 x 32 block (clamp-to-edge) with N(0, 0.02) noise. Drawn with
    torch's CUDA generator (imported here; the caller imports torch before the engine)."""
    import torch
    dev = "cuda"
    g = torch.Generator(device=dev).manual_seed(seed)
    yy = torch.arange(height, device=dev, dtype=torch.float32)[:, None] / height
    xx = torch.arange(width, device=dev, dtype=torch.float32)[None, :] / width
    ref = torch.zeros(n, planes, height, width, device=dev)
    for _ in range(6):
        a = torch.rand(n, planes, 1, 1, device=dev, generator=g) * 0.1 + 0.05
        fx = torch.rand(n, planes, 1, 1, device=dev, generator=g) * 3.5 + 0.5
        fy = torch.rand(n, planes, 1, 1, device=dev, generator=g) * 3.5 + 0.5
        ph = torch.rand(n, planes, 1, 1, device=dev, generator=g) * 6.283
        ref += a * torch.sin(6.283 * (fx * xx + fy * yy) + ph)
    ref = (ref + 0.5).clamp(0, 1)
    nby, nbx = (height + 31) // 32, (width + 31) // 32
    sh = torch.randint(-2, 3, (2, nby, nbx), device=dev, generator=g)
    dy = sh[0].repeat_interleave(32, 0).repeat_interleave(32, 1)[:height, :width]
    dx = sh[1].repeat_interleave(32, 0).repeat_interleave(32, 1)[:height, :width]
    iy = (torch.arange(height, device=dev)[:, None] + dy).clamp(0, height - 1)
    ix = (torch.arange(width, device=dev)[None, :] + dx).clamp(0, width - 1)
    tst = (ref[:, :, iy, ix] + 0.02 * torch.randn(n, planes, height, width, device=dev, generator=g)).clamp(0, 1)
    return tst.contiguous(), ref.contiguous()


# Workload presets (frame size, bit depth, frames, generator knobs).
CONFIGS = {
    # plain IV-PSNR + PSNR, 8-bit 1080p
    "ivpsnr_1080p8": dict(width=1920, height=1080, bit_depth=8, frames=3, luma_offset=3.0),
    # IV-PSNR + WS-PSNR on ERP, 10-bit 4096x2048 (luma offset x4 = +12, beyond the 10-level clip)
    "ivpsnr_erp_4k10": dict(width=4096, height=2048, bit_depth=10, frames=8, luma_offset=3.0),
    # SSIM / MS-SSIM / IV-SSIM, 10-bit 1080p
    "ssim_1080p10": dict(width=1920, height=1080, bit_depth=10, frames=16, luma_offset=3.0),
}


def config_pair(name: str, frame: int, **overrides) -> FramePair:
    c = dict(CONFIGS[name])
    c.update(overrides)
    return make_pair(c["width"], c["height"], c["bit_depth"], seed=1 + frame,
                     luma_offset=c.get("luma_offset", 3.0))
