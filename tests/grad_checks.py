"""Checks of the differentiable SSIM family (qmiv_ssim_backward, qmiv_msssim_backward,
qmiv_ivssim_backward), run as `python -m tests.grad_checks` in a process of its own by
tests/test_gpu_grad.py: torch is imported first, so the engine shares torch's HIP runtime there,
while the main GPU test process never loads torch.

Checks:
  * forward on pictures holding k / 1023 equals the integer engine (the reference's values) within
    1e-6 relative (SSIM, IV-SSIM);
  * gradients against an FP64 torch restatement differentiated by autograd (1e-6 relative to the
    gradient's scale) - for IV-SSIM with the winners from a numpy brute-force shift search;
  * gradients against FP64 central finite differences of the engine's own forward (1e-4
    relative), at pels chosen away from zero gradient;
  * the torch.autograd wrappers scale the saved gradients by the incoming ones.
"""
import sys
import traceback

import numpy as np
import torch
import torch.nn.functional as F

from ivpsnr_amd import IVSSIM, SSIM, Engine, default_params
from ivpsnr_amd import torch_ssim

from .helpers import ssim_window_table

MAXV = 1023.0


def field(shape, seed, shift=2):
    """Smooth random pictures in [0, 1] (ref) and a distorted, block-shifted copy (tst)."""
    g = np.random.default_rng(seed)
    c, h, w = shape
    yy, xx = np.mgrid[0:h, 0:w]
    ref = np.zeros(shape)
    for p in range(c):
        for _ in range(5):
            fx, fy, ph = g.uniform(0.5, 3) / w, g.uniform(0.5, 3) / h, g.uniform(0, 2 * np.pi)
            ref[p] += g.uniform(0.05, 0.15) * np.sin(2 * np.pi * (fx * xx + fy * yy) + ph)
    ref = np.clip(ref + 0.5, 0, 1)
    tst = np.roll(ref, (g.integers(-shift, shift + 1), g.integers(-shift, shift + 1)), axis=(1, 2))
    tst = np.clip(tst + g.normal(0, 0.02, shape), 0, 1)
    return tst, ref


def quantised(a):
    return np.rint(a * MAXV) / MAXV


def cuda(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")


WIN = torch.tensor(ssim_window_table(), dtype=torch.float64).reshape(1, 1, 10, 10)


def ssim_ref(t, r, calc_l=True, maxv=MAXV):
    """FP64 SSIM per plane of (C, H, W) tensors in [0, 1] (xStructSim::CalcPel on real samples)."""
    t, r = t[:, None] * maxv, r[:, None] * maxv
    c1, c2 = (0.01 * maxv) ** 2, (0.03 * maxv) ** 2
    h, w = t.shape[-2:]
    mom = lambda x: F.conv2d(x, WIN)[..., : h - 10, : w - 10]
    mt, mr = mom(t), mom(r)
    vt, vr, cov = mom(t * t) - mt * mt, mom(r * r) - mr * mr, mom(t * r) - mt * mr
    s = (2 * cov + c2) / (vt + vr + c2)
    if calc_l:
        s = s * (2 * mt * mr + c1) / (mt * mt + mr * mr + c1)
    return s.sum(dim=(1, 2, 3)) / (w * h)


def msssim_ref(t, r):
    wts = [0.0448, 0.2856, 0.3001, 0.2363, 0.1333]
    out = torch.ones(t.shape[0], dtype=torch.float64)
    for l in range(5):
        out = out * ssim_ref(t, r, calc_l=l not in (0, 4)) ** wts[l]
        if l < 4:
            t = F.avg_pool2d(t[:, None], 2)[:, 0]
            r = F.avg_pool2d(r[:, None], 2)[:, 0]
    return out


def cps_winners(x, y, g, sgn, B=2, ws=(4, 1, 1)):
    """Per-pel first strict minimum of sum_c w_c (x_c + sgn g_c - y_c(p + d))^2 over the window,
    raster order, clamped candidates (xCorrespPixelShiftSTD.cpp:18-75), as a window index."""
    c, h, w = x.shape
    errs = []
    yy, xx = np.mgrid[0:h, 0:w]
    for dy in range(-B, B + 1):
        for dx in range(-B, B + 1):
            qy, qx = np.clip(yy + dy, 0, h - 1), np.clip(xx + dx, 0, w - 1)
            e = sum(ws[k] * (x[k] + sgn * g[k] - y[k][qy, qx]) ** 2 for k in range(3))
            errs.append(e)
    return np.argmin(np.stack(errs), axis=0)


def gcd_of(q0, q1):
    """GCD R2T = clip(rint(sum(P1 - P0) / area), +-rint(0.01 max)) on integer pictures."""
    md = int(np.rint(np.float32(0.01) * np.float32(MAXV)))
    return [int(np.clip(np.rint((q1[c].astype(np.int64) - q0[c]).sum() / q0[c].size), -md, md)) for c in range(3)]


def ivssim_ref(t, r, best0, best1, gcd, B=2):
    """FP64 IV-SSIM of (3, H, W) tensors with fixed winners: T2R = SSIM(P0, SCP1), R2T = SSIM(P1, SCP0)."""
    _, h, w = t.shape
    n = 2 * B + 1
    yy, xx = np.mgrid[0:h, 0:w]

    def scp(src, best, sgn):
        sy = np.clip(yy + best // n - B, 0, h - 1)
        sx = np.clip(xx + best % n - B, 0, w - 1)
        idx = torch.tensor(sy * w + sx).reshape(-1)
        out = []
        for c in range(3):
            v = src[c].reshape(-1)[idx].reshape(h, w) * MAXV - sgn * gcd[c]
            out.append(torch.clamp(v, 0, MAXV) / MAXV)
        return torch.stack(out)

    scp1 = scp(r, best0, 1)
    scp0 = scp(t, best1, -1)
    wts = torch.tensor([4.0, 1.0, 1.0], dtype=torch.float64) / 6.0
    t2r = (ssim_ref(t, scp1) * wts).sum()
    r2t = (ssim_ref(r, scp0) * wts).sum()
    return torch.minimum(t2r, r2t) if bool(r2t < t2r) else t2r, t2r, r2t


def engine_grads(engine, kind, tst, ref, **kw):
    c, h, w = tst.shape
    t, r = cuda(tst), cuda(ref)
    gt, gr = torch.zeros_like(t), torch.zeros_like(r)
    torch.cuda.synchronize()
    if kind == "ssim":
        v = engine.ssim_backward(w, h, c, t.data_ptr(), r.data_ptr(), MAXV, kw.get("calc_l", True), gt.data_ptr(),
                                 gr.data_ptr())
    elif kind == "msssim":
        v, _ = engine.msssim_backward(w, h, c, t.data_ptr(), r.data_ptr(), MAXV, gt.data_ptr(), gr.data_ptr())
    else:
        v = engine.ivssim_backward(w, h, t.data_ptr(), r.data_ptr(), MAXV, kw.get("params"), gt.data_ptr(),
                                   gr.data_ptr())
    return v, gt.double().cpu().numpy(), gr.double().cpu().numpy()


def engine_value(engine, kind, tst, ref, plane=None):
    c, h, w = tst.shape
    t, r = cuda(tst), cuda(ref)
    torch.cuda.synchronize()
    if kind == "ssim":
        return engine.ssim_backward(w, h, c, t.data_ptr(), r.data_ptr(), MAXV)[plane]
    if kind == "msssim":
        return engine.msssim_backward(w, h, c, t.data_ptr(), r.data_ptr(), MAXV)[0][plane]
    return engine.ivssim_backward(w, h, t.data_ptr(), r.data_ptr(), MAXV)["ivssim"]


def test_forward_equals_integer_engine(engine):
    c, h, w = 3, 72, 96
    tst, ref = (quantised(a) for a in field((c, h, w), 1))
    q0, q1 = (np.rint(a * MAXV).astype("<u2") for a in (tst, ref))
    p0 = engine.picture(w, h, 10, 444, data=q0.reshape(-1))
    p1 = engine.picture(w, h, 10, 444, data=q1.reshape(-1))
    r = engine.frame(p0, p1, SSIM | IVSSIM, default_params())
    v, _, _ = engine_grads(engine, "ssim", tst, ref)
    np.testing.assert_allclose(v, list(r.ssim), rtol=1e-6)
    iv, _, _ = engine_grads(engine, "ivssim", tst, ref)
    np.testing.assert_allclose([iv["ivssim"], iv["t2r"], iv["r2t"]], [r.ivssim, r.ivssim_t2r, r.ivssim_r2t], rtol=1e-6)


def test_ssim_gradient_against_fp64_autograd(engine, calc_l):
    c, h, w = 3, 40, 56
    tst, ref = field((c, h, w), 2)
    v, gt, gr = engine_grads(engine, "ssim", tst, ref, calc_l=calc_l)
    t = torch.tensor(tst, dtype=torch.float32).double().requires_grad_()
    r = torch.tensor(ref, dtype=torch.float32).double().requires_grad_()
    vr = ssim_ref(t, r, calc_l)
    np.testing.assert_allclose(v, vr.detach().numpy(), rtol=1e-9)
    vr.sum().backward()
    for got, want in ((gt, t.grad.numpy()), (gr, r.grad.numpy())):
        assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


def test_msssim_gradient_against_fp64_autograd(engine):
    c, h, w = 2, 192, 208
    tst, ref = field((c, h, w), 3)
    v, gt, gr = engine_grads(engine, "msssim", tst, ref)
    t = torch.tensor(tst, dtype=torch.float32).double().requires_grad_()
    r = torch.tensor(ref, dtype=torch.float32).double().requires_grad_()
    vr = msssim_ref(t, r)
    np.testing.assert_allclose(v, vr.detach().numpy(), rtol=1e-6)
    vr.sum().backward()
    for got, want in ((gt, t.grad.numpy()), (gr, r.grad.numpy())):
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


def test_ivssim_gradient_against_fp64_autograd(engine, B=2, h=48, w=64):
    """B = 2 takes the adjoint's LDS-tiled gather (64 x 4 tiles with a 2-pel halo), other search
    ranges its general form. 200 x 72 spans three full tiles and a partial one in x."""
    c = 3
    tst, ref = (quantised(a) for a in field((c, h, w), 4))
    prm = default_params()
    prm.search_range = B
    iv, gt, gr = engine_grads(engine, "ivssim", tst, ref, params=prm)
    q0 = np.rint(np.float32(tst) * MAXV).astype(np.int64)
    q1 = np.rint(np.float32(ref) * MAXV).astype(np.int64)
    gcd = gcd_of(q0, q1)
    best0 = cps_winners(q0, q1, gcd, +1, B)
    best1 = cps_winners(q1, q0, gcd, -1, B)
    t = torch.tensor(tst, dtype=torch.float32).double().requires_grad_()
    r = torch.tensor(ref, dtype=torch.float32).double().requires_grad_()
    vr, t2r, r2t = ivssim_ref(t, r, best0, best1, gcd, B)
    np.testing.assert_allclose([iv["ivssim"], iv["t2r"], iv["r2t"]], [vr.item(), t2r.item(), r2t.item()], rtol=1e-6)
    vr.backward()
    for got, want in ((gt, t.grad.numpy()), (gr, r.grad.numpy())):
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


def test_gradient_against_finite_differences(engine, kind, shape):
    tst, ref = field(shape, 5)
    if kind == "ivssim":
        tst, ref = quantised(tst), quantised(ref)
    tst32 = tst.astype(np.float32)
    v, gt, _ = engine_grads(engine, kind, tst32, ref)
    plane = 0
    rng = np.random.default_rng(9)
    g0 = gt[plane] if kind != "ivssim" else gt.sum(axis=0)
    # pels with a clearly non-zero gradient, away from the borders
    c, h, w = shape
    cand = [(y, x) for y in range(6, h - 6) for x in range(6, w - 6) if abs(g0[y, x]) > 0.2 * np.abs(g0).max()]
    pick = [cand[i] for i in rng.choice(len(cand), size=min(6, len(cand)), replace=False)]
    # IV-SSIM: small enough to keep the quantised pictures; MS-SSIM: its powers curve more
    step = {"ssim": 1e-3, "msssim": 2e-4, "ivssim": 1e-4}[kind]
    for (y, x) in pick:
        p = plane if kind != "ivssim" else int(np.argmax(np.abs(gt[:, y, x])))
        up, dn = tst32.copy(), tst32.copy()
        up[p, y, x] += np.float32(step)
        dn[p, y, x] -= np.float32(step)
        delta = float(up[p, y, x]) - float(dn[p, y, x])
        fd = (engine_value(engine, kind, up, ref, p if kind != "ivssim" else None)
              - engine_value(engine, kind, dn, ref, p if kind != "ivssim" else None)) / delta
        assert abs(fd - gt[p, y, x]) <= 1e-4 * abs(gt[p, y, x]), (kind, y, x, fd, gt[p, y, x])


def test_autograd_wrappers(engine):
    a, b = field((3, 40, 48), 7)
    c, d = field((3, 40, 48), 8)
    t = cuda(np.stack([a, c])).requires_grad_()
    r = cuda(np.stack([b, d]))
    wts = torch.tensor([[0.5, 1.0, 2.0], [1.5, -1.0, 0.25]], dtype=torch.float64, device="cuda")
    v = torch_ssim.ssim(t, r)
    (v * wts).sum().backward()
    for i, (x, y) in enumerate(((a, b), (c, d))):
        _, gt, _ = engine_grads(engine, "ssim", x, y)
        want = gt * wts[i].cpu().numpy()[:, None, None]
        np.testing.assert_allclose(t.grad[i].double().cpu().numpy(), want, rtol=1e-6, atol=1e-12)
    t.grad = None
    iv = torch_ssim.ivssim(t, r)
    assert iv.shape == (2, 1)
    iv.sum().backward()
    assert torch.isfinite(t.grad).all() and t.grad.abs().sum() > 0


# ---- the differentiable workload at its size: 8 x 3 x 1024 x 1024 fp32 -------------------------

def mom_fp64(x, win):
    """Windowed sums of (C, H, W) fp64 x over the interior: out[:, i, j] = sum_{a,b < 10}
    win[a][b] x[:, i + a, j + b], i < H - 10, j < W - 10 (dy-outer, dx-inner; plain tensor ops, so
    autograd differentiates them on the device without a convolution library)."""
    h, w = x.shape[-2:]
    acc = None
    for a in range(10):
        for b in range(10):
            term = win[a][b] * x[:, a:a + h - 10, b:b + w - 10]
            acc = term if acc is None else acc + term
    return acc


WIN_ROWS = [[float(v) for v in row] for row in ssim_window_table()]


def ssim_dev(t, r, calc_l=True, maxv=MAXV):
    """FP64 SSIM per plane of (C, H, W) device tensors in [0, 1] (xStructSim::CalcPel on real samples,
    interior pels, over W H)."""
    t, r = t * maxv, r * maxv
    c1, c2 = (0.01 * maxv) ** 2, (0.03 * maxv) ** 2
    h, w = t.shape[-2:]
    mt, mr = mom_fp64(t, WIN_ROWS), mom_fp64(r, WIN_ROWS)
    vt = mom_fp64(t * t, WIN_ROWS) - mt * mt
    vr = mom_fp64(r * r, WIN_ROWS) - mr * mr
    cov = mom_fp64(t * r, WIN_ROWS) - mt * mr
    s = (2 * cov + c2) / (vt + vr + c2)
    if calc_l:
        s = s * (2 * mt * mr + c1) / (mt * mt + mr * mr + c1)
    return s.sum(dim=(1, 2)) / (w * h)


def msssim_dev(t, r):
    """MS-SSIM per plane with the float 2 x 2 mean pyramid (odd rows / columns dropped)."""
    wts = [0.0448, 0.2856, 0.3001, 0.2363, 0.1333]
    out = torch.ones(t.shape[0], dtype=torch.float64, device=t.device)
    for lev in range(5):
        out = out * ssim_dev(t, r, calc_l=lev not in (0, 4)) ** wts[lev]
        if lev < 4:
            c, h, w = t.shape
            t = t[:, : h // 2 * 2, : w // 2 * 2].reshape(c, h // 2, 2, w // 2, 2).mean(dim=(2, 4))
            r = r[:, : h // 2 * 2, : w // 2 * 2].reshape(c, h // 2, 2, w // 2, 2).mean(dim=(2, 4))
    return out


def winners_dev(x, y, g, sgn, B=2, ws=(4, 1, 1)):
    """First strict minimum (raster order, clamped candidates) of sum_c w_c (x_c + sgn g_c -
    y_c(p + d))^2 over the window, by brute force on integer (C, H, W) device tensors (int64): the
    key err * 32 + index makes the first minimum the unique one."""
    _, h, w = x.shape
    yy = torch.arange(h, device=x.device)[:, None]
    xx = torch.arange(w, device=x.device)[None, :]
    keys = []
    n = 2 * B + 1
    for dy in range(-B, B + 1):
        for dx in range(-B, B + 1):
            qy, qx = (yy + dy).clamp(0, h - 1), (xx + dx).clamp(0, w - 1)
            e = sum(ws[k] * (x[k] + sgn * g[k] - y[k][qy, qx]) ** 2 for k in range(3))
            keys.append(e * 32 + (dy + B) * n + (dx + B))
    return torch.argmin(torch.stack(keys), dim=0)


def ivssim_dev(t, r, best0, best1, gcd, B=2):
    """FP64 IV-SSIM with fixed winners: T2R = SSIM(P0, SCP1), R2T = SSIM(P1, SCP0), 4:1:1, min."""
    _, h, w = t.shape
    n = 2 * B + 1
    yy = torch.arange(h, device=t.device)[:, None]
    xx = torch.arange(w, device=t.device)[None, :]

    def scp(src, best, sgn):
        sy = (yy + torch.div(best, n, rounding_mode="floor") - B).clamp(0, h - 1)
        sx = (xx + best % n - B).clamp(0, w - 1)
        return torch.stack([torch.clamp(src[c][sy, sx] * MAXV - sgn * gcd[c], 0, MAXV) / MAXV for c in range(3)])

    wts = torch.tensor([4.0, 1.0, 1.0], dtype=torch.float64, device=t.device) / 6.0
    t2r = (ssim_dev(t, scp(r, best0, 1)) * wts).sum()
    r2t = (ssim_dev(r, scp(t, best1, -1)) * wts).sum()
    return (r2t if bool(r2t < t2r) else t2r), t2r, r2t


def grad_close(got, want, rel):
    return float((got - want).abs().max()) <= rel * float(want.abs().max())


def test_workload_batch_1024(engine):
    """The differentiable workload at its size, through the torch.autograd wrappers: IV-SSIM and
    MS-SSIM of the 8 x 3 x 1024 x 1024 batch, forward and backward.
      * forward on the quantised batch equals the integer engine (the reference's values) within
        1e-6 (SSIM per plane, IV-SSIM with both directions);
      * gradients of every picture against the FP64 device restatement, differentiated by autograd
        (IV-SSIM with brute-force winners), within 1e-5 of the gradient's scale;
      * gradients at pels of pictures 0 and 7 against FP64 central finite differences of the
        engine's forward, within 1e-4."""
    from ivpsnr_amd.synth import grad_batch
    n, c, h, w = 8, 3, 1024, 1024
    tst, ref = grad_batch(n, c, h, w, seed=5)
    tst = torch.round(tst * MAXV) / MAXV   # the shift search sees exactly these samples
    ref = torch.round(ref * MAXV) / MAXV
    t = tst.clone().requires_grad_()
    iv = torch_ssim.ivssim(t, ref)
    assert iv.shape == (n, 1)
    iv.sum().backward()
    g_iv = t.grad.detach().clone()
    t.grad = None
    ms = torch_ssim.msssim(t, ref)
    assert ms.shape == (n, c)
    ms.sum().backward()
    g_ms = t.grad.detach().clone()
    # forward against the integer engine
    q0 = torch.round(tst * MAXV).to(torch.int32)
    q1 = torch.round(ref * MAXV).to(torch.int32)
    ssim_float = torch_ssim.ssim(tst, ref)
    for i in range(n):
        a = q0[i].cpu().numpy().astype("<u2").reshape(-1)
        b = q1[i].cpu().numpy().astype("<u2").reshape(-1)
        p0 = engine.picture(w, h, 10, 444, data=a)
        p1 = engine.picture(w, h, 10, 444, data=b)
        r = engine.frame(p0, p1, SSIM | IVSSIM, default_params())
        p0.close()
        p1.close()
        np.testing.assert_allclose(ssim_float[i].cpu().numpy(), list(r.ssim), rtol=1e-6)
        np.testing.assert_allclose(iv[i, 0].item(), r.ivssim, rtol=1e-6)
    # gradients against the FP64 restatement, picture by picture
    for i in range(n):
        x0 = q0[i].to(torch.int64)
        x1 = q1[i].to(torch.int64)
        md = int(np.rint(np.float32(0.01) * np.float32(MAXV)))
        gcd = [int(np.clip(np.rint(float((x1[k] - x0[k]).sum()) / (h * w)), -md, md)) for k in range(3)]
        best0 = winners_dev(x0, x1, gcd, +1)
        best1 = winners_dev(x1, x0, gcd, -1)
        tt = tst[i].double().requires_grad_()
        vr, t2r, r2t = ivssim_dev(tt, ref[i].double(), best0, best1, gcd)
        np.testing.assert_allclose(iv[i, 0].item(), vr.item(), rtol=1e-6)
        vr.backward()
        assert grad_close(g_iv[i].double(), tt.grad, 1e-5), ("ivssim", i)
        tt = tst[i].double().requires_grad_()
        vm = msssim_dev(tt, ref[i].double())
        np.testing.assert_allclose(ms[i].detach().cpu().numpy(), vm.detach().cpu().numpy(), rtol=1e-6)
        vm.sum().backward()
        assert grad_close(g_ms[i].double(), tt.grad, 1e-5), ("msssim", i)
        print(f"  picture {i}: IV-SSIM {vr.item():.9f}, MS-SSIM {vm.detach().cpu().numpy()}", flush=True)
    # central finite differences of the engine's forward at pels of pictures 0 and 7
    rng = np.random.default_rng(21)
    for i in (0, n - 1):
        for kind, grad, step in (("ivssim", g_iv[i], 1e-4), ("msssim", g_ms[i], 2e-4)):
            gi = grad.double().cpu().numpy()
            mag = np.abs(gi).max(axis=0) if kind == "ivssim" else np.abs(gi[0])
            ys, xs = np.nonzero(mag[8:h - 8, 8:w - 8] > 0.3 * mag.max())
            pick = rng.choice(len(ys), size=min(4, len(ys)), replace=False)
            base = tst[i].cpu().numpy().astype(np.float32)
            rf = ref[i].cpu().numpy()
            for k in pick:
                y, x = int(ys[k]) + 8, int(xs[k]) + 8
                p = int(np.argmax(np.abs(gi[:, y, x]))) if kind == "ivssim" else 0
                up, dn = base.copy(), base.copy()
                up[p, y, x] += np.float32(step)
                dn[p, y, x] -= np.float32(step)
                delta = float(up[p, y, x]) - float(dn[p, y, x])
                pl = None if kind == "ivssim" else p
                fd = (engine_value(engine, kind, up, rf, pl) - engine_value(engine, kind, dn, rf, pl)) / delta
                assert abs(fd - gi[p, y, x]) <= 1e-4 * abs(gi[p, y, x]), (kind, i, y, x, fd, gi[p, y, x])


def test_ivssim_batch_equals_single(engine):
    """qmiv_ivssim_backward_batch: every pair's values and gradients equal qmiv_ivssim_backward on it
    alone, bit for bit, with more pairs than streams, one pair of identical pictures (tied
    directions), and pairs asking for one gradient or none."""
    h, w = 72, 200
    pairs = [field((3, h, w), 300 + i) for i in range(5)]
    pairs[2] = (pairs[2][1].copy(), pairs[2][1].copy())
    tst = [cuda(t) for t, _ in pairs]
    ref = [cuda(r) for _, r in pairs]
    want_t = [True, True, False, True, False]
    want_r = [True, False, True, True, False]
    gt = [torch.zeros_like(t) for t in tst]
    gr = [torch.zeros_like(r) for r in ref]
    torch.cuda.synchronize()
    got = engine.ivssim_backward_batch(w, h, [t.data_ptr() for t in tst], [r.data_ptr() for r in ref], MAXV, None,
                                       [g.data_ptr() if k else 0 for g, k in zip(gt, want_t)],
                                       [g.data_ptr() if k else 0 for g, k in zip(gr, want_r)])
    for i in range(len(pairs)):
        st, sr = torch.zeros_like(tst[i]), torch.zeros_like(ref[i])
        one = engine.ivssim_backward(w, h, tst[i].data_ptr(), ref[i].data_ptr(), MAXV, None,
                                     st.data_ptr() if want_t[i] else 0, sr.data_ptr() if want_r[i] else 0)
        torch.cuda.synchronize()
        assert got[i] == one, (i, got[i], one)
        if want_t[i]:
            assert torch.equal(gt[i], st), ("grad_tst", i)
        if want_r[i]:
            assert torch.equal(gr[i], sr), ("grad_ref", i)
        print(f"  pair {i}: IV-SSIM {got[i]['ivssim']:.9f}", flush=True)


CHECKS = [
    ("forward_equals_integer_engine", test_forward_equals_integer_engine, {}),
    ("ssim_gradient_fp64_autograd[L]", test_ssim_gradient_against_fp64_autograd, {"calc_l": True}),
    ("ssim_gradient_fp64_autograd[CS]", test_ssim_gradient_against_fp64_autograd, {"calc_l": False}),
    ("msssim_gradient_fp64_autograd", test_msssim_gradient_against_fp64_autograd, {}),
    ("ivssim_gradient_fp64_autograd", test_ivssim_gradient_against_fp64_autograd, {}),
    ("ivssim_gradient_fp64_autograd[200x72]", test_ivssim_gradient_against_fp64_autograd, {"h": 72, "w": 200}),
    ("ivssim_gradient_fp64_autograd[B=1]", test_ivssim_gradient_against_fp64_autograd, {"B": 1}),
    ("ivssim_gradient_fp64_autograd[B=3]", test_ivssim_gradient_against_fp64_autograd, {"B": 3}),
    ("finite_differences[ssim]", test_gradient_against_finite_differences, {"kind": "ssim", "shape": (3, 40, 48)}),
    ("finite_differences[msssim]", test_gradient_against_finite_differences,
     {"kind": "msssim", "shape": (1, 180, 184)}),
    ("finite_differences[ivssim]", test_gradient_against_finite_differences, {"kind": "ivssim", "shape": (3, 40, 48)}),
    ("ivssim_batch_equals_single", test_ivssim_batch_equals_single, {}),
    ("autograd_wrappers", test_autograd_wrappers, {}),
    ("workload_batch_8x3x1024x1024", test_workload_batch_1024, {}),
]


def main() -> int:
    engine = Engine(0)
    failed = 0
    for name, fn, kw in CHECKS:
        print(f"[check] {name}", flush=True)
        try:
            fn(engine, **kw)
            print(f"PASS {name}", flush=True)
        except Exception:  # report every check, then fail the process
            failed += 1
            print(f"FAIL {name}\n{traceback.format_exc()}", flush=True)
    engine.close()
    print(f"{len(CHECKS) - failed} passed, {failed} failed", flush=True)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
