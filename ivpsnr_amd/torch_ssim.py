"""torch.autograd wrappers of the engine's differentiable SSIM family (qmiv_ssim_backward,
qmiv_msssim_backward, qmiv_ivssim_backward).

Inputs are float32 CUDA tensors of shape (planes, H, W) or (N, planes, H, W) with values in
[0, 1] standing for samples x * max_value. The engine computes the forward values and the
gradients in one pass (fp64 inside); backward scales the saved gradients by the incoming ones.
Values: per plane for SSIM / MS-SSIM, one scalar per picture for IV-SSIM.

One HIP runtime serves a process: PyTorch wheels carry their own libamdhip64 under the soname the
engine library asks for, so whichever loads first serves both. This module imports torch before
the engine library is loaded, so the engine runs on torch's runtime here.
"""
from __future__ import annotations

import torch

from .engine import Engine, default_params

_engines: dict[int, Engine] = {}


def _engine(device: torch.device) -> Engine:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx not in _engines:
        _engines[idx] = Engine(idx)
    return _engines[idx]


def _prep(t: torch.Tensor) -> torch.Tensor:
    if not t.is_cuda or t.dtype != torch.float32:
        raise ValueError("expected float32 CUDA tensors")
    return t.contiguous()


def _batched(x):
    return x if x.dim() == 4 else x.unsqueeze(0)


class _SSIMFamily(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tst, ref, kind, max_value, calc_l, params):
        tst, ref = _prep(tst), _prep(ref)
        squeeze = tst.dim() == 3
        tb, rb = _batched(tst), _batched(ref)
        if tb.shape != rb.shape:
            raise ValueError("tst and ref differ in shape")
        n, planes, h, w = tb.shape
        need = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        gt = torch.empty_like(tb) if need else None
        gr = torch.empty_like(rb) if need else None
        eng = _engine(tb.device)
        torch.cuda.current_stream(tb.device).synchronize()  # the engine runs on its own stream
        vals = []
        if kind == "ivssim" and n > 1:
            # the pictures of the batch together (qmiv_ivssim_backward_batch)
            if planes != 3:
                raise ValueError("IV-SSIM needs 3 planes")
            res = eng.ivssim_backward_batch(w, h, [tb[i].data_ptr() for i in range(n)],
                                            [rb[i].data_ptr() for i in range(n)], max_value, params,
                                            [gt[i].data_ptr() for i in range(n)] if need else None,
                                            [gr[i].data_ptr() for i in range(n)] if need else None)
            vals = [[r["ivssim"]] for r in res]
        for i in range(n if not vals else 0):
            ptr_gt = gt[i].data_ptr() if need else 0
            ptr_gr = gr[i].data_ptr() if need else 0
            if kind == "ssim":
                v = eng.ssim_backward(w, h, planes, tb[i].data_ptr(), rb[i].data_ptr(), max_value, calc_l, ptr_gt, ptr_gr)
            elif kind == "msssim":
                v, _ = eng.msssim_backward(w, h, planes, tb[i].data_ptr(), rb[i].data_ptr(), max_value, ptr_gt, ptr_gr)
            else:
                if planes != 3:
                    raise ValueError("IV-SSIM needs 3 planes")
                v = [eng.ivssim_backward(w, h, tb[i].data_ptr(), rb[i].data_ptr(), max_value, params, ptr_gt,
                                         ptr_gr)["ivssim"]]
            vals.append(v)
        out = torch.tensor(vals, dtype=torch.float64, device=tb.device)
        ctx.save_for_backward(gt, gr)
        ctx.squeeze = squeeze
        return out[0] if squeeze else out

    @staticmethod
    def backward(ctx, grad_out):
        gt, gr = ctx.saved_tensors
        if gt is None:
            return None, None, None, None, None, None
        go = grad_out if not ctx.squeeze else grad_out.unsqueeze(0)
        if go.shape[1] == gt.shape[1]:   # per plane
            scale = go.to(torch.float32)[:, :, None, None]
        else:                            # one value per picture
            scale = go.to(torch.float32)[:, :1, None, None]
        dt, dr = gt * scale, gr * scale
        if ctx.squeeze:
            dt, dr = dt[0], dr[0]
        return dt, dr, None, None, None, None


def ssim(tst, ref, max_value: float = 1023.0, calc_l: bool = True):
    """SSIM per plane (shape (planes,) or (N, planes))."""
    return _SSIMFamily.apply(tst, ref, "ssim", float(max_value), calc_l, None)


def msssim(tst, ref, max_value: float = 1023.0):
    """MS-SSIM per plane (float 2 x 2 mean pyramid)."""
    return _SSIMFamily.apply(tst, ref, "msssim", float(max_value), True, None)


def ivssim(tst, ref, max_value: float = 1023.0, params=None):
    """IV-SSIM of 3-plane pictures (shape (1,) or (N, 1)); tst is InputFile0, ref InputFile1."""
    return _SSIMFamily.apply(tst, ref, "ivssim", float(max_value), True, params or default_params())
