"""Check-and-tuning harness of the relaxed differentiable enfuse: a
torch.autograd.Function over the C ABI's forward and hand-written adjoint
(enbl_diff_enfuse_forward / enbl_diff_enfuse_backward).  Device buffers come
from torch tensors; all arithmetic runs in the HIP library."""
import ctypes

import torch


class DiffParams(ctypes.Structure):
    """enbl_diff_params (include/enblend_amd.h)."""
    _fields_ = [("w_exposure", ctypes.c_double), ("w_saturation", ctypes.c_double),
                ("w_contrast", ctypes.c_double), ("exposure_optimum", ctypes.c_double),
                ("exposure_width", ctypes.c_double)]


PARAM_NAMES = ("w_exposure", "w_saturation", "w_contrast", "exposure_optimum", "exposure_width")


def _params(p):
    v = [float(x) for x in (p.detach().cpu().tolist() if torch.is_tensor(p) else p)]
    return DiffParams(*v)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def forward(enbl, ctx, imgs, params, levels, window=5):
    """imgs: (K, h, w, 3) float64 contiguous CUDA tensor -> fused (h, w, 3) float64."""
    K, h, w, _ = imgs.shape
    out = torch.empty((h, w, 3), dtype=torch.float64, device=imgs.device)
    prm = _params(params)
    torch.cuda.synchronize(imgs.device)   # inputs written on torch's stream; the library runs on its own
    enbl.check(enbl.lib().enbl_diff_enfuse_forward(ctx.h, w, h, K, _vp(imgs), ctypes.byref(prm), levels, window,
                                                   _vp(out)))
    ctx.sync()
    return out


def backward(enbl, ctx, imgs, params, levels, gout, window=5):
    """Returns (d/d imgs, d/d params) for the output gradient gout."""
    K, h, w, _ = imgs.shape
    gimgs = torch.empty_like(imgs)
    gparams = torch.empty(5, dtype=torch.float64, device=imgs.device)
    prm = _params(params)
    gout = gout.contiguous()
    torch.cuda.synchronize(imgs.device)
    enbl.check(enbl.lib().enbl_diff_enfuse_backward(ctx.h, w, h, K, _vp(imgs), ctypes.byref(prm), levels, window,
                                                    _vp(gout), _vp(gimgs), _vp(gparams)))
    ctx.sync()
    return gimgs, gparams


class DiffEnfuse(torch.autograd.Function):
    """fused = DiffEnfuse.apply(imgs, params, enbl, ctx, levels, window)."""

    @staticmethod
    def forward(fctx, imgs, params, enbl, ctx, levels, window):
        imgs = imgs.contiguous()
        fctx.save_for_backward(imgs, params)
        fctx.enbl, fctx.ctx, fctx.levels, fctx.window = enbl, ctx, levels, window
        return forward(enbl, ctx, imgs, params, levels, window)

    @staticmethod
    def backward(fctx, gout):
        imgs, params = fctx.saved_tensors
        gimgs, gparams = backward(fctx.enbl, fctx.ctx, imgs, params, fctx.levels, gout, fctx.window)
        return gimgs, gparams.to(params.device, params.dtype), None, None, None, None
