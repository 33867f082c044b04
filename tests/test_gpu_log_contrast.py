"""GPU contrast by Laplacian-of-Gaussian edges (enfuse.hip edge_* kernels)
against the oracle: the weight image bit for bit, with and without local
contrast enhancement, under both the clamping and the fill-in functor, and
the whole enfuse in that mode."""
import ctypes

import numpy as np
import pytest

from oracle_binding import F32, U8, U16, enfuse_params as orc_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _image(rng, itype, h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.35 * np.sin(xx / 4.0) * np.cos(yy / 3.0) + 0.1 * (rng.random((h, w)) - 0.5)
    rgb = np.clip(base[..., None] * rng.uniform(0.7, 1.2, size=3)[None, None, :], 0, 1)
    if itype == F32:
        return rgb.astype(np.float32)
    m = 255 if itype == U8 else 65535
    return np.rint(rgb * m).astype(np.uint8 if itype == U8 else np.uint16)


EDGE_CASES = [
    dict(edge_scale=0.6),
    dict(edge_scale=1.3, min_curvature=-0.02),
    dict(edge_scale=0.9, lce_scale=2.5, lce_factor=0.6),
    dict(edge_scale=0.8, lce_scale=1.2, lce_factor=0.2, min_curvature=0.05, contrast_window=5),
    dict(edge_scale=1.0, min_curvature=0.1, contrast_window=3),
]


def _scaled(kw, itype):
    """min_curvature is given as a fraction of the scalar maximum here and
    passed instantiated for the image type, as the reference does for a
    percentage."""
    kw = dict(kw)
    if "min_curvature" in kw:
        mx = {U8: 255.0, U16: 65535.0, F32: 1.0}[itype]
        v = kw["min_curvature"] * mx
        kw["min_curvature"] = float(int(v)) if itype != F32 else float(np.float32(v))
    return kw


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("case", range(len(EDGE_CASES)))
def test_edge_contrast_mask_matches_oracle(enbl, gpu_ctx, oracle, itype, case):
    rng = np.random.default_rng(900 + 10 * itype + case)
    h, w = 67, 301
    img = _image(rng, itype, h, w)
    a = np.where(rng.random((h, w)) < 0.85, 255, 0).astype(np.uint8)
    kw = dict(w_exposure=0.0, w_contrast=1.0, **_scaled(EDGE_CASES[case], itype))
    print(f"edge contrast mask itype={itype} case={kw}", flush=True)
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(img).cuda(), torch.from_numpy(a).cuda()
    dm = torch.full((h, w), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    np.testing.assert_array_equal(dm.cpu().numpy(), want)


@pytest.mark.parametrize("itype", [U8, U16, F32])
def test_edge_contrast_enfuse_matches_oracle(enbl, gpu_ctx, oracle, itype):
    rng = np.random.default_rng(950 + itype)
    h, w, K = 96, 130, 3
    imgs = [_image(rng, itype, h, w) for _ in range(K)]
    alphas = [np.where(rng.random((h, w)) < 0.95, 255, 0).astype(np.uint8) for _ in range(K)]
    L = oracle.auto_levels(w, h)
    kw = dict(w_exposure=1.0 if itype != F32 else 0.0, w_saturation=0.2, w_contrast=0.7,
              **_scaled(dict(edge_scale=0.7, lce_scale=1.5, lce_factor=0.3, min_curvature=0.04), itype))
    print(f"edge contrast enfuse itype={itype} {w}x{h} K={K}", flush=True)
    want, want_a = oracle.enfuse(imgs, alphas, itype, orc_params(**kw), L, dither=1, seed=3)
    dimgs = [torch.from_numpy(i).cuda() for i in imgs]
    das = [torch.from_numpy(x).cuda() for x in alphas]
    out = torch.zeros_like(dimgs[0])
    oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    enbl.enfuse(gpu_ctx, itype, dimgs, das, enbl.enfuse_params(**kw), L, out, oa, dither=1, seed=3)
    gpu_ctx.sync()
    np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
    got = out.cpu().numpy()
    if itype == F32:
        sel = want_a != 0
        np.testing.assert_allclose(got[sel], want[sel], rtol=2e-6, atol=0)
    else:
        np.testing.assert_array_equal(got, want)


def test_edge_kernel_longer_than_line(enbl, gpu_ctx):
    h, w = 6, 40
    dimg = torch.zeros((h, w, 3), dtype=torch.uint16, device="cuda")
    da = torch.full((h, w), 255, dtype=torch.uint8, device="cuda")
    dm = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    kw = dict(w_exposure=0.0, w_contrast=1.0, edge_scale=2.0)
    rc = enbl.lib().enbl_enfuse_mask(gpu_ctx.h, U16, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                     ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w)
    assert rc == enbl.ENBL_E_PRECONDITION
