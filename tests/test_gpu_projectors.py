"""GPU grayscale projectors (--gray-projector) of the exposure and contrast
weights against the oracle."""
import ctypes

import numpy as np
import pytest

from oracle_binding import F32, U8, U16, enfuse_params as orc_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NP = {U8: np.uint8, U16: np.uint16, F32: np.float32}
MAX = {U8: 255, U16: 65535, F32: 1.0}


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _img(rng, itype, h, w):
    if itype == F32:
        return rng.uniform(0, 1, size=(h, w, 3)).astype(np.float32)
    return rng.integers(0, MAX[itype] + 1, size=(h, w, 3)).astype(NP[itype])


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("proj", range(8))
def test_projector_masks_match_oracle(enbl, gpu_ctx, oracle, itype, proj):
    rng = np.random.default_rng(500 + 10 * itype + proj)
    h, w = 41, 150
    img = _img(rng, itype, h, w)
    a = np.where(rng.random((h, w)) < 0.9, 255, 0).astype(np.uint8)
    mixer = (0.3, 0.5, 0.2) if proj == 7 else (0.0, 0.0, 0.0)
    # integer images: exposure (table lookup by gray code) and contrast;
    # float images: contrast only (the device exp of the float exposure
    # weight differs from glibc's in the last bit)
    kw = dict(w_exposure=1.0 if itype != F32 else 0.0, w_contrast=0.6, contrast_window=5, gray_projector=proj,
              mixer=mixer)
    print(f"projector {proj} itype={itype}", flush=True)
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(img).cuda(), torch.from_numpy(a).cuda()
    dm = torch.full((h, w), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    np.testing.assert_array_equal(dm.cpu().numpy(), want)


@pytest.mark.parametrize("proj", [1, 6])
def test_projector_enfuse_matches_oracle(enbl, gpu_ctx, oracle, proj):
    rng = np.random.default_rng(600 + proj)
    h, w, K = 70, 97, 3
    imgs = [_img(rng, U16, h, w) for _ in range(K)]
    alphas = [np.full((h, w), 255, np.uint8) for _ in range(K)]
    L = oracle.auto_levels(w, h)
    kw = dict(w_exposure=1.0, w_saturation=0.2, w_contrast=0.4, gray_projector=proj)
    want, want_a = oracle.enfuse(imgs, alphas, U16, orc_params(**kw), L, dither=1, seed=2)
    dimgs = [torch.from_numpy(i).cuda() for i in imgs]
    das = [torch.from_numpy(x).cuda() for x in alphas]
    out = torch.zeros_like(dimgs[0])
    oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    enbl.enfuse(gpu_ctx, U16, dimgs, das, enbl.enfuse_params(**kw), L, out, oa, dither=1, seed=2)
    gpu_ctx.sync()
    np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_mixer_weights_rejected(enbl, gpu_ctx):
    d = torch.zeros((4, 4, 3), dtype=torch.uint8, device="cuda")
    a = torch.full((4, 4), 255, dtype=torch.uint8, device="cuda")
    m = torch.zeros((4, 4), dtype=torch.float32, device="cuda")
    kw = dict(w_exposure=1.0, gray_projector=7, mixer=(1.0, -0.5, 1.0))
    rc = enbl.lib().enbl_enfuse_mask(gpu_ctx.h, U8, 4, 4, _vp(d), 12, _vp(a), 4,
                                     ctypes.byref(enbl.enfuse_params(**kw)), _vp(m), 4)
    assert rc == enbl.ENBL_E_ARG
