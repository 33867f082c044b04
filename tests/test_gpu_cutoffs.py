"""GPU --exposure-cutoff and --entropy-cutoff against the oracle: weight
images bit for bit and one whole enfuse."""
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


def _cut(itype, lo, hi):
    lo, hi = lo * MAX[itype], hi * MAX[itype]
    return (float(int(lo)), float(int(hi))) if itype != F32 else (float(np.float32(lo)), float(np.float32(hi)))


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("kind", ["exposure", "entropy", "both"])
def test_cutoff_masks_match_oracle(enbl, gpu_ctx, oracle, itype, kind):
    rng = np.random.default_rng(40 + itype + len(kind))
    h, w = 57, 203
    img = _img(rng, itype, h, w)
    a = np.where(rng.random((h, w)) < 0.9, 255, 0).astype(np.uint8)
    kw = dict(w_exposure=1.0 if itype != F32 else 0.0, w_saturation=0.2, w_entropy=0.5, entropy_window=5)
    if kind in ("exposure", "both"):
        kw["w_exposure"] = 1.0
        kw["exposure_cutoff"] = _cut(itype, 0.1, 0.85)
    if kind in ("entropy", "both"):
        kw["entropy_cutoff"] = _cut(itype, 0.25, 0.75)
    print(f"cutoff mask itype={itype} kind={kind}", flush=True)
    want = oracle.enfuse_mask(img, a, itype, orc_params(**kw))
    dimg, da = torch.from_numpy(img).cuda(), torch.from_numpy(a).cuda()
    dm = torch.full((h, w), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_enfuse_mask(gpu_ctx.h, itype, w, h, _vp(dimg), 3 * w, _vp(da), w,
                                           ctypes.byref(enbl.enfuse_params(**kw)), _vp(dm), w))
    gpu_ctx.sync()
    got = dm.cpu().numpy()
    if itype == F32 and kw["w_exposure"] > 0.0:
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)   # device exp in the float exposure weight
    else:
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("itype", [U8, U16])
def test_cutoff_enfuse_matches_oracle(enbl, gpu_ctx, oracle, itype):
    rng = np.random.default_rng(70 + itype)
    h, w, K = 80, 111, 3
    imgs = [_img(rng, itype, h, w) for _ in range(K)]
    alphas = [np.full((h, w), 255, np.uint8) for _ in range(K)]
    L = oracle.auto_levels(w, h)
    kw = dict(w_exposure=1.0, w_saturation=0.2, w_entropy=0.3, entropy_window=3,
              exposure_cutoff=_cut(itype, 0.05, 0.95), entropy_cutoff=_cut(itype, 0.1, 0.9))
    want, want_a = oracle.enfuse(imgs, alphas, itype, orc_params(**kw), L, dither=1, seed=9)
    dimgs = [torch.from_numpy(i).cuda() for i in imgs]
    das = [torch.from_numpy(x).cuda() for x in alphas]
    out = torch.zeros_like(dimgs[0])
    oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    enbl.enfuse(gpu_ctx, itype, dimgs, das, enbl.enfuse_params(**kw), L, out, oa, dither=1, seed=9)
    gpu_ctx.sync()
    np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_exposure_cutoff_batch_falls_back_per_stack(enbl, gpu_ctx, oracle):
    """The batched entry accepts the cutoff and runs such stacks one by one."""
    rng = np.random.default_rng(77)
    h, w, K = 48, 64, 3
    imgs = [_img(rng, U8, h, w) for _ in range(K)]
    alphas = [np.full((h, w), 255, np.uint8) for _ in range(K)]
    L = oracle.auto_levels(w, h)
    kw = dict(w_exposure=1.0, w_saturation=0.2, exposure_cutoff=(20.0, 240.0))
    want, want_a = oracle.enfuse(imgs, alphas, U8, orc_params(**kw), L, dither=1, seed=4)
    dimgs = [torch.from_numpy(i).cuda() for i in imgs]
    das = [torch.from_numpy(x).cuda() for x in alphas]
    out = torch.zeros_like(dimgs[0])
    oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    enbl.enfuse_batch(gpu_ctx, U8, [(dimgs, das, out, oa, L, 4)], enbl.enfuse_params(**kw), dither=1)
    gpu_ctx.sync()
    np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
    np.testing.assert_array_equal(out.cpu().numpy(), want)
