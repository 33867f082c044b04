"""enbl_gaussian_pyramid_noalloc (per-level alpha in the library's static
device scratch array) against enbl_gaussian_pyramid (context scratch), bit
for bit, on two calls in a row, the second smaller; and its size limit."""
import ctypes

import numpy as np
import pytest

from oracle_binding import F32, F64, I16, I32, U8, U16

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

PT = {U8: I16, U16: I32, F32: F64}
FRAC = {U8: 7, U16: 7, F32: 0}


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _img(rng, itype, h, w, C):
    if itype == U8:
        return rng.integers(0, 256, size=(h, w, C)).astype(np.uint8)
    if itype == U16:
        return rng.integers(0, 65536, size=(h, w, C)).astype(np.uint16)
    return np.exp(rng.normal(0, 2, size=(h, w, C))).astype(np.float32)


def _both(enbl, ctx, itype, C, img, alpha, wrap):
    h, w = img.shape[0], img.shape[1]
    L = enbl.auto_levels(w, h)
    dimg = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    da = None if alpha is None else torch.from_numpy(np.ascontiguousarray(alpha)).cuda()
    ref = enbl.Pyramid(ctx, PT[itype], C, w, h, L)
    new = enbl.Pyramid(ctx, PT[itype], C, w, h, L)
    torch.cuda.synchronize()
    enbl.check(enbl.lib().enbl_gaussian_pyramid(ctx.h, int(wrap), itype, FRAC[itype], _vp(dimg), w * C, _vp(da),
                                                w, ref.h))
    enbl.check(enbl.lib().enbl_gaussian_pyramid_noalloc(ctx.h, int(wrap), itype, FRAC[itype], _vp(dimg), w * C,
                                                        _vp(da), w, new.h))
    ctx.sync()
    try:
        for l in range(L):
            a, b = ref.read(l), new.read(l)
            assert a.dtype == b.dtype
            np.testing.assert_array_equal(b.view(np.uint8), a.view(np.uint8), err_msg=f"level {l}")
    finally:
        ref.close()
        new.close()


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("wrap", [False, True])
def test_noalloc_gaussian_pyramid_equals_allocating_one(enbl, gpu_ctx, itype, wrap):
    rng = np.random.default_rng(900 + itype + 10 * wrap)
    for (h, w, C) in [(1500, 2900, 3), (173, 301, 1)]:
        print(f"noalloc gaussian pyramid itype={itype} wrap={wrap} {w}x{h}x{C}", flush=True)
        img = _img(rng, itype, h, w, C)
        alpha = np.where(rng.random((h, w)) < 0.85, 255, 0).astype(np.uint8)
        _both(enbl, gpu_ctx, itype, C, img, alpha, wrap)
        _both(enbl, gpu_ctx, itype, C, img, None, wrap)


def test_noalloc_gaussian_pyramid_rejects_oversized_alpha(enbl, gpu_ctx):
    """Level-1 alpha of 5000 x 4000: 2 * 5120 * 4000 bytes exceed the 16 MiB
    the library holds; rejected before any launch."""
    h, w = 8000, 10000
    print(f"noalloc gaussian pyramid size limit {w}x{h}", flush=True)
    img = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    alpha = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    p = enbl.Pyramid(gpu_ctx, I16, 1, w, h, 3)
    torch.cuda.synchronize()
    rc = enbl.lib().enbl_gaussian_pyramid_noalloc(gpu_ctx.h, 0, U8, 7, _vp(img), w, _vp(alpha), w, p.h)
    assert rc == enbl.ENBL_E_SIZE
    assert "scratch" in enbl.lib().enbl_last_error().decode()
    # without alpha no scratch is needed
    enbl.check(enbl.lib().enbl_gaussian_pyramid_noalloc(gpu_ctx.h, 0, U8, 7, _vp(img), w, None, 0, p.h))
    gpu_ctx.sync()
    p.close()
