"""The device transform of float images into double pyramids and back
(1 + log1p(v) or 1 / (1 - v), and expm1(c - 1) or 1 - 1 / c;
src/fixmath.h:165-186, :320-329) through enbl_copy_to_pyramid and
enbl_copy_from_pyramid_if, over floats of every exponent and sign: within 2
units in the last place of the double pyramid value computed with the C
library, and the same float back as the C library's inverse of the device's
pyramid values."""
import ctypes

import numpy as np
import pytest

from oracle_binding import F32, F64

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _floats(n, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    v = bits.view(np.float32)
    v = np.where(np.isfinite(v), v, np.float32(1.5))
    special = np.array([0.0, -0.0, 1e-45, 2 ** -30, 2 ** -29, 2 ** -28, 1e-7, 0.5, 1.0, 2.0, 3.4028235e38,
                        np.inf, -1e-45, -0.5, -1.0, -3.4028235e38, -np.inf], np.float32)
    v[:len(special)] = special
    return v


def test_transform_against_libm(enbl, gpu_ctx):
    w, h = 1024, 1024
    v = _floats(w * h, 11).reshape(h, w)
    dimg = torch.from_numpy(v).cuda()
    p = enbl.Pyramid(gpu_ctx, F64, 1, w, h, 1)
    torch.cuda.synchronize()
    print(f"copy_to_pyramid float {w}x{h}", flush=True)
    enbl.check(enbl.lib().enbl_copy_to_pyramid(gpu_ctx.h, F32, 0, _vp(dimg), w, p.h))
    gpu_ctx.sync()
    got = p.read(0)[0]
    d = v.astype(np.float64)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        want = np.where(d >= 0, 1.0 + np.log1p(np.maximum(d, 0)), 1.0 / (1.0 - d))
    fin = np.isfinite(want)
    np.testing.assert_array_equal(np.isfinite(got), fin)
    np.testing.assert_array_equal(got[~fin], want[~fin])
    ulp = np.spacing(np.abs(want[fin]))
    err = np.abs(got[fin] - want[fin]) / ulp
    print("max ulp", err.max())
    assert err.max() <= 2.0
    # back to float under a full mask
    dm = torch.full((h, w), 255, dtype=torch.uint8, device="cuda")
    out = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    print("copy_from_pyramid_if float", flush=True)
    enbl.check(enbl.lib().enbl_copy_from_pyramid_if(gpu_ctx.h, p.h, 0, 0, _vp(dm), w, _vp(out), F32, w, 0, 0))
    gpu_ctx.sync()
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        back = np.where(got >= 1.0, np.expm1(got - 1.0), 1.0 - 1.0 / got).astype(np.float32)
    np.testing.assert_array_equal(out.cpu().numpy(), back)
    p.close()
