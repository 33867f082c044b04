"""The GPU path against the committed golden records of tests/golden, bit
for bit for integer images, NFT masks and distance transforms, within 2e-6
relative for float images (device log1p/expm1/exp against glibc)."""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

import cases  # noqa: E402

with open(os.path.join(HERE, "golden.json")) as _f:
    GOLDEN = json.load(_f)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("c", cases.ENBLEND, ids=lambda c: c["name"])
def test_enblend_golden_gpu(enbl, gpu_ctx, c):
    g = GOLDEN[c["name"]]
    w, h = c["w"], c["h"]
    white, wa, black, ba = cases.enblend_inputs(c["seed"], c["itype"], w, h)
    assert cases.digest(white, wa, black, ba) == g["inputs"]
    print(f"golden {c['name']}", flush=True)
    dw, dwa, db, dba = (_dev(x) for x in (white, wa, black, ba))
    mask = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    lib = enbl.lib()
    enbl.check(lib.enbl_nft_mask(gpu_ctx.h, w, h, _vp(dwa), w, _vp(dba), w, 2, _vp(mask), w))
    gpu_ctx.sync()
    assert cases.digest(mask.cpu().numpy()) == g["mask"]
    levels = enbl.auto_levels(w, h)
    assert levels == g["levels"]
    enbl.check(lib.enbl_enblend_pair(gpu_ctx.h, c["itype"], 3, w, h, levels, int(c["wrap"]), _vp(dw), 3 * w, _vp(dwa), w,
                                     _vp(db), 3 * w, _vp(dba), w, _vp(mask), w, cases.DITHER_DET, cases.SEED))
    gpu_ctx.sync()
    assert cases.digest(db.cpu().numpy()) == g["out"]
    assert cases.digest(dba.cpu().numpy()) == g["out_alpha"]


@pytest.mark.parametrize("c", cases.ENFUSE, ids=lambda c: c["name"])
def test_enfuse_golden_gpu(enbl, gpu_ctx, c):
    g = GOLDEN[c["name"]]
    w, h = c["w"], c["h"]
    imgs, alphas = cases.enfuse_inputs(c["seed"], c["itype"], c["K"], w, h, c["holes"])
    assert cases.digest(*(imgs + alphas)) == g["inputs"]
    print(f"golden {c['name']}", flush=True)
    dimgs = [_dev(i) for i in imgs]
    das = [_dev(a) for a in alphas]
    out = torch.zeros_like(dimgs[0])
    oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    levels = enbl.auto_levels(w, h)
    enbl.enfuse(gpu_ctx, c["itype"], dimgs, das, enbl.enfuse_params(**c["prm"]), levels, out, oa,
                dither=cases.DITHER_DET, seed=cases.SEED)
    gpu_ctx.sync()
    assert cases.digest(oa.cpu().numpy()) == g["out_alpha"]
    got = out.cpu().numpy()
    if c["itype"] == cases.F32:
        want = np.load(os.path.join(HERE, c["name"] + ".npy"), allow_pickle=False)
        sel = oa.cpu().numpy() != 0
        np.testing.assert_allclose(got[sel], want[sel], rtol=2e-6, atol=0)
    else:
        assert cases.digest(got) == g["out"]


@pytest.mark.parametrize("c", cases.NFT, ids=lambda c: c["name"])
def test_nft_golden_gpu(enbl, gpu_ctx, c):
    g = GOLDEN[c["name"]]
    w, h = c["w"], c["h"]
    _, wa, _, ba = cases.enblend_inputs(c["seed"], cases.U8, w, h)
    assert cases.digest(wa, ba) == g["inputs"]
    print(f"golden {c['name']}", flush=True)
    dwa, dba = _dev(wa), _dev(ba)
    mask = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    dist = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    lib = enbl.lib()
    rc = lib.enbl_nft_mask(gpu_ctx.h, w, h, _vp(dwa), w, _vp(dba), w, 2, _vp(mask), w)
    assert (rc == 0) == (g["rc"] == 0)
    enbl.check(lib.enbl_distance_transform(gpu_ctx.h, w, h, _vp(dwa), w, 2, _vp(dist), w))
    gpu_ctx.sync()
    assert cases.digest(mask.cpu().numpy()) == g["mask"]
    assert cases.digest(dist.cpu().numpy()) == g["distance"]
