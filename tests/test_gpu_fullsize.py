"""Every benchmark workload at its full size through the HIP path, against the
oracle on the same inputs:

* 3 x 1024x768 uint8 exposures, exposure + saturation, 8 levels (bit-exact);
* the enblend step of 2 x 4096x3072 RGBA uint16 on the 7168x3072 ROI with
  its nearest-feature-transform seam mask (both bit-exact), and the whole
  enblend run from the two decoded layers to the 16-bit export (bit-exact);
* 8 x 6000x4000 float32 exposures with exposure, saturation, contrast (5x5)
  and entropy (7x7) weights, 10 levels of double pyramids (2e-6 relative,
  non-finite pattern and alpha identical);
* the 64 uint16 stacks (640x480 to 4096x3072, 3 to 9 exposures): batched ==
  one enbl_enfuse per stack on every stack, and both == the oracle on three
  sampled stacks;
* the relaxed differentiable enfuse of 5 x 2048x1536 float exposures: the
  GPU adjoint against the oracle's FP64 adjoint.

Each test prints its workload before the GPU work starts."""
import ctypes
import importlib.util
import os
import sys

import numpy as np
import pytest

from oracle_binding import F32, U8, U16, enfuse_params as orc_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load(name, rel):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, rel))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def scenes():
    return _load("enblend_scenes", os.path.join("enblend-enfuse_amd", "scenes.py"))


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _rel_err(got, want):
    """Largest relative difference over finite pixels; inf when the
    non-finite patterns differ or a zero is not reproduced."""
    g, c = got.astype(np.float64), want.astype(np.float64)
    if not (np.array_equal(np.isnan(g), np.isnan(c)) and np.array_equal(np.isinf(g), np.isinf(c))
            and np.array_equal(g[np.isinf(c)], c[np.isinf(c)])):
        return float("inf")
    fin = np.isfinite(c)
    g, c = g[fin], c[fin]
    zero = c == 0
    if np.any(g[zero] != 0):
        return float("inf")
    nz = ~zero
    return float(np.max(np.abs(g[nz] - c[nz]) / np.abs(c[nz]))) if nz.any() else 0.0


def test_enfuse_3x1024x768_u8_full(enbl, gpu_ctx, oracle, scenes):
    print("full size: enfuse 3 x 1024x768 uint8, exposure 1.0 + saturation 0.2", flush=True)
    host = scenes.exposure_stack(1024, 768, 3, 0, bits=8)
    h, w = host[0].shape[:2]
    L = enbl.auto_levels(w, h)
    al = [np.full((h, w), 255, np.uint8) for _ in host]
    imgs = [torch.from_numpy(i).cuda() for i in host]
    das = [torch.from_numpy(a).cuda() for a in al]
    out = torch.zeros_like(imgs[0])
    oa = torch.zeros((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    enbl.enfuse(gpu_ctx, enbl.U8, imgs, das, enbl.enfuse_params(w_exposure=1.0, w_saturation=0.2), L, out, oa,
                dither=enbl.DITHER_DET, seed=1)
    gpu_ctx.sync()
    want, want_a = oracle.enfuse(host, al, U8, orc_params(w_exposure=1.0, w_saturation=0.2), L, dither=1, seed=1)
    np.testing.assert_array_equal(oa.cpu().numpy(), want_a)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_enblend_step_2x4096x3072_u16_full(enbl, gpu_ctx, oracle, scenes):
    print("full size: NFT seam mask + enblend step, 2 x 4096x3072 RGBA uint16, 7168x3072 ROI", flush=True)
    white, wa, black, ba, _ = scenes.enblend_pair(seed=1)
    h, w, C = white.shape
    levels = enbl.auto_levels(w, h)
    lib = enbl.lib()
    d_white, d_wa = torch.from_numpy(white).cuda(), torch.from_numpy(wa).cuda()
    d_black, d_ba = torch.from_numpy(black).cuda(), torch.from_numpy(ba).cuda()
    d_mask = torch.empty((h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    enbl.check(lib.enbl_nft_mask(gpu_ctx.h, w, h, _vp(d_wa), w, _vp(d_ba), w, 2, _vp(d_mask), w))
    gpu_ctx.sync()
    mask = d_mask.cpu().numpy()
    want_mask, rc = oracle.nft_mask(wa, ba)
    assert rc == 0
    np.testing.assert_array_equal(mask, want_mask)
    enbl.check(lib.enbl_enblend_pair(gpu_ctx.h, enbl.U16, C, w, h, levels, 0, _vp(d_white), w * C, _vp(d_wa), w,
                                     _vp(d_black), w * C, _vp(d_ba), w, _vp(d_mask), w, enbl.DITHER_DET, 1))
    gpu_ctx.sync()
    want, want_a = oracle.enblend_pair(white, wa, black, ba, mask, levels, False, U16, 1, 1)
    np.testing.assert_array_equal(d_ba.cpu().numpy(), want_a)
    np.testing.assert_array_equal(d_black.cpu().numpy(), want)


def test_enblend_layers_7168x3072_u16_full(enbl, gpu_ctx, oracle, scenes):
    print("full size: enblend of 2 decoded RGBA uint16 layers 4096x3072 on a 7168x3072 canvas", flush=True)
    white, wa, black, ba, _ = scenes.enblend_pair(seed=1)
    H, W = wa.shape
    wimg = 4096
    off = W - wimg

    def rgba(img, a, x0):
        o = np.empty((H, wimg, 4), np.uint16)
        o[..., :3] = img[:, x0:x0 + wimg]
        o[..., 3] = a[:, x0:x0 + wimg].astype(np.uint16) * 257
        return o

    host_layers = [(rgba(black, ba, 0), 0, 0, enbl.U16), (rgba(white, wa, off), off, 0, enbl.U16)]
    dl = [(torch.from_numpy(a).cuda(), x, y, ft) for a, x, y, ft in host_layers]
    tab = enbl.layer_table(dl)
    lib = enbl.lib()
    d_b = torch.empty((H, W * 3), dtype=torch.uint16, device="cuda")
    d_ba = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    d_out = torch.empty((H, W * 4), dtype=torch.uint16, device="cuda")
    bb = enbl.Rect()
    counts = (ctypes.c_int * 3)()
    ct = ctypes.c_int(-1)
    torch.cuda.synchronize()
    enbl.check(lib.enbl_enblend_layers(gpu_ctx.h, enbl.U16, 3, W, H, 2, tab, 0, 0, 0, 1, 2, enbl.DITHER_DET, 1,
                                       _vp(d_b), W * 3, _vp(d_ba), W, ctypes.byref(bb), counts))
    enbl.check(lib.enbl_checkpoint(gpu_ctx.h, enbl.U16, 3, W, H, _vp(d_b), W * 3, _vp(d_ba), W, enbl.U16, _vp(d_out),
                                   W * 4, ctypes.byref(ct)))
    gpu_ctx.sync()
    got = d_out.cpu().numpy().reshape(H, W, 4)
    del dl
    taken = [0, 0]
    _, bimg, bal, bbb, taken = oracle.assemble(host_layers, U16, 3, W, H, taken, True)
    _, wimg_, wal, wbb, taken = oracle.assemble(host_layers, U16, 3, W, H, taken, True)
    rc, _, bimg, bal, bbb, _ = oracle.enblend_step(bimg, bal, bbb, wimg_, wal, wbb, U16, seed=1)
    assert rc == 0
    want, _ = oracle.checkpoint(bimg, bal, U16, U16)
    assert list(counts) == [1, 0, 0]
    np.testing.assert_array_equal(got, want)


def test_enfuse_8x6000x4000_f32_full(enbl, gpu_ctx, oracle, scenes):
    print("full size: enfuse 8 x 6000x4000 float32, exposure + saturation + contrast 5x5 + entropy 7x7", flush=True)
    W, H = 6000, 4000
    imgs, target = scenes.float_scene(torch, W, H, [k - 3.5 for k in range(8)], seed=2)
    del target
    kw = dict(w_exposure=1.0, w_saturation=0.2, w_contrast=0.5, w_entropy=0.3, contrast_window=5, entropy_window=7)
    alphas = [torch.full((H, W), 255, dtype=torch.uint8, device="cuda") for _ in imgs]
    out = torch.zeros_like(imgs[0])
    oa = torch.zeros((H, W), dtype=torch.uint8, device="cuda")
    L = enbl.auto_levels(W, H)
    torch.cuda.synchronize()
    enbl.enfuse(gpu_ctx, enbl.F32, imgs, alphas, enbl.enfuse_params(**kw), L, out, oa, dither=enbl.DITHER_DET,
                seed=1)
    gpu_ctx.sync()
    got, got_a = out.cpu().numpy(), oa.cpu().numpy()
    host = [t.cpu().numpy() for t in imgs]
    del imgs, alphas, out, oa
    torch.cuda.empty_cache()
    al = [np.full((H, W), 255, np.uint8) for _ in host]
    print("full size: oracle enfuse of the 8 float exposures (about 40 s on 16 threads)", flush=True)
    want, want_a = oracle.enfuse(host, al, F32, orc_params(**kw), L, dither=1, seed=1)
    np.testing.assert_array_equal(got_a, want_a)
    assert _rel_err(got, want) <= 2e-6


def test_batch_64_stacks_u16_full(enbl, gpu_ctx, oracle, scenes):
    print("full size: 64 uint16 stacks, batched vs one enbl_enfuse per stack vs the oracle", flush=True)
    plan = scenes.batch_plan()
    stacks = []
    for s, (w, h, K, ev) in enumerate(plan):
        imgs = scenes.device_stack(torch, w, h, K, 1000 + s, ev, bits=16)
        alphas = [torch.full((h, w), 255, dtype=torch.uint8, device="cuda") for _ in imgs]
        stacks.append((imgs, alphas, enbl.auto_levels(w, h)))
    prm = enbl.enfuse_params(w_exposure=1.0, w_saturation=0.2)
    plain = [(torch.zeros_like(i[0]), torch.zeros(i[0].shape[:2], dtype=torch.uint8, device="cuda"))
             for i, _, _ in stacks]
    batched = [(torch.full_like(i[0], 77), torch.full(i[0].shape[:2], 9, dtype=torch.uint8, device="cuda"))
               for i, _, _ in stacks]
    torch.cuda.synchronize()
    for (imgs, alphas, L), (o, oa) in zip(stacks, plain):
        enbl.enfuse(gpu_ctx, enbl.U16, imgs, alphas, prm, L, o, oa, dither=enbl.DITHER_DET, seed=1)
    enbl.enfuse_batch(gpu_ctx, enbl.U16, [(imgs, alphas, o, oa, L, 1) for (imgs, alphas, L), (o, oa)
                                          in zip(stacks, batched)], prm, dither=enbl.DITHER_DET)
    gpu_ctx.sync()
    for s, ((a, aa), (b, ba)) in enumerate(zip(plain, batched)):
        assert torch.equal(aa, ba), f"stack {s} {plan[s]}: alpha"
        assert torch.equal(a, b), f"stack {s} {plan[s]}: image"
    sizes = [w * h * K for (w, h, K, _) in plan]
    picks = sorted({int(np.argmax(sizes)), int(np.argmin(sizes)),
                    next(i for i, p in enumerate(plan) if p[2] == 3 and i != int(np.argmin(sizes)))})
    oprm = orc_params(w_exposure=1.0, w_saturation=0.2)
    for s in picks:
        imgs, alphas, L = stacks[s]
        print(f"full size: oracle enfuse of stack {s} {plan[s][:3]}", flush=True)
        want, want_a = oracle.enfuse([i.cpu().numpy() for i in imgs], [a.cpu().numpy() for a in alphas], U16, oprm,
                                     L, dither=1, seed=1)
        np.testing.assert_array_equal(batched[s][1].cpu().numpy(), want_a)
        np.testing.assert_array_equal(batched[s][0].cpu().numpy(), want)


def test_adjoint_5x2048x1536_full(enbl, gpu_ctx, oracle, scenes):
    print("full size: adjoint of the relaxed enfuse, 5 x 2048x1536 float64, against the oracle's adjoint", flush=True)
    diff = _load("enblend_diff", os.path.join("enblend-enfuse_amd", "diff.py"))
    f32, target = scenes.float_scene(torch, 2048, 1536, [-4, -2, 0, 2, 4], seed=4)
    imgs = torch.stack([t.to(torch.float64) for t in f32]).contiguous()
    K, h, w, _ = imgs.shape
    L = enbl.auto_levels(w, h)
    prm = torch.tensor([1.0, 0.2, 0.5, 0.5, 0.2], dtype=torch.float64)
    out = diff.forward(enbl, gpu_ctx, imgs, prm, L)
    gout = 2.0 * (out - target)
    gi, gp = diff.backward(enbl, gpu_ctx, imgs, prm, L, gout)
    gpu_ctx.sync()
    ih, gh = imgs.cpu().numpy(), gout.cpu().numpy()
    want_out = oracle.diff_forward(ih, prm.numpy(), L)
    np.testing.assert_allclose(out.cpu().numpy(), want_out, rtol=1e-12, atol=1e-14 * np.abs(want_out).max())
    ogi, ogp = oracle.diff_backward(ih, prm.numpy(), L, gh)
    gi_h, gp_h = gi.cpu().numpy(), gp.cpu().numpy()
    big = float(np.abs(ogi).max())
    assert np.abs(gi_h - ogi).max() <= 1e-13 * big
    nz = np.abs(ogi) > 1e-6 * big
    assert np.max(np.abs(gi_h[nz] - ogi[nz]) / np.abs(ogi[nz])) <= 1e-10
    np.testing.assert_allclose(gp_h, ogp, rtol=1e-10, atol=0)
