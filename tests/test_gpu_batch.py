"""GPU parity of batched enfuse (enbl_enfuse_batch) against the plain
per-stack path (enbl_enfuse) and the oracle, bit-exact."""
import numpy as np
import pytest

from oracle_binding import F32, U8, U16, enfuse_params as orc_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _stack(rng, itype, K, h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.4 * np.sin(xx / 7.0 + rng.uniform(0, 3)) * np.cos(yy / 5.0)
    tint = rng.uniform(0.7, 1.3, size=3)
    imgs = []
    for k in range(K):
        rad = base[..., None] * tint[None, None, :] * 2.0 ** ((k - (K - 1) / 2) * 1.0)
        if itype == F32:   # linear radiance, values above 1 included
            imgs.append((rad * (1.0 + 0.01 * rng.normal(0, 1, rad.shape))).astype(np.float32))
            continue
        m = 255 if itype == U8 else 65535
        v = np.clip(rad, 0, 1) ** (1 / 2.2) * m + rng.normal(0, 1, rad.shape)
        imgs.append(np.clip(np.rint(v), 0, m).astype(np.uint8 if itype == U8 else np.uint16))
    return imgs


def _alpha(rng, h, w, kind):
    if kind == "full":
        return np.full((h, w), 255, np.uint8)
    if kind == "holes":
        return np.where(rng.random((h, w)) < 0.85, 255, 0).astype(np.uint8)
    a = np.zeros((h, w), np.uint8)   # a band: the union alpha misses parts of the canvas
    a[:, : max(1, (w * 2) // 3)] = 255
    return a


# (w, h, K, levels or None for auto, alpha kind)
GEOMS = [
    (83, 61, 3, None, "holes"),
    (1, 1, 2, None, "full"),
    (2, 1, 1, None, "full"),
    (3, 9, 3, 3, "full"),            # coarsest level 1 pixel wide
    (5, 2, 2, 2, "holes"),           # coarsest level 1 pixel high
    (130, 67, 5, None, "band"),
    (64, 48, 1, 1, "holes"),
    (200, 150, 17, None, "holes"),   # more images than the weights pass keeps in registers
    (37, 23, 4, 2, "holes"),
    (257, 129, 2, None, "full"),
]


def _make_batch(rng, itype, geoms, enbl):
    host, dev = [], []
    for (w, h, K, L, ak) in geoms:
        imgs = _stack(rng, itype, K, h, w)
        alphas = [_alpha(rng, h, w, ak) for _ in range(K)]
        levels = L if L is not None else enbl.auto_levels(w, h)
        seed = int(rng.integers(0, 1 << 40))
        fill = rng.integers(0, 255)
        odt = {U8: torch.uint8, U16: torch.uint16, F32: torch.float32}[itype]
        out = torch.full((h, w, 3), int(fill), dtype=odt, device="cuda")
        oa = torch.full((h, w), 7, dtype=torch.uint8, device="cuda")
        host.append((imgs, alphas, levels, seed, int(fill)))
        dev.append(([torch.from_numpy(i).cuda() for i in imgs], [torch.from_numpy(a).cuda() for a in alphas],
                    out, oa, levels, seed))
    torch.cuda.synchronize()
    return host, dev


def _plain(enbl, ctx, itype, dev, kw):
    res = []
    for (imgs, alphas, out, oa, levels, seed) in dev:
        o2 = torch.empty_like(out).copy_(out)
        a2 = torch.empty_like(oa).copy_(oa)
        torch.cuda.synchronize()
        enbl.enfuse(ctx, itype, imgs, alphas, enbl.enfuse_params(**kw), levels, o2, a2, dither=1, seed=seed)
        ctx.sync()
        res.append((o2.cpu().numpy(), a2.cpu().numpy()))
    return res


PARAMS = [
    dict(w_exposure=1.0, w_saturation=0.2),
    dict(w_exposure=0.7, w_saturation=0.0, exposure_function=3, exposure_optimum=0.6, exposure_width=0.25),
    dict(w_exposure=0.0, w_saturation=1.0),
]


@pytest.mark.parametrize("itype", [U8, U16])
@pytest.mark.parametrize("case", range(len(PARAMS)))
def test_batch_matches_plain_and_oracle(enbl, gpu_ctx, oracle, itype, case):
    rng = np.random.default_rng(900 + itype + 10 * case)
    kw = PARAMS[case]
    host, dev = _make_batch(rng, itype, GEOMS, enbl)
    print(f"batch itype={itype} case={case} stacks={len(dev)}", flush=True)
    plain = _plain(enbl, gpu_ctx, itype, dev, kw)
    enbl.enfuse_batch(gpu_ctx, itype, dev, enbl.enfuse_params(**kw), dither=1)
    gpu_ctx.sync()
    for s, ((imgs, alphas, levels, seed, fill), d, (po, pa)) in enumerate(zip(host, dev, plain)):
        got, got_a = d[2].cpu().numpy(), d[3].cpu().numpy()
        np.testing.assert_array_equal(got_a, pa, err_msg=f"stack {s} alpha")
        np.testing.assert_array_equal(got, po, err_msg=f"stack {s}")
        want, want_a = oracle.enfuse(imgs, alphas, itype, orc_params(**kw), levels, dither=1, seed=seed)
        sel = want_a != 0
        np.testing.assert_array_equal(got_a, want_a)
        np.testing.assert_array_equal(got[sel], want[sel], err_msg=f"stack {s} vs oracle")
        assert np.all(got[~sel] == fill)


def test_batch_graph_replay_reads_current_inputs(enbl, gpu_ctx):
    """The second call replays the captured graph: same pointers, new pixel
    values; every replay must equal the plain path on the current inputs."""
    rng = np.random.default_rng(77)
    itype = U16
    kw = PARAMS[0]
    host, dev = _make_batch(rng, itype, GEOMS, enbl)
    for rep in range(3):
        if rep:
            for (imgs, alphas, out, oa, levels, seed) in dev:
                for t in imgs:
                    t.copy_(torch.from_numpy(_stack(rng, itype, 1, t.shape[0], t.shape[1])[0]).cuda())
            torch.cuda.synchronize()
        print(f"batch replay {rep}", flush=True)
        plain = _plain(enbl, gpu_ctx, itype, dev, kw)
        enbl.enfuse_batch(gpu_ctx, itype, dev, enbl.enfuse_params(**kw), dither=1)
        gpu_ctx.sync()
        for s, (d, (po, pa)) in enumerate(zip(dev, plain)):
            np.testing.assert_array_equal(d[3].cpu().numpy(), pa)
            np.testing.assert_array_equal(d[2].cpu().numpy(), po, err_msg=f"replay {rep} stack {s}")


STACK_WEIGHT_PARAMS = [
    dict(w_exposure=1.0, w_saturation=0.2, w_contrast=0.5, w_entropy=0.3, contrast_window=5, entropy_window=3),
    dict(w_exposure=1.0, w_saturation=0.2, hard_mask=True),
    dict(w_exposure=0.8, w_contrast=0.4, contrast_window=3, gray_projector=2),   # + an exposure cutoff
]


@pytest.mark.parametrize("itype", [U8, U16, F32])
@pytest.mark.parametrize("case", range(len(STACK_WEIGHT_PARAMS)), ids=["contrast-entropy", "hard-mask", "cutoff-projector"])
def test_batch_stack_weights_match_plain_and_oracle(enbl, gpu_ctx, oracle, itype, case):
    """Weights the batched weights pass does not cover (contrast, entropy,
    hard masks, cutoffs, other projectors) are made stack by stack, then every
    stack goes through the batched pyramid passes: bit-exact with one
    enbl_enfuse per stack and with the oracle, twice in a row (the second
    call replays the captured graph of the pyramid passes).  Float stacks:
    bit-exact with enbl_enfuse, within 2e-6 relative of the oracle (whose
    log/exp transform is the C library's) with the alpha exact."""
    rng = np.random.default_rng(78 + 10 * itype + case)
    kw = dict(STACK_WEIGHT_PARAMS[case])
    if case == 2:
        mx = {U8: 255.0, U16: 65535.0, F32: 1.0}[itype]
        lo, hi = 0.05 * mx, 0.95 * mx
        kw["exposure_cutoff"] = (lo, hi) if itype == F32 else (float(int(lo)), float(int(hi)))
    # (local windows need images at least as large as the window)
    host, dev = _make_batch(rng, itype, [g for g in GEOMS if min(g[0], g[1]) >= 7], enbl)
    for rep in range(2):
        print(f"batch stack weights itype={itype} case={case} call {rep}", flush=True)
        plain = _plain(enbl, gpu_ctx, itype, dev, kw)
        enbl.enfuse_batch(gpu_ctx, itype, dev, enbl.enfuse_params(**kw), dither=1)
        gpu_ctx.sync()
        for s, ((imgs, alphas, levels, seed, fill), d, (po, pa)) in enumerate(zip(host, dev, plain)):
            got, got_a = d[2].cpu().numpy(), d[3].cpu().numpy()
            np.testing.assert_array_equal(got_a, pa, err_msg=f"stack {s} alpha")
            np.testing.assert_array_equal(got, po, err_msg=f"stack {s}")
            if rep == 0:
                want, want_a = oracle.enfuse(imgs, alphas, itype, orc_params(**kw), levels, dither=1, seed=seed)
                sel = want_a != 0
                np.testing.assert_array_equal(got_a, want_a)
                if itype == F32:
                    np.testing.assert_allclose(got[sel], want[sel], rtol=2e-6, atol=0, err_msg=f"stack {s} vs oracle")
                else:
                    np.testing.assert_array_equal(got[sel], want[sel], err_msg=f"stack {s} vs oracle")


def test_batch_float_special_values_match_plain(enbl, gpu_ctx):
    """Float stacks with negative samples (the reciprocal branch of the log
    transform), +inf, NaN and huge values: the batched pyramid passes give
    exactly what one enbl_enfuse per stack gives (NaN where it gives NaN)."""
    rng = np.random.default_rng(81)
    host, dev = _make_batch(rng, F32, [(83, 61, 3, None, "holes"), (130, 67, 4, None, "full"), (37, 23, 2, 2, "band")],
                            enbl)
    for (imgs, alphas, out, oa, levels, seed) in dev:
        for k, t in enumerate(imgs):
            a = t.cpu().numpy()
            a[3, 5, 0] = -0.5
            a[7, 2, 1] = np.inf
            a[10 + k, 11, 2] = np.nan
            a[1, 1, :] = 3.0e38
            a[2, 9, 2] = -0.0
            t.copy_(torch.from_numpy(a).cuda())
    torch.cuda.synchronize()
    kw = dict(w_exposure=1.0, w_saturation=0.2, w_contrast=0.3, contrast_window=3)
    print("batch float special values", flush=True)
    plain = _plain(enbl, gpu_ctx, F32, dev, kw)
    enbl.enfuse_batch(gpu_ctx, F32, dev, enbl.enfuse_params(**kw), dither=1)
    gpu_ctx.sync()
    for s_, (d, (po, pa)) in enumerate(zip(dev, plain)):
        np.testing.assert_array_equal(d[3].cpu().numpy(), pa, err_msg=f"stack {s_} alpha")
        np.testing.assert_array_equal(d[2].cpu().numpy(), po, err_msg=f"stack {s_}")


def test_batch_autotune_keeps_results(enbl, gpu_ctx):
    rng = np.random.default_rng(79)
    kw = PARAMS[0]
    host, dev = _make_batch(rng, U16, [(300, 200, 3, None, "holes"), (97, 131, 4, None, "band")], enbl)
    plain = _plain(enbl, gpu_ctx, U16, dev, kw)
    saved = enbl.batch_tuning(gpu_ctx)
    try:
        best = enbl.batch_autotune(gpu_ctx, U16, dev, enbl.enfuse_params(**kw), dither=1, repeats=1)
        print(f"autotune: {best}", flush=True)
        assert best == enbl.batch_tuning(gpu_ctx)
        assert best["rows"][0] in (4, 8, 16, 32) and best["rows"][1] in (8, 16, 32, 64) and best["rows"][2] in (4, 8, 16, 32)
        enbl.enfuse_batch(gpu_ctx, U16, dev, enbl.enfuse_params(**kw), dither=1)
        gpu_ctx.sync()
        for d, (po, pa) in zip(dev, plain):
            np.testing.assert_array_equal(d[3].cpu().numpy(), pa)
            np.testing.assert_array_equal(d[2].cpu().numpy(), po)
    finally:
        enbl.batch_tuning(gpu_ctx, saved)


def test_batch_rejects_bad_stacks(enbl, gpu_ctx):
    rng = np.random.default_rng(80)
    host, dev = _make_batch(rng, U8, [(20, 10, 2, None, "full")], enbl)
    imgs, alphas, out, oa, levels, seed = dev[0]
    with pytest.raises(enbl.EnblError):
        enbl.enfuse_batch(gpu_ctx, U8, [(imgs, alphas, out, oa, 30, seed)], enbl.enfuse_params())
    with pytest.raises(enbl.EnblError):
        enbl.enfuse_batch(gpu_ctx, U8, [(imgs, alphas, out, oa, 0, seed)], enbl.enfuse_params())
    with pytest.raises(enbl.EnblError):
        enbl.batch_tuning(gpu_ctx, {"rows": [3, 8, 8], "blocks_per_cu": [8, 8, 8]})
    # reduce() of a 1-pixel-wide level is a precondition violation, as in the reference
    host, dev = _make_batch(rng, U8, [(1, 7, 2, 3, "full")], enbl)
    with pytest.raises(enbl.EnblError, match="too small"):
        enbl.enfuse_batch(gpu_ctx, U8, dev, enbl.enfuse_params())
