#!/usr/bin/env python3
"""Benchmark of the enblend/enfuse pyramid path on one MI355X.

Headline workload (one step): enfuse of 8 float32 RGB exposures at
6000x4000 (one synthetic scene, EV -3.5..+3.5), weights exposure 1.0 +
saturation 0.2 + contrast 0.5 (5x5) + entropy 0.3 (7x7), 10 levels of double
pyramids (src/enfuse.h:1217-1609).  Inputs are resident in HBM before timing
starts.

Prints one JSON line.  `value` is input Mpix (8 x 6000 x 4000 per step) per
second of wall time over the timed steps; `host_resident` is the same step
with pinned host buffers uploaded and downloaded around it.  `roofline` is
the dominant kernel's algorithmic bytes per launch over its HIP-event launch
duration, measured on the library's stream in a profiled pass of the same
workload; `pyramid_kernels` is the same fraction over all pyramid kernels.
`cpu_baseline` times the oracle (the C++ restatement of the reference's
algorithm, OpenMP pointwise stages, single-threaded SKIPSM passes) on the
host cores for the full workload, whose output is also the parity check.
Further parts: the enblend step of two 4096x3072 RGBA uint16 images with
its seam mask (`enblend_c1`), the whole enblend run from decoded layers,
the 64-stack batch against the plain path, the small 8-bit enfuse and the
adjoint check.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
METRIC = "Mpix/s fused per GPU; % HBM peak; batched vs plain Mpix/s; adjoint rel err"


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def measured_traffic(kernel, traffic_file):
    """HBM bytes per launch of `kernel` from a committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated on a plain copy), or None."""
    try:
        with open(os.path.join(ROOT, traffic_file)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def cpu_baseline(white, wa, black, ba, mask, levels, repeats):
    """Time the oracle's enblend step on the host (checker code, bench only)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import U16, load_oracle
    o = load_oracle()
    times = []
    out = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        out = o.enblend_pair(white, wa, black, ba, mask, levels, False, U16, 1, 1)
        times.append(time.perf_counter() - t0)
    times.sort()
    return times[len(times) // 2], o.lib.orc_num_threads(), out


def batch_stacks(enbl, scenes, torch, dev):
    """The 64-stack uint16 workload: per stack K exposures (3-9) of one scene,
    sizes from 640x480 to 4096x3072, all-255 alphas, generated on the device."""
    stacks = []
    mpix = 0.0
    for s, (w, h, K, ev) in enumerate(scenes.batch_plan()):
        imgs = scenes.device_stack(torch, w, h, K, 1000 + s, ev, bits=16)
        alphas = [torch.full((h, w), 255, dtype=torch.uint8, device=dev) for _ in imgs]
        stacks.append((imgs, alphas, enbl.auto_levels(w, h)))
        mpix += K * w * h / 1e6
    torch.cuda.synchronize()
    return stacks, mpix


def _kernels(rep, runs=1):
    return {k: {"ms": round(v["ms"] / runs, 3), "launches": v["launches"] // runs,
                "GB_per_s": round(v["own_bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in sorted(rep.items())}


def batch_compare(args, enbl, scenes, torch, steps, warmup):
    """Batched (enbl_enfuse_batch) vs plain per-stack (enbl_enfuse) GPU
    throughput on the 64-stack uint16 workload, with bit-exact parity of the
    two outputs on every stack."""
    dev = torch.device("cuda:0")
    print("bench: generating the 64-stack uint16 batch", file=sys.stderr, flush=True)
    stacks, mpix = batch_stacks(enbl, scenes, torch, dev)
    ctx = enbl.Context(0)
    prm = enbl.enfuse_params(w_exposure=1.0, w_saturation=0.2)
    tuning_src = "built-in defaults"
    if args.tuning and os.path.exists(args.tuning):
        with open(args.tuning) as f:
            enbl.batch_tuning(ctx, json.load(f)["tuning"])
        tuning_src = os.path.relpath(args.tuning, ROOT)
    plain_out = [(torch.zeros_like(i[0]), torch.zeros(i[0].shape[:2], dtype=torch.uint8, device=dev)) for i, _, _ in stacks]
    batch_out = [(torch.zeros_like(i[0]), torch.zeros(i[0].shape[:2], dtype=torch.uint8, device=dev)) for i, _, _ in stacks]
    batch = [(imgs, alphas, o, oa, L, 1) for (imgs, alphas, L), (o, oa) in zip(stacks, batch_out)]

    def run_plain():
        for (imgs, alphas, L), (o, oa) in zip(stacks, plain_out):
            enbl.enfuse(ctx, enbl.U16, imgs, alphas, prm, L, o, oa, dither=enbl.DITHER_DET, seed=1)

    def run_batch():
        enbl.enfuse_batch(ctx, enbl.U16, batch, prm, dither=enbl.DITHER_DET)

    if args.autotune:
        print("bench: autotuning the batched passes", file=sys.stderr, flush=True)
        best = enbl.batch_autotune(ctx, enbl.U16, batch, prm, dither=enbl.DITHER_DET, repeats=2)
        with open(args.autotune, "w") as f:
            json.dump({"workload": "64-stack uint16 batch (scenes.batch_plan)", "tuning": best}, f, indent=1)
        tuning_src = "autotuned: " + os.path.relpath(args.autotune, ROOT)

    def timed(fn, n, label):
        print(f"bench: {label}: {warmup} warmup + {n} timed runs", file=sys.stderr, flush=True)
        for _ in range(warmup):
            fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        ctx.sync()
        return (time.perf_counter() - t0) * 1e3 / n

    ms_plain = timed(run_plain, steps, "plain per-stack enfuse")
    ms_batch = timed(run_batch, steps, "batched enfuse")
    parity = all(torch.equal(a, b) and torch.equal(aa, ba) for (a, aa), (b, ba) in zip(plain_out, batch_out))
    ctx.profile(True)
    run_batch()
    rep_b = ctx.profile_report()
    ctx.profile(False)
    ctx.profile(True)
    run_plain()
    rep_p = ctx.profile_report()
    ctx.profile(False)
    tuning = enbl.batch_tuning(ctx)
    ctx.close()
    return {"workload": "64 exposure stacks, uint16 RGB, K=3-9, 640x480..4096x3072, exposure 1.0 + saturation 0.2",
            "input_mpix": round(mpix, 1), "plain_ms": round(ms_plain, 3), "batched_ms": round(ms_batch, 3),
            "plain_mpix_per_s": round(mpix / (ms_plain * 1e-3), 1),
            "batched_mpix_per_s": round(mpix / (ms_batch * 1e-3), 1),
            "batched_vs_plain": round(ms_plain / ms_batch, 3), "parity_batched_eq_plain": bool(parity),
            "tuning": tuning, "tuning_source": tuning_src,
            "batched_kernels": _kernels(rep_b), "plain_kernels": _kernels(rep_p)}


def adjoint_c4(args, enbl, scenes, torch, fd_points=256, adam_steps=200, cpu=True):
    """Differentiable enfuse, 5 x 2048x1536 float32 exposures (EV -4..+4),
    loss = sum of squared differences from the Reinhard tone map of the scene
    radiance: hand-written adjoint vs FP64 central differences of the GPU's
    FP64 forward on 256 random input samples and all 5 parameters, then Adam on
    the 5 parameters.  Steps are 1e-4 for the parameters and 1e-4 relative for
    the image samples: the darkest exposure holds values near 6e-5, where an
    absolute 1e-4 step crosses zero.  Each difference of two losses is summed
    elementwise as (O1 - O2)(O1 + O2 - 2T)."""
    import numpy as np
    diff = _load("enblend_diff", os.path.join(ROOT, "enblend-enfuse_amd", "diff.py"))
    dev = torch.device("cuda:0")
    print("bench: generating the 5 x 2048x1536 float scene", file=sys.stderr, flush=True)
    f32, target = scenes.float_scene(torch, 2048, 1536, [-4, -2, 0, 2, 4], seed=4)
    imgs = torch.stack([t.to(torch.float64) for t in f32]).contiguous()
    K, h, w, _ = imgs.shape
    L = enbl.auto_levels(w, h)
    ctx = enbl.Context(0)
    prm0 = torch.tensor([1.0, 0.2, 0.5, 0.5, 0.2], dtype=torch.float64)

    def loss_diff(i1, p1, i2, p2):
        """loss(i1, p1) - loss(i2, p2) = sum (O1 - O2)(O1 + O2 - 2 T), elementwise,
        so the difference of two nearby losses keeps its precision."""
        o1 = diff.forward(enbl, ctx, i1, p1, L)
        o2 = diff.forward(enbl, ctx, i2, p2, L)
        return float(torch.sum((o1 - o2) * (o1 + o2 - 2.0 * target)))

    out = diff.forward(enbl, ctx, imgs, prm0, L)
    t0 = time.perf_counter()
    for _ in range(3):
        out = diff.forward(enbl, ctx, imgs, prm0, L)
    fwd_ms = (time.perf_counter() - t0) * 1e3 / 3
    gout = 2.0 * (out - target)
    gi, gp = diff.backward(enbl, ctx, imgs, prm0, L, gout)
    t0 = time.perf_counter()
    for _ in range(3):
        gi, gp = diff.backward(enbl, ctx, imgs, prm0, L, gout)
    bwd_ms = (time.perf_counter() - t0) * 1e3 / 3
    gi_h, gp_h = gi.cpu().numpy(), gp.cpu().numpy()
    print(f"bench: finite differences on {fd_points} samples + 5 parameters", file=sys.stderr, flush=True)
    rng = np.random.default_rng(4)
    # Parameters: central differences at h = 1e-4, 5e-5, 2.5e-5, Richardson
    # extrapolated.  Image samples: Ridders' extrapolation of central
    # differences (a Neville tableau over steps shrinking by 1.4 from
    # h0 = 0.1 |v|, keeping the entry of smallest estimated error), which
    # balances the truncation error of the steep terms of dark pixels (the
    # local standard deviation of dark windows, their saturation) against the
    # rounding noise of the 47-Mpixel forward.  The saturation's max/min is
    # the one kink of the relaxed model: no step reaches another channel's
    # value (h0 <= d / 2, d the distance to the nearest other channel of the
    # pixel); a sample level with another channel (d = 0) is a kink, checked
    # against the matching one-sided derivative and listed apart with both
    # one-sided values as evidence.
    pscale = max(float(np.abs(gp_h).max()), 1e-300)
    iscale = float(np.abs(gi_h).max())
    floor_i = 1e-6 * iscale
    smooth, kinks = [], []

    def record(g, est, err, floor, what, extra=None):
        den = max(abs(est), floor)
        rec = {"what": what, "adjoint": float(g), "fd": float(est), "rel_err": abs(g - est) / den,
               "fd_err_estimate": abs(err) / den}
        if extra:
            rec.update(extra)
        smooth.append(rec)

    for q in range(5):
        fds = []
        for hs in (1e-4, 5e-5, 2.5e-5):
            p1, p2 = prm0.clone(), prm0.clone()
            p1[q] += hs
            p2[q] -= hs
            fds.append(loss_diff(imgs, p1, imgs, p2) / (2 * hs))
        r1, r2 = (4.0 * fds[1] - fds[0]) / 3.0, (4.0 * fds[2] - fds[1]) / 3.0
        record(gp_h[q], r2, r1 - r2, 1e-6 * pscale, f"param {q}")
    i1, i2 = imgs.clone(), imgs.clone()

    def central(k, y, x, c, old, hs):
        i1[k, y, x, c] = old + hs
        i2[k, y, x, c] = old - hs
        v = loss_diff(i1, prm0, i2, prm0) / (2 * hs)
        i1[k, y, x, c] = old
        i2[k, y, x, c] = old
        return v

    def ridders(f, h0, ntab=10, con=1.4):
        """Ridders' extrapolation of f(h) -> f(0) (f even in h): (estimate, error)."""
        a = [[0.0] * ntab for _ in range(ntab)]
        a[0][0] = f(h0)
        hh, err, ans = h0, float("inf"), a[0][0]
        for i in range(1, ntab):
            hh /= con
            a[0][i] = f(hh)
            fac = con * con
            for j in range(1, i + 1):
                a[j][i] = (a[j - 1][i] * fac - a[j - 1][i - 1]) / (fac - 1.0)
                fac *= con * con
                errt = max(abs(a[j][i] - a[j - 1][i]), abs(a[j][i] - a[j - 1][i - 1]))
                if errt <= err:
                    err, ans = errt, a[j][i]
            if abs(a[i][i] - a[i - 1][i - 1]) >= 2.0 * err:
                break
        return ans, err

    def one_sided(k, y, x, c, old, hs, sign):
        # (L(v + s h) - L(v)) / (s h), extrapolated over h, h/2 (first order)
        vals = []
        for hh in (hs, hs / 2):
            i1[k, y, x, c] = old + sign * hh
            vals.append(loss_diff(i1, prm0, imgs, prm0) / (sign * hh))
            i1[k, y, x, c] = old
        return 2.0 * vals[1] - vals[0]

    for _ in range(fd_points):
        k, y, x, c = (int(rng.integers(0, K)), int(rng.integers(0, h)), int(rng.integers(0, w)), int(rng.integers(0, 3)))
        old = float(imgs[k, y, x, c])
        pix = [float(imgs[k, y, x, o]) for o in range(3)]
        d = min(abs(old - pix[o]) for o in range(3) if o != c)
        what = f"image {k} ({y},{x}) ch {c} value {old:.3e}"
        g = float(gi_h[k, y, x, c])
        if d == 0.0:
            hs = 1e-3 * max(abs(old), 1e-6)
            left, right = one_sided(k, y, x, c, old, hs, -1.0), one_sided(k, y, x, c, old, hs, 1.0)
            side = min(abs(g - v) / max(abs(v), floor_i) for v in (left, right))
            kinks.append({"what": what, "adjoint": g, "fd_left": left, "fd_right": right, "rel_err_matching_side": side})
            continue
        h0 = min(0.1 * max(abs(old), 1e-6), d / 2.0)
        est, err = ridders(lambda hh: central(k, y, x, c, old, hh), h0)
        record(g, est, err, floor_i, what, {"h0": h0})
    # Directional derivatives along random directions over every sample of
    # every image: V = n I with one N(0,1) factor n per pixel for all three
    # channels, so a step scales each pixel and never reorders its channels
    # (no max/min kink is crossed), and the loss changes coherently over 47 M
    # samples, far above the rounding noise of the forward.  The full-size
    # adjoint is checked to the extrapolation's precision: <dL/dI, V> against
    # Ridders' central differences of L(I + e V).
    gen = torch.Generator(device=dev).manual_seed(44)
    directional = []
    gi_dev = gi.to(dev) if hasattr(gi, "to") else torch.from_numpy(gi_h).to(dev)
    for t in range(3):
        V = torch.randn(imgs.shape[:3] + (1,), generator=gen, dtype=torch.float64, device=dev) * imgs
        ad = float(torch.sum(gi_dev * V))

        def dirf(e):
            return loss_diff(imgs + e * V, prm0, imgs - e * V, prm0) / (2 * e)

        # (the loss is strongly curved along such scalings: central
        # differences at 1e-4 are off by up to 6e-4, so the steps start at 2e-5)
        est, err = ridders(dirf, 2e-5)
        directional.append({"adjoint": ad, "fd": est, "rel_err": abs(ad - est) / abs(est),
                            "fd_err_estimate": abs(err) / abs(est)})
    # The oracle's adjoint (oracle/oracle_diff.cpp, FP64 on the host) on the
    # same full-size inputs and output gradient: the CPU baseline of the
    # backward pass and its parity check.
    oracle_check = None
    if cpu:
        print("bench: oracle adjoint on the full workload", file=sys.stderr, flush=True)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_binding import load_oracle
        o = load_oracle()
        ih, gh = imgs.cpu().numpy(), gout.cpu().numpy()
        t0 = time.perf_counter()
        ogi, ogp = o.diff_backward(ih, prm0.numpy(), L, gh)
        osec = time.perf_counter() - t0
        big = float(np.abs(ogi).max())
        nz = np.abs(ogi) > 1e-6 * big
        oracle_check = {"gimg_rel_err_max": float(np.max(np.abs(gi_h[nz] - ogi[nz]) / np.abs(ogi[nz]))),
                        "gimg_abs_err_max_over_max": float(np.abs(gi_h - ogi).max() / big),
                        "gparam_rel_err_max": float(np.max(np.abs(gp_h - ogp) / np.abs(ogp))),
                        "cpu_baseline": {"value": round(osec * 1e3, 1), "unit": "ms (backward)",
                                         "cores": o.lib.orc_num_threads(), "kind": "port",
                                         "sample": "the full 5 x 2048x1536 backward, one run"}}
    print(f"bench: {adam_steps} Adam steps on the 5 parameters", file=sys.stderr, flush=True)
    prm = prm0.clone().to(dev).requires_grad_(True)
    opt = torch.optim.Adam([prm], lr=0.01)
    losses = []
    t0 = time.perf_counter()
    for _ in range(adam_steps):
        opt.zero_grad()
        o = diff.DiffEnfuse.apply(imgs, prm, enbl, ctx, L, 5)
        loss = torch.sum((o - target) ** 2)
        loss.backward()
        opt.step()
        with torch.no_grad():
            prm[:3].clamp_(min=0.0)
            prm[3].clamp_(0.0, 1.0)
            prm[4].clamp_(min=0.02)
        losses.append(float(loss.detach()))
    adam_s = time.perf_counter() - t0
    ctx.close()
    mpix = K * w * h / 1e6
    return {"workload": "differentiable enfuse (relaxed float variant), 5 x 2048x1536 RGB, EV -4..+4, FP64 pyramids, "
                        f"{L} levels, L2 loss to the Reinhard tone map",
            "adjoint_rel_err_max": max(r["rel_err"] for r in smooth),
            "adjoint_within_1e-4": sum(1 for r in smooth if r["rel_err"] <= 1e-4),
            "adjoint_within_1e-4_or_3x_fd_error": sum(1 for r in smooth
                                                       if r["rel_err"] <= max(1e-4, 3.0 * r["fd_err_estimate"])),
            "adjoint_rel_err_max_fd_resolved": max([r["rel_err"] for r in smooth if r["fd_err_estimate"] <= 1e-5],
                                                   default=0.0),
            "fd_resolved": sum(1 for r in smooth if r["fd_err_estimate"] <= 1e-5),
            "directional": directional, "vs_oracle_adjoint": oracle_check,
            "adjoint_rel_err_params": [r["rel_err"] for r in smooth if r["what"].startswith("param")],
            "fd_points": fd_points + 5, "fd_smooth": len(smooth), "fd_kinks": kinks,
            "fd_err_estimate_max": max(r["fd_err_estimate"] for r in smooth),
            "fd_worst": sorted(smooth, key=lambda r: -r["rel_err"])[:3],
            "fd": "central differences of the GPU's FP64 forward. Parameters: h = 1e-4, 5e-5, 2.5e-5, Richardson "
                  "extrapolated. Image samples (values down to 6e-5): Ridders' extrapolation of central differences "
                  "from h0 = 0.1 |value| down by 1.4 per step, the tableau entry of smallest estimated error, no "
                  "step reaching another channel's value (the saturation's max/min kink). adjoint_rel_err_max is "
                  "over every point that is not a kink; a sample equal to another channel of its pixel is a kink, "
                  "checked against the matching one-sided derivative and listed with both one-sided values. "
                  "Denominator floor 1e-6 x the largest gradient component",
            "forward_ms": round(fwd_ms, 3), "backward_ms": round(bwd_ms, 3),
            "forward_mpix_per_s": round(mpix / (fwd_ms * 1e-3), 1),
            "adam_steps": adam_steps, "adam_s": round(adam_s, 2), "loss_first": losses[0], "loss_last": losses[-1],
            "params_initial": prm0.tolist(), "params_final": [float(v) for v in prm.detach().cpu()]}


def _rel_err_float(got, want):
    """Largest |got - want| / |want| over the pixels, with the non-finite
    pattern required to match exactly and want == 0 requiring got == 0."""
    import numpy as np
    g = got.astype(np.float64)
    c = want.astype(np.float64)
    fin_c = np.isfinite(c)
    if not (np.array_equal(np.isnan(g), np.isnan(c)) and np.array_equal(np.isinf(g), np.isinf(c))
            and np.array_equal(g[np.isinf(c)], c[np.isinf(c)])):
        return float("inf")
    g, c = g[fin_c], c[fin_c]
    zero = c == 0
    if np.any(g[zero] != 0):
        return float("inf")
    nz = ~zero
    return float(np.max(np.abs(g[nz] - c[nz]) / np.abs(c[nz]))) if nz.any() else 0.0


def _time_enfuse(enbl, ctx, itype, imgs, alphas, prm, L, out, oa, steps, warmup):
    def run():
        enbl.enfuse(ctx, itype, imgs, alphas, prm, L, out, oa, dither=enbl.DITHER_DET, seed=1)

    for _ in range(warmup):
        run()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.profile(True)
    run()
    rep = ctx.profile_report()
    ctx.profile(False)
    return ms, rep


def enfuse_c0(args, enbl, scenes, np, torch, cpu=True):
    """Plain enfuse of the 3-exposure 1024x768 RGB uint8 stack (EV -2, 0, +2;
    exposure 1.0, saturation 0.2, contrast 0; 8 levels; Int16 fixed-point
    pyramids): GPU throughput, bit-exact parity with the oracle on the whole
    workload, and the oracle timed on the host cores."""
    ctx = enbl.Context(0)
    dev = torch.device("cuda:0")
    prm = enbl.enfuse_params(w_exposure=1.0, w_saturation=0.2)
    host = scenes.exposure_stack(1024, 768, 3, 0, bits=8)
    imgs = [torch.from_numpy(i).to(dev) for i in host]
    h, w = imgs[0].shape[0], imgs[0].shape[1]
    alphas = [torch.full((h, w), 255, dtype=torch.uint8, device=dev) for _ in imgs]
    out = torch.zeros_like(imgs[0])
    oa = torch.zeros((h, w), dtype=torch.uint8, device=dev)
    L = enbl.auto_levels(w, h)
    mpix = len(imgs) * w * h / 1e6
    torch.cuda.synchronize()
    print("bench: enfuse 3 x 1024x768 uint8", file=sys.stderr, flush=True)
    ms, rep = _time_enfuse(enbl, ctx, enbl.U8, imgs, alphas, prm, L, out, oa, max(args.steps, 20), args.warmup)
    got, got_a = out.cpu().numpy(), oa.cpu().numpy()
    ctx.close()
    res = {"workload": "enfuse, 3 x 1024x768 RGB uint8, EV -2/0/+2, exposure 1.0 + saturation 0.2, "
                       f"{L} levels, Int16 fixed-point pyramids",
           "input_mpix": round(mpix, 3), "ms_per_step": round(ms, 3), "mpix_per_s": round(mpix / (ms * 1e-3), 1),
           "kernels": _kernels(rep)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_binding import U8, enfuse_params, load_oracle
        o = load_oracle()
        oprm = enfuse_params(w_exposure=1.0, w_saturation=0.2)
        al = [np.full((h, w), 255, np.uint8) for _ in host]
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            want, want_a = o.enfuse(host, al, U8, oprm, L, dither=1, seed=1)
            times.append(time.perf_counter() - t0)
        times.sort()
        res["parity"] = bool(np.array_equal(got, want) and np.array_equal(got_a, want_a))
        res["cpu_baseline"] = {"value": round(mpix / times[2], 3), "unit": "Mpix/s", "cores": o.lib.orc_num_threads(),
                               "kind": "port", "sample": f"full workload, median of 5 runs ({times[2]:.3f} s each)"}
    return res


def enblend_layers_c1(args, enbl, scenes, np, torch, cpu=True):
    """The whole enblend run of the headline scene from decoded input layers:
    2 interleaved RGBA uint16 images of 4096x3072 at x = 0 and x = 3072 of a
    7168x3072 canvas (alpha bands 0 / 65535), one at a time (no
    pre-assembly), the fine NFT seam mask without optimiser, then the
    checkpoint conversion for a 16-bit output.  One step = enbl_enblend_layers
    (import with alpha thresholds, the two assemblies, overlap inspection,
    seam mask, mask bounds, ROI and levels, the blend step) + enbl_checkpoint;
    layers already resident in HBM.  Parity: the oracle's assemble + step +
    checkpoint on the same layers, bit for bit."""
    white, wa, black, ba, _ = scenes.enblend_pair(seed=1)
    H, W = wa.shape
    wimg = 4096
    off = W - wimg

    def rgba(img, a, x0):
        out = np.empty((H, wimg, 4), np.uint16)
        out[..., :3] = img[:, x0:x0 + wimg]
        out[..., 3] = a[:, x0:x0 + wimg].astype(np.uint16) * 257
        return out

    host_layers = [(rgba(black, ba, 0), 0, 0, enbl.U16), (rgba(white, wa, off), off, 0, enbl.U16)]
    dev = torch.device("cuda:0")
    dl = [(torch.from_numpy(a).to(dev), x, y, ft) for a, x, y, ft in host_layers]
    tab = enbl.layer_table(dl)
    ctx = enbl.Context(0)
    L = enbl.lib()
    d_b = torch.empty((H, W * 3), dtype=torch.uint16, device=dev)
    d_ba = torch.empty((H, W), dtype=torch.uint8, device=dev)
    d_out = torch.empty((H, W * 4), dtype=torch.uint16, device=dev)
    bb = enbl.Rect()
    counts = (ctypes.c_int * 3)()
    ct = ctypes.c_int(-1)
    vp = ctypes.c_void_p

    def run():
        enbl.check(L.enbl_enblend_layers(ctx.h, enbl.U16, 3, W, H, 2, tab, 0, 0, 0, 1, 2, enbl.DITHER_DET, 1,
                                         vp(d_b.data_ptr()), W * 3, vp(d_ba.data_ptr()), W, ctypes.byref(bb),
                                         counts))
        enbl.check(L.enbl_checkpoint(ctx.h, enbl.U16, 3, W, H, vp(d_b.data_ptr()), W * 3, vp(d_ba.data_ptr()), W,
                                     enbl.U16, vp(d_out.data_ptr()), W * 4, ctypes.byref(ct)))

    print("bench: enblend from decoded layers (assembly, seam mask, blend, checkpoint)", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    run()
    ctx.sync()
    got = d_out.cpu().numpy()
    for _ in range(args.warmup):
        run()
    ctx.sync()
    steps = max(5, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.profile(True)
    run()
    rep = ctx.profile_report()
    ctx.profile(False)
    ctx.close()
    mpix = 2 * wimg * H / 1e6
    res = {"workload": "enblend of 2 decoded RGBA uint16 layers 4096x3072 on a 7168x3072 canvas: import + "
                       "assembly, overlap, NFT seam mask, mask bounds, ROI, 10-level blend, 16-bit checkpoint",
           "input_mpix": round(mpix, 3), "ms_per_step": round(ms, 3), "mpix_per_s": round(mpix / (ms * 1e-3), 1),
           "steps": steps, "blend_steps": list(counts), "union_box": bb.tuple(), "kernels": _kernels(rep)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_binding import U16, load_oracle
        o = load_oracle()
        t0 = time.perf_counter()
        taken = [0, 0]
        _, bimg, bal, bbb, taken = o.assemble(host_layers, U16, 3, W, H, taken, True)
        _, wimg_, wal, wbb, taken = o.assemble(host_layers, U16, 3, W, H, taken, True)
        rc, _, bimg, bal, bbb, _ = o.enblend_step(bimg, bal, bbb, wimg_, wal, wbb, U16, seed=1)
        want, _ = o.checkpoint(bimg, bal, U16, U16)
        cpu_s = time.perf_counter() - t0
        res["parity"] = bool(rc == 0 and np.array_equal(got.reshape(H, W, 4), want))
        res["cpu_baseline"] = {"value": round(mpix / cpu_s, 3), "unit": "Mpix/s", "cores": o.lib.orc_num_threads(),
                               "kind": "port", "sample": f"full workload, one run ({cpu_s:.2f} s)"}
    return res


C2_EVS = [k - 3.5 for k in range(8)]


def c2_params(mod):
    return mod.enfuse_params(w_exposure=1.0, w_saturation=0.2, w_contrast=0.5, w_entropy=0.3,
                             contrast_window=5, entropy_window=7)


C2_WORKLOAD = ("enfuse, 8 x 6000x4000 RGB float32, EV -3.5..+3.5 one EV apart, exposure 1.0 + saturation 0.2 + "
               "contrast 0.5 (5x5 local std-dev) + entropy 0.3 (7x7), 10 levels, double pyramids")
C2_TRAFFIC = "profiles/enfuse_f32_6000x4000x8_traffic.md"
C1_TRAFFIC = "profiles/enblend_pair_u16_4096x3072x2_traffic.md"
# Kernels of the pyramid build, merge and collapse (the reduce / expand
# passes and the passes that fuse them) as the library's profile names them.
PYRAMID_KERNEL_PREFIXES = ("reduce", "expand", "laplacian", "collapse", "merge", "tail", "to_pyramid", "from_pyramid")


def _roofline(rep, runs, traffic_file):
    """Roofline object for the dominant kernel (largest total time) of a
    profiled pass: its own algorithmic bytes per launch (each input read once,
    each output written once) over its mean HIP-event launch duration."""
    total_ms = sum(v["ms"] for v in rep.values())
    dom_name, dom = max(rep.items(), key=lambda kv: kv[1]["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    avg_own = dom["own_bytes"] / dom["launches"]
    achieved = avg_own / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom_name, traffic_file)
    return {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "achieved_basis": "the kernel's own algorithmic bytes per launch (each input read once, each output "
                              "written once) over its mean HIP-event launch duration on the library's stream",
            "algorithmic_bytes_per_launch": round(avg_own), "launch_ms": round(avg_ms, 5),
            "launches_per_step": dom["launches"] // runs,
            "traffic_source": traffic_file if traffic else None,
            "dram_GB_per_s": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
            "kernel_ms_per_step": round(total_ms / runs, 4)}


def _pyramid_kernels(rep, runs):
    """HBM fraction of the pyramid kernels together: their own bytes over
    their summed HIP-event time."""
    sel = {k: v for k, v in rep.items() if k.startswith(PYRAMID_KERNEL_PREFIXES)}
    ms = sum(v["ms"] for v in sel.values())
    own = sum(v["own_bytes"] for v in sel.values())
    gbs = own / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"kernels": sorted(sel), "ms_per_step": round(ms / runs, 4), "own_bytes_per_step": round(own / runs),
            "GB_per_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def enfuse_c2(args, enbl, scenes, np, torch, cpu=True):
    """The headline: plain enfuse of 8 float32 RGB exposures at 6000x4000
    (radiance 1e-3..1e2, EV -3.5..+3.5 one EV apart, no clipping); weights
    exposure 1.0, saturation 0.2, contrast 0.5 (5x5 local standard
    deviation), entropy 0.3 (7x7 window, per-channel exact-value histograms
    as the reference computes them); 10 levels of double pyramids; full alpha.
    One step = enbl_enfuse on the 8 device-resident exposures (weights,
    normalisation, the 8 Laplacian and mask pyramids, the merge of every
    level, collapse and the conversion into the float output).  Also: the same
    step with host-resident inputs and outputs (pinned buffers, PCIe
    included), and the oracle on the full workload for the CPU baseline and
    the parity check (2e-6 relative, non-finite pattern equal)."""
    dev = torch.device("cuda:0")
    W, H = 6000, 4000
    print("bench: generating the 8 x 6000x4000 float scene", file=sys.stderr, flush=True)
    imgs, target = scenes.float_scene(torch, W, H, C2_EVS, seed=2)
    del target
    alphas = [torch.full((H, W), 255, dtype=torch.uint8, device=dev) for _ in imgs]
    out = torch.zeros_like(imgs[0])
    oa = torch.zeros((H, W), dtype=torch.uint8, device=dev)
    prm = c2_params(enbl)
    L = enbl.auto_levels(W, H)
    mpix = len(imgs) * W * H / 1e6
    torch.cuda.synchronize()
    ctx = enbl.Context(0)

    def run():
        enbl.enfuse(ctx, enbl.F32, imgs, alphas, prm, L, out, oa, dither=enbl.DITHER_DET, seed=1)

    print(f"bench: enfuse 8 x 6000x4000 float32: {args.warmup} warmup + {args.steps} timed steps", file=sys.stderr,
          flush=True)
    run()
    ctx.sync()
    got = out.cpu().numpy()
    got_a = oa.cpu().numpy()
    for _ in range(args.warmup):
        run()
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    ctx.sync()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    nprof = 2
    ctx.profile(True)
    for _ in range(nprof):
        run()
    rep = ctx.profile_report()
    ctx.profile(False)
    res = {
        "metric": METRIC,
        "value": round(mpix / (ms * 1e-3), 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": C2_WORKLOAD, "input_mpix_per_step": mpix, "levels": L,
                   "inputs": "resident in HBM when the timed region starts"},
        "roofline": _roofline(rep, nprof, C2_TRAFFIC),
        "pyramid_kernels": _pyramid_kernels(rep, nprof),
        "kernels": {k: {"ms_per_step": round(v["ms"] / nprof, 4), "launches_per_step": v["launches"] // nprof,
                        "GB_per_s": round(v["own_bytes"] / (v["ms"] * 1e-3) / 1e9, 1)}
                    for k, v in sorted(rep.items())},
    }
    if not args.no_host:
        # Host-resident inputs and outputs: the 8 exposures and alphas uploaded
        # from pinned host memory, the step, the output image and alpha
        # downloaded (what a caller handing over host buffers sees).
        print("bench: host-resident (PCIe-inclusive) steps", file=sys.stderr, flush=True)
        h_in = [t.cpu().pin_memory() for t in imgs + alphas]
        h_out = [torch.empty(t.shape, dtype=t.dtype).pin_memory() for t in (out, oa)]
        d_in = imgs + alphas

        def host_step():
            for d, hh in zip(d_in, h_in):
                d.copy_(hh, non_blocking=True)
            torch.cuda.synchronize()
            run()
            ctx.sync()
            for hh, d in zip(h_out, (out, oa)):
                hh.copy_(d, non_blocking=True)
            torch.cuda.synchronize()

        host_step()
        hruns = 3
        t0 = time.perf_counter()
        for _ in range(hruns):
            host_step()
        hms = (time.perf_counter() - t0) * 1e3 / hruns
        hbytes = sum(t.numel() * t.element_size() for t in h_in + h_out)
        res["host_resident"] = {"ms_per_step": round(hms, 3), "mpix_per_s": round(mpix / (hms * 1e-3), 1),
                                "host_bytes_per_step": hbytes, "runs": hruns,
                                "note": "pinned host buffers in and out around the same step (PCIe included)"}
        del h_in, h_out
    host = [t.cpu().numpy() for t in imgs]
    ctx.close()
    del imgs, alphas, out, oa
    torch.cuda.empty_cache()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_binding import F32, load_oracle
        import oracle_binding
        o = load_oracle()
        al = [np.full((H, W), 255, np.uint8) for _ in host]
        print("bench: oracle enfuse on the full 8 x 6000x4000 workload", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        want, want_a = o.enfuse(host, al, F32, c2_params(oracle_binding), L, dither=1, seed=1)
        cpu_s = time.perf_counter() - t0
        err = _rel_err_float(got, want)
        res["parity_rel_err_max"] = err
        res["parity"] = bool(err <= 2e-6 and np.array_equal(got_a, want_a))
        res["parity_sample"] = "the full workload; non-finite pattern must match"
        res["cpu_baseline"] = {"value": round(mpix / cpu_s, 3), "unit": "Mpix/s", "cores": o.lib.orc_num_threads(),
                               "kind": "port", "sample": f"the full 8 x 6000x4000 workload, one run ({cpu_s:.1f} s)"}
    return res


def enblend_c1(args, enbl, scenes, np, torch, cpu=True):
    """One enblend blend step of two 4096x3072 RGBA uint16 images sharing one
    synthetic scene, 25% horizontal overlap, on their 7168x3072 union canvas
    with 10 pyramid levels (src/enblend.h:378-510): mask Gaussian pyramid,
    two Laplacian pyramids under the images' alphas, blend, collapse and the
    dithered copy back into the 16-bit image.  The seam mask is the
    nearest-feature transform of the two alphas (`nft`, timed apart and
    added in `step_with_nft_ms`).  Every timed step blends its own staged copy
    of the black image and alpha (the step updates them in place), so each
    one sees the stated inputs."""
    print("bench: generating the 2 x 4096x3072 RGBA u16 scene", file=sys.stderr, flush=True)
    white, wa, black, ba, _ = scenes.enblend_pair(seed=1)
    h, w, C = white.shape
    levels = enbl.auto_levels(w, h)
    mpix = 2 * 4096 * 3072 / 1e6
    dev = torch.device("cuda:0")
    d_white = torch.from_numpy(white).to(dev)
    d_wa = torch.from_numpy(wa).to(dev)
    d_black0 = torch.from_numpy(black).to(dev)
    d_ba0 = torch.from_numpy(ba).to(dev)
    d_mask = torch.empty((h, w), dtype=torch.uint8, device=dev)
    steps = max(args.steps, 5)
    nprof = 3
    # staged pristine copies: parity, warmup, timed and profiled steps
    ncopies = 1 + args.warmup + steps + nprof
    stage = [(d_black0.clone(), d_ba0.clone()) for _ in range(ncopies)]
    torch.cuda.synchronize()
    ctx = enbl.Context(0)
    L = enbl.lib()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())

    print("bench: nearest-feature-transform seam mask", file=sys.stderr, flush=True)

    def nft():
        enbl.check(L.enbl_nft_mask(ctx.h, w, h, vp(d_wa), w, vp(d_ba0), w, 2, vp(d_mask), w))

    nft()
    ctx.sync()
    nft_runs = 5
    t0 = time.perf_counter()
    for _ in range(nft_runs):
        nft()
    ctx.sync()
    nft_ms = (time.perf_counter() - t0) * 1e3 / nft_runs
    ctx.profile(True)
    nft()
    nft_rep = ctx.profile_report()
    ctx.profile(False)
    mask = d_mask.cpu().numpy()

    def step(bk, bka):
        enbl.check(L.enbl_enblend_pair(ctx.h, enbl.U16, C, w, h, levels, 0,
                                       vp(d_white), w * C, vp(d_wa), w, vp(bk), w * C, vp(bka), w,
                                       vp(d_mask), w, enbl.DITHER_DET, 1))

    print("bench: enblend step for parity", file=sys.stderr, flush=True)
    it = iter(stage)
    b, bal = next(it)
    step(b, bal)
    ctx.sync()
    gpu_out, gpu_alpha = b.cpu().numpy(), bal.cpu().numpy()
    for _ in range(args.warmup):
        step(*next(it))
    ctx.sync()
    torch.cuda.synchronize()
    print(f"bench: timing {steps} enblend steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        step(*next(it))
    ctx.sync()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.profile(True)
    for _ in range(nprof):
        step(*next(it))
    rep = ctx.profile_report()
    ctx.profile(False)
    res = {"workload": "enblend step, 2 x 4096x3072 RGBA uint16, 25% overlap, 7168x3072 ROI, NFT seam mask, "
                       f"{levels} levels, Int32 fixed-point pyramids",
           "input_mpix_per_step": mpix, "steps": steps, "ms_per_step": round(ms, 4),
           "mpix_per_s": round(mpix / (ms * 1e-3), 1),
           "step_with_nft_ms": round(ms + nft_ms, 4),
           "step_with_nft_mpix_per_s": round(mpix / ((ms + nft_ms) * 1e-3), 1),
           "roofline": _roofline(rep, nprof, C1_TRAFFIC),
           "pyramid_kernels": _pyramid_kernels(rep, nprof),
           "kernels": {k: {"ms_per_step": round(v["ms"] / nprof, 4), "launches_per_step": v["launches"] // nprof,
                           "GB_per_s": round(v["own_bytes"] / (v["ms"] * 1e-3) / 1e9, 1)}
                       for k, v in sorted(rep.items())},
           "nft": {"workload": f"nearest-feature-transform seam mask of the two alphas, {w}x{h}, Euclidean, "
                               "float Felzenszwalb-Huttenlocher transforms", "ms": round(nft_ms, 3),
                   "mpix_per_s": round(w * h / 1e6 / (nft_ms * 1e-3), 1), "runs": nft_runs,
                   "kernels": _kernels(nft_rep)}}
    if not args.no_host:
        print("bench: PCIe-inclusive enblend steps", file=sys.stderr, flush=True)
        d_black, d_ba = stage[0]
        h_in = [torch.from_numpy(a).pin_memory() for a in (white, wa, black, ba, mask)]
        h_out = [torch.empty(t.shape, dtype=t.dtype).pin_memory() for t in (d_black, d_ba)]
        d_in = [d_white, d_wa, d_black, d_ba, d_mask]

        def pcie_step():
            for d, hh in zip(d_in, h_in):
                d.copy_(hh, non_blocking=True)
            torch.cuda.synchronize()
            step(d_black, d_ba)
            ctx.sync()
            for hh, d in zip(h_out, (d_black, d_ba)):
                hh.copy_(d, non_blocking=True)
            torch.cuda.synchronize()

        pcie_step()
        pcie_runs = 5
        t0 = time.perf_counter()
        for _ in range(pcie_runs):
            pcie_step()
        pcie_ms = (time.perf_counter() - t0) * 1e3 / pcie_runs
        res["host_resident"] = {"ms_per_step": round(pcie_ms, 3), "mpix_per_s": round(mpix / (pcie_ms * 1e-3), 1),
                                "host_bytes_per_step": sum(t.numel() * t.element_size() for t in h_in + h_out),
                                "runs": pcie_runs,
                                "note": "pinned host buffers in and out around the same step (PCIe included)"}
    ctx.close()
    del stage, d_white, d_wa, d_black0, d_ba0, d_mask
    torch.cuda.empty_cache()
    if cpu:
        print("bench: CPU baseline of the enblend step (oracle)", file=sys.stderr, flush=True)
        cpu_s, threads, (cpu_out, cpu_alpha) = cpu_baseline(white, wa, black, ba, mask, levels, args.cpu_repeats)
        res["cpu_baseline"] = {"value": round(mpix / cpu_s, 3), "unit": "Mpix/s", "cores": threads, "kind": "port",
                               "sample": f"full workload, median of {args.cpu_repeats} runs ({cpu_s:.2f} s each)"}
        res["parity"] = bool(np.array_equal(cpu_out, gpu_out) and np.array_equal(cpu_alpha, gpu_alpha))
        from oracle_binding import load_oracle
        o = load_oracle()
        t0 = time.perf_counter()
        cpu_mask, cpu_rc = o.nft_mask(wa, ba)
        nft_cpu_s = time.perf_counter() - t0
        res["nft"]["parity_vs_oracle"] = bool(cpu_rc == 0 and np.array_equal(cpu_mask, mask))
        res["nft"]["cpu_baseline"] = {"ms": round(nft_cpu_s * 1e3, 1), "cores": o.lib.orc_num_threads(),
                                      "kind": "port", "sample": "full mask, one run"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="leave out the oracle (CPU baseline and parity)")
    ap.add_argument("--no-host", action="store_true", help="leave out the host-resident (PCIe-inclusive) steps")
    ap.add_argument("--workload", default="enfuse-c2",
                    choices=["enfuse-c2", "enblend-c1", "enblend-layers", "enfuse-c0", "batch-c3", "adjoint-c4",
                             "copy-calibration"],
                    help="enfuse-c2: the headline line with every part (default); the others print one part alone")
    ap.add_argument("--no-c1", action="store_true", help="leave the enblend step out of the line")
    ap.add_argument("--no-enfuse", action="store_true", help="leave the small enfuse workload out of the line")
    ap.add_argument("--no-layers", action="store_true",
                    help="leave the enblend run from decoded layers (assembly through checkpoint) out of the line")
    ap.add_argument("--no-batch", action="store_true", help="leave the batched-vs-plain comparison out of the line")
    ap.add_argument("--no-adjoint", action="store_true", help="leave the adjoint check and tuning run out of the line")
    ap.add_argument("--headline-only", action="store_true", help="the headline workload alone")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--tuning", default=os.path.join(ROOT, "enblend-enfuse_amd", "batch_tuning.json"),
                    help="launch shapes of the batched passes (JSON written by --autotune)")
    ap.add_argument("--autotune", default=None, help="run the autotuner on the batch first and write its JSON here")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    enbl = _load("enblend_enfuse_amd", os.path.join(ROOT, "enblend-enfuse_amd", "__init__.py"))
    scenes = _load("enblend_scenes", os.path.join(ROOT, "enblend-enfuse_amd", "scenes.py"))
    cpu = not args.no_cpu
    if args.workload == "enblend-c1":
        print(json.dumps(enblend_c1(args, enbl, scenes, np, torch, cpu=cpu)), flush=True)
        return
    if args.workload == "enblend-layers":
        print(json.dumps(enblend_layers_c1(args, enbl, scenes, np, torch, cpu=cpu)), flush=True)
        return
    if args.workload == "enfuse-c0":
        print(json.dumps(enfuse_c0(args, enbl, scenes, np, torch, cpu=cpu)), flush=True)
        return
    if args.workload == "batch-c3":
        print(json.dumps(batch_compare(args, enbl, scenes, torch, args.batch_steps, 1)), flush=True)
        return
    if args.workload == "adjoint-c4":
        print(json.dumps(adjoint_c4(args, enbl, scenes, torch, cpu=cpu)), flush=True)
        return
    if args.workload == "copy-calibration":
        # A plain device copy of known size: the reference reading for the
        # FETCH_SIZE / WRITE_SIZE counters that scale kernel traffic.
        n = 1 << 30
        a = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(args.steps):
            b.copy_(a)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / args.steps
        print(json.dumps({"workload": "copy-calibration", "bytes_read_per_copy": n, "bytes_written_per_copy": n,
                          "copies": args.steps + 1, "ms_per_copy": round(ms, 4),
                          "GB_per_s": round(2 * n / (ms * 1e-3) / 1e9, 1)}), flush=True)
        return
    result = enfuse_c2(args, enbl, scenes, np, torch, cpu=cpu)
    if not args.headline_only:
        if not args.no_c1:
            result["enblend_c1"] = enblend_c1(args, enbl, scenes, np, torch, cpu=cpu)
        if not args.no_layers:
            result["enblend_layers"] = enblend_layers_c1(args, enbl, scenes, np, torch, cpu=cpu)
        if not args.no_batch:
            result["batch"] = batch_compare(args, enbl, scenes, torch, args.batch_steps, 1)
        if not args.no_enfuse:
            result["enfuse_c0"] = enfuse_c0(args, enbl, scenes, np, torch, cpu=cpu)
        if not args.no_adjoint:
            result["adjoint"] = adjoint_c4(args, enbl, scenes, torch, cpu=cpu)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
