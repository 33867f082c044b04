"""Deterministic synthetic inputs for the benchmark and parity harness.

Scene model: base radiance 0.2 + a linear gradient + 40 isotropic Gaussian
blobs (sigma 5-80 px, amplitude U(0,2)) + 20 hard-edged rectangles (value
U(0,1.5)), per-channel tint U(0.7,1.3).  Integer exposures are 2^EV times the
radiance, gamma 1/2.2, clipped to [0,1], quantised, plus Gaussian noise of
sigma 1 LSB.  Everything derives from numpy's PCG64 seeded by (seed, purpose),
so every machine regenerates identical arrays.
"""
import numpy as np


def _rng(seed, purpose):
    return np.random.default_rng([int(seed), int(purpose)])


def radiance(w, h, seed, rng_purpose=0):
    """(h, w, 3) float32 radiance."""
    rng = _rng(seed, rng_purpose)
    g = rng.uniform(0.2, 0.8)
    base = (0.2 + g * (np.arange(w, dtype=np.float32) / max(w - 1, 1)))[None, :].repeat(h, 0)
    r = base.astype(np.float32)
    for _ in range(40):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        s = rng.uniform(5, 80)
        amp = rng.uniform(0, 2)
        x0, x1 = int(max(0, cx - 4 * s)), int(min(w, cx + 4 * s + 1))
        y0, y1 = int(max(0, cy - 4 * s)), int(min(h, cy + 4 * s + 1))
        if x0 >= x1 or y0 >= y1:
            continue
        xs = (np.arange(x0, x1, dtype=np.float32) - cx) ** 2
        ys = (np.arange(y0, y1, dtype=np.float32) - cy) ** 2
        r[y0:y1, x0:x1] += amp * np.exp(-(ys[:, None] + xs[None, :]) / (2 * s * s)).astype(np.float32)
    for _ in range(20):
        x0, x1 = sorted(rng.integers(0, w, size=2))
        y0, y1 = sorted(rng.integers(0, h, size=2))
        r[y0:y1 + 1, x0:x1 + 1] = rng.uniform(0, 1.5)
    tint = rng.uniform(0.7, 1.3, size=3).astype(np.float32)
    return r[:, :, None] * tint[None, None, :]


def expose_int(rad, ev, bits, seed, purpose):
    """Integer exposure of a radiance map: gamma 1/2.2, clip, quantise, noise."""
    rng = _rng(seed, purpose)
    m = (1 << bits) - 1
    v = np.clip(rad * (2.0 ** ev), 0, 1) ** (1 / 2.2) * m
    v = v + rng.standard_normal(v.shape, dtype=np.float32)
    v = np.clip(np.rint(v), 0, m)
    return v.astype(np.uint8 if bits == 8 else np.uint16)


def convex_polygon_alpha(w, h, rng, x0, x1):
    """0/255 alpha of a random convex polygon (hull of 8-16 points) inside
    columns [x0, x1)."""
    n = int(rng.integers(8, 17))
    cx, cy = (x0 + x1) / 2, h / 2
    ang = np.sort(rng.uniform(0, 2 * np.pi, size=n))
    rx, ry = (x1 - x0) * rng.uniform(0.45, 0.6), h * rng.uniform(0.45, 0.6)
    px, py = cx + rx * np.cos(ang), cy + ry * np.sin(ang)
    inside = np.ones((h, w), dtype=bool)
    yy = np.arange(h, dtype=np.float64)[:, None]
    xx = np.arange(w, dtype=np.float64)[None, :]
    for k in range(n):
        ax, ay = px[k], py[k]
        bx, by = px[(k + 1) % n], py[(k + 1) % n]
        inside &= (bx - ax) * (yy - ay) - (by - ay) * (xx - ax) >= 0
    inside[:, :x0] = False
    inside[:, x1:] = False
    return np.where(inside, 255, 0).astype(np.uint8)


def enblend_pair(seed=1, w_img=4096, h=3072, overlap=0.25, bits=16):
    """Two RGB images sharing one scene on a union canvas with polygon alphas
    and a 0/255 seam mask.  Returns (white, white_alpha, black, black_alpha,
    mask) on the canvas; black is the left image, white the right one.

    The seam mask stands in for the nearest-feature-transform mask the
    reference derives from the two alphas: the seam runs down the middle of
    the overlap with a smooth random wiggle (the mask is an input of the
    blend, so any 0/255 mask exercises the same path)."""
    off = int(round(w_img * (1 - overlap)))
    W = off + w_img
    rng = _rng(seed, 7)
    rad = radiance(W, h, seed)
    black = expose_int(rad[:, :w_img], 0.0, bits, seed, 11)
    gain = 1.0 + rng.uniform(-0.02, 0.02)
    white = expose_int(rad[:, off:] * gain, 0.0, bits, seed, 12)
    dt = black.dtype
    B = np.zeros((h, W, 3), dt)
    Wt = np.zeros((h, W, 3), dt)
    B[:, :w_img] = black
    Wt[:, off:] = white
    ba = convex_polygon_alpha(W, h, rng, 0, w_img)
    wa = convex_polygon_alpha(W, h, rng, off, W)
    mid = (off + w_img) / 2
    yy = np.arange(h)
    seam = mid + 0.1 * (w_img - off) * np.sin(2 * np.pi * yy / h * rng.uniform(1, 3) + rng.uniform(0, 6.3))
    mask = np.where(np.arange(W)[None, :] >= seam[:, None], 255, 0).astype(np.uint8)
    return Wt, wa, B, ba, mask


def exposure_stack(w, h, K, seed, bits=8, ev_step=2.0):
    """K integer exposures (2^EV times one radiance, EV centred on 0),
    e.g. the 3-exposure 1024x768 uint8 stack: exposure_stack(1024, 768, 3, 0)."""
    rad = radiance(w, h, seed)
    evs = [(k - (K - 1) / 2) * ev_step for k in range(K)]
    return [expose_int(rad, ev, bits, seed, 100 + k) for k, ev in enumerate(evs)]


BATCH_SIZES = [(640, 480), (1280, 720), (1920, 1080), (2448, 2048), (3840, 2160), (4096, 3072)]


def batch_plan(n_stacks=64, seed=3):
    """Geometry of the batched workload: per stack (w, h, K, EV spacing)."""
    rng = _rng(seed, 9)
    plan = []
    for s in range(n_stacks):
        w, h = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
        K = int(rng.integers(3, 10))
        plan.append((w, h, K, float(rng.uniform(1.0, 2.0))))
    return plan


def device_stack(torch, w, h, K, seed, ev_step, bits=16, device="cuda"):
    """The same scene model generated on the device (torch), for workloads
    too large to synthesise on the host in bounded time.  Returns K (h, w, 3)
    tensors of uint16 (bits=16) or uint8."""
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) * 7919 + 13)

    def U(a, b, n=1):
        return torch.rand(n, generator=g, device=device, dtype=torch.float32) * (b - a) + a

    ys = torch.arange(h, device=device, dtype=torch.float32)[:, None]
    xs = torch.arange(w, device=device, dtype=torch.float32)[None, :]
    r = 0.2 + U(0.2, 0.8) * xs / max(w - 1, 1) + 0.0 * ys
    for _ in range(40):
        cx, cy = U(0, w), U(0, h)
        s = U(5, 80)
        r = r + U(0, 2) * torch.exp(-((xs - cx) ** 2 + (ys - cy) ** 2) / (2 * s * s))
    for _ in range(20):
        x0, x1 = sorted(int(v) for v in (U(0, w, 2).tolist()))
        y0, y1 = sorted(int(v) for v in (U(0, h, 2).tolist()))
        r[y0:y1 + 1, x0:x1 + 1] = U(0, 1.5)
    tint = U(0.7, 1.3, 3)
    rad = r[:, :, None] * tint[None, None, :]
    m = (1 << bits) - 1
    out = []
    for k in range(K):
        ev = (k - (K - 1) / 2) * ev_step
        v = torch.clamp(rad * (2.0 ** ev), 0, 1) ** (1 / 2.2) * m
        v = v + torch.randn(v.shape, generator=g, device=device)
        v = torch.clamp(torch.round(v), 0, m)
        out.append(v.to(torch.int32).to(torch.uint16 if bits == 16 else torch.uint8).contiguous())
    return out


def float_scene(torch, w, h, evs, seed, device="cuda"):
    """Float exposures of one device-generated scene: radiance remapped
    log-uniformly to [1e-3, 1e2], exposures R * 2^EV as linear float32 with no
    clipping, and the Reinhard global tone map R / (1 + R) of the radiance.
    Returns (list of (h, w, 3) float32 tensors, target (h, w, 3) float64)."""
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) * 7919 + 17)

    def U(a, b, n=1):
        return torch.rand(n, generator=g, device=device, dtype=torch.float64) * (b - a) + a

    ys = torch.arange(h, device=device, dtype=torch.float64)[:, None]
    xs = torch.arange(w, device=device, dtype=torch.float64)[None, :]
    r = 0.2 + U(0.2, 0.8) * xs / max(w - 1, 1) + 0.0 * ys
    for _ in range(40):
        cx, cy = U(0, w), U(0, h)
        s = U(5, 80)
        r = r + U(0, 2) * torch.exp(-((xs - cx) ** 2 + (ys - cy) ** 2) / (2 * s * s))
    for _ in range(20):
        x0, x1 = sorted(int(v) for v in (U(0, w, 2).tolist()))
        y0, y1 = sorted(int(v) for v in (U(0, h, 2).tolist()))
        r[y0:y1 + 1, x0:x1 + 1] = U(0, 1.5)
    rad = r[:, :, None] * U(0.7, 1.3, 3)[None, None, :]
    lo, hi = float(rad.min()), float(rad.max())
    rad = 1e-3 * torch.pow(torch.tensor(1e5, dtype=torch.float64, device=device), (rad - lo) / max(hi - lo, 1e-12))
    imgs = [(rad * (2.0 ** ev)).to(torch.float32).contiguous() for ev in evs]
    return imgs, (rad / (1.0 + rad)).contiguous()
