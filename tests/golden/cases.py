"""Small seeded cases of the pyramid path whose oracle outputs are committed
as golden digests (golden.json, and the float output as a .npy array).

Inputs are regenerated from numpy PCG64 seeds; their SHA-256 digests are part
of the golden record, so a change of the generator shows up as an input
mismatch rather than as a silent change of the expected outputs."""
import hashlib

import numpy as np

U8, U16, F32 = 10, 11, 12
NP_TYPE = {U8: np.uint8, U16: np.uint16, F32: np.float32}


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode() + str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def _smooth(rng, h, w, itype, k, K):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    base = 0.5 + 0.4 * np.sin(xx / rng.uniform(5, 11)) * np.cos(yy / rng.uniform(4, 9))
    rad = base[..., None] * rng.uniform(0.7, 1.3, size=3)[None, None, :] * 2.0 ** ((k - (K - 1) / 2) * 1.5)
    if itype == F32:
        return (rad + rng.normal(0, 1e-3, rad.shape)).astype(np.float32)
    m = 255 if itype == U8 else 65535
    v = np.clip(rad, 0, 1) ** (1 / 2.2) * m + rng.normal(0, 1, rad.shape)
    return np.clip(np.rint(v), 0, m).astype(NP_TYPE[itype])


def _ellipse_alpha(h, w, cx, cy, rx, ry):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    return np.where(((xx - cx) / rx) ** 2 + ((yy - cy) / ry) ** 2 <= 1.0, 255, 0).astype(np.uint8)


def enblend_inputs(seed, itype, w, h):
    """white, white alpha, black, black alpha: two overlapping ellipses."""
    rng = np.random.default_rng(seed)
    white = _smooth(rng, h, w, itype, 0, 1)
    black = _smooth(rng, h, w, itype, 0, 1)
    wa = _ellipse_alpha(h, w, 0.62 * w, 0.5 * h, 0.38 * w, 0.45 * h)
    ba = _ellipse_alpha(h, w, 0.38 * w, 0.5 * h, 0.38 * w, 0.47 * h)
    return white, wa, black, ba


def enfuse_inputs(seed, itype, K, w, h, holes):
    rng = np.random.default_rng(seed)
    imgs = [_smooth(rng, h, w, itype, k, K) for k in range(K)]
    alphas = [np.where(rng.random((h, w)) < (0.9 if holes else 2.0), 255, 0).astype(np.uint8) for _ in range(K)]
    return imgs, alphas


ENBLEND = [
    dict(name="enblend_u16_160x96", seed=101, itype=U16, w=160, h=96, wrap=False),
    dict(name="enblend_u8_wrap_150x90", seed=102, itype=U8, w=150, h=90, wrap=True),
]

ENFUSE = [
    dict(name="enfuse_u8_k3", seed=201, itype=U8, K=3, w=96, h=64, holes=False,
         prm=dict(w_exposure=1.0, w_saturation=0.2)),
    dict(name="enfuse_u16_k4_contrast_entropy", seed=202, itype=U16, K=4, w=80, h=72, holes=True,
         prm=dict(w_exposure=1.0, w_saturation=0.2, w_contrast=0.5, w_entropy=0.3, contrast_window=5,
                  entropy_window=3)),
    dict(name="enfuse_u8_k3_hard_mask", seed=203, itype=U8, K=3, w=90, h=70, holes=True,
         prm=dict(w_exposure=1.0, w_saturation=0.2, hard_mask=True)),
    dict(name="enfuse_f32_k3_contrast_entropy", seed=204, itype=F32, K=3, w=64, h=48, holes=False,
         prm=dict(w_exposure=1.0, w_saturation=0.2, w_contrast=0.5, w_entropy=0.3, contrast_window=5,
                  entropy_window=7)),
]

NFT = [dict(name="nft_mask_200x120", seed=301, w=200, h=120)]

DITHER_DET = 1
SEED = 7
