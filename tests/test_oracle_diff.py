"""The FP64 oracle of the relaxed differentiable enfuse: its transposed
pyramid operators against the forward operators (dot-product identity), and
its adjoint against central finite differences of its own forward."""
import numpy as np
import pytest

from oracle_binding import F64


def _scene(rng, K, h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    base = 0.3 + 0.25 * np.sin(xx / 3.0 + rng.uniform(0, 3)) * np.cos(yy / 2.5) + 0.05 * rng.random((h, w))
    tint = rng.uniform(0.7, 1.3, size=3)
    return np.stack([np.clip(base[..., None] * tint * 2.0 ** (k - (K - 1) / 2), 0, None) for k in range(K)])


@pytest.mark.parametrize("fw,fh", [(2, 2), (3, 5), (9, 4), (16, 13), (33, 20), (7, 1 + 1)])
def test_expand_transpose(oracle, fw, fh):
    rng = np.random.default_rng(fw * 100 + fh)
    cw, ch = (fw + 1) // 2, (fh + 1) // 2
    x = rng.standard_normal((ch, cw))
    y = rng.standard_normal((fh, fw))
    ex = oracle.expand(x, np.zeros((fh, fw)), True, False, F64)        # expand(x) added to 0
    lhs = float(np.sum(ex * y))
    rhs = float(np.sum(x * oracle.expand_T(y, cw, ch)))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


@pytest.mark.parametrize("sw,sh", [(2, 2), (3, 5), (9, 4), (16, 13), (33, 20), (5, 2)])
def test_reduce_transpose(oracle, sw, sh):
    rng = np.random.default_rng(sw * 100 + sh + 7)
    x = rng.standard_normal((sh, sw))
    dw, dh = (sw + 1) // 2, (sh + 1) // 2
    y = rng.standard_normal((dh, dw))
    rx, _ = oracle.reduce(x, np.full((sh, sw), 255, np.uint8), False, F64)
    lhs = float(np.sum(rx * y))
    rhs = float(np.sum(x * oracle.reduce_T(y, sw, sh)))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_transposes_fail_when_wrong(oracle):
    """The dot-product test detects a transpose that is off by one tap."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5, 6))
    y = rng.standard_normal((10, 11))
    ex = oracle.expand(x, np.zeros((10, 11)), True, False, F64)
    wrong = oracle.expand_T(np.roll(y, 1, axis=1), 6, 5)
    assert abs(np.sum(ex * y) - np.sum(x * wrong)) > 1e-6


def _loss_and_grad(oracle, imgs, prm, L, target):
    out = oracle.diff_forward(imgs, prm, L)
    loss = float(np.sum((out - target) ** 2))
    gi, gp = oracle.diff_backward(imgs, prm, L, 2.0 * (out - target))
    return loss, gi, gp


@pytest.mark.parametrize("K,h,w,L", [(3, 19, 23, 3), (2, 12, 9, 2), (4, 16, 16, 4)])
def test_adjoint_matches_finite_differences(oracle, K, h, w, L):
    rng = np.random.default_rng(K * 1000 + h * 10 + w)
    imgs = _scene(rng, K, h, w)
    prm = np.array([1.0, 0.3, 0.5, 0.45, 0.25])
    target = rng.random((h, w, 3))
    loss, gi, gp = _loss_and_grad(oracle, imgs, prm, L, target)
    step = 1e-5

    def fd(f):
        return (f(step) - f(-step)) / (2 * step)

    for q in range(5):
        def f(d, q=q):
            p2 = prm.copy()
            p2[q] += d
            return float(np.sum((oracle.diff_forward(imgs, p2, L) - target) ** 2))
        g = fd(f)
        assert abs(gp[q] - g) <= 1e-6 * max(abs(g), 1e-3), (q, gp[q], g)
    for _ in range(25):
        k, y, x, c = rng.integers(0, K), rng.integers(0, h), rng.integers(0, w), rng.integers(0, 3)
        def f(d):
            i2 = imgs.copy()
            i2[k, y, x, c] += d
            return float(np.sum((oracle.diff_forward(i2, prm, L) - target) ** 2))
        g = fd(f)
        assert abs(gi[k, y, x, c] - g) <= 1e-6 * max(abs(g), 1e-3), ((k, y, x, c), gi[k, y, x, c], g)


def test_adjoint_check_can_fail(oracle):
    """A planted mistake in the oracle's inputs (a wrong sign on one parameter
    gradient) is caught by the same comparison."""
    rng = np.random.default_rng(11)
    imgs = _scene(rng, 3, 14, 15)
    prm = np.array([1.0, 0.3, 0.5, 0.45, 0.25])
    target = rng.random((14, 15, 3))
    _, _, gp = _loss_and_grad(oracle, imgs, prm, 3, target)
    step = 1e-5
    p_hi, p_lo = prm.copy(), prm.copy()
    p_hi[3] += step
    p_lo[3] -= step
    g = (np.sum((oracle.diff_forward(imgs, p_hi, 3) - target) ** 2) -
         np.sum((oracle.diff_forward(imgs, p_lo, 3) - target) ** 2)) / (2 * step)
    assert abs(-gp[3] - g) > 1e-6 * max(abs(g), 1e-3)
