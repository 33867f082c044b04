"""Closed-form (direct convolution) statements of reduce and expand.

These are written from the mathematics of the Burt-Adelson filters, not from
the SKIPSM state machines, so that they pin the oracle independently:

* reduce (src/pyramid.h:166-573, :601-848): 5x5 binomial [1 4 6 4 1]^2 taps
  centred at source (2i, 2j).  With alpha, taps outside the image or under
  alpha = 0 are dropped and the weighted sum is divided by the in-mask weight
  sum (C++ truncating '/'); without alpha the borders replicate the edge
  sample and the sum is divided by 256.  Wraparound is periodic in x.
* expand (src/pyramid.h:873-1419): polyphase taps [1 6 1] for even and [4 4]
  for odd outputs per axis (the vertical odd phase is [1 1]), out-of-image
  taps dropped, divided by the reference's SKIPSM scale constant for the
  output position.  Those constants are the known-answer table
  EXPAND_SCALES below, transcribed from the macro arguments at
  src/pyramid.h:1199-1411.

All integer arithmetic is exact (Python/numpy int64); double inputs that
hold integers make every sum exact, so only the final division rounds.
"""
import numpy as np

W5 = np.array([1, 4, 6, 4, 1], dtype=np.int64)


def trunc_div(s, d):
    """C++ '/' on integers (truncation toward zero)."""
    s = np.asarray(s, dtype=np.int64)
    d = np.asarray(d, dtype=np.int64)
    q = np.abs(s) // d
    return np.where(s < 0, -q, q)


def reduce_alpha(x, alpha, wrap):
    """Exact closed form of reduce with alpha on an integer array.
    Returns (sum S, weight A) arrays per destination pixel."""
    h, w = x.shape
    dh, dw = (h + 1) // 2, (w + 1) // 2
    S = np.zeros((dh, dw), np.int64)
    A = np.zeros((dh, dw), np.int64)
    m = (alpha != 0).astype(np.int64)
    xv = x.astype(np.int64) * m
    for i in range(dh):
        for j in range(dw):
            for ky in range(-2, 3):
                r = 2 * i + ky
                if r < 0 or r >= h:
                    continue
                for kx in range(-2, 3):
                    c = 2 * j + kx
                    if wrap:
                        c %= w
                    elif c < 0 or c >= w:
                        continue
                    wt = W5[ky + 2] * W5[kx + 2]
                    S[i, j] += wt * xv[r, c]
                    A[i, j] += wt * m[r, c]
    return S, A


def reduce_plain(x, wrap):
    """Exact closed-form sum of reduce without alpha (before the /256)."""
    h, w = x.shape
    dh, dw = (h + 1) // 2, (w + 1) // 2
    S = np.zeros((dh, dw), np.int64)
    xv = x.astype(np.int64)
    for i in range(dh):
        for j in range(dw):
            for ky in range(-2, 3):
                r = min(max(2 * i + ky, 0), h - 1)
                for kx in range(-2, 3):
                    c = 2 * j + kx
                    c = c % w if wrap else min(max(c, 0), w - 1)
                    S[i, j] += W5[ky + 2] * W5[kx + 2] * xv[r, c]
    return S


# SKIPSM_EXPAND* scale constants (out00, out10, out01, out11) by source row
# class, source column class and boundary mode, src/pyramid.h:1199-1411.
# Row classes: "second" = source row 0 of a source with >= 2 rows,
# "main" = interior source rows, "last" = last source row, "only" = 1-row
# source.  Column classes likewise.  Modes: "none", "even" (wraparound,
# even destination width), "odd" (wraparound, odd destination width).
_ALL = ("none", "even", "odd")
EXPAND_SCALES = {}


def _put(row, col, modes, vals):
    for m in modes:
        EXPAND_SCALES[(row, col, m)] = vals


_put("second", "second", ("none",), (49, 56, 14, 16))
_put("second", "second", ("even",), (56, 56, 16, 16))
_put("second", "second", ("odd",), (77, 56, 22, 16))
_put("second", "main", _ALL, (56, 56, 16, 16))
_put("second", "end", ("none",), (49, 28, 14, 8))
_put("second", "end", ("even",), (56, 56, 16, 16))
_put("second", "end", ("odd",), (77, None, 22, None))
_put("second", "only", _ALL, (42, 28, 12, 8))
_put("main", "second", ("none",), (56, 64, 14, 16))
_put("main", "second", ("even",), (64, 64, 16, 16))
_put("main", "second", ("odd",), (88, 64, 22, 16))
_put("main", "main", _ALL, (64, 64, 16, 16))
_put("main", "end", ("none",), (56, 32, 14, 8))
_put("main", "end", ("even",), (64, 64, 16, 16))
_put("main", "end", ("odd",), (88, None, 22, None))
_put("main", "only", _ALL, (48, 32, 12, 8))
_put("last", "second", ("none",), (49, 56, 7, 8))
_put("last", "second", ("even",), (56, 56, 8, 8))
_put("last", "second", ("odd",), (77, 56, 11, 8))
_put("last", "main", _ALL, (56, 56, 8, 8))
_put("last", "end", ("none",), (49, 28, 7, 4))
_put("last", "end", ("even",), (56, 56, 8, 8))
_put("last", "end", ("odd",), (77, 56, 11, 8))
_put("last", "only", _ALL, (42, 28, 6, 4))
_put("only", "second", ("none",), (42, 48, 7, 8))
_put("only", "second", ("even",), (48, 48, 8, 8))
_put("only", "second", ("odd",), (66, 48, 11, 8))
_put("only", "main", _ALL, (48, 48, 8, 8))
_put("only", "end", ("none",), (42, 24, 7, 4))
_put("only", "end", ("even",), (48, 48, 8, 8))
_put("only", "end", ("odd",), (66, 48, 11, 8))
_put("only", "only", _ALL, (36, 24, 6, 4))


def _cls(k, n):
    if n == 1:
        return "only"
    if k == 0:
        return "second"
    if k == n - 1:
        return "end"
    return "main"


def expand_scale(X, Y, sw, sh, wrap, dw):
    """Scale constant for destination pixel (X, Y)."""
    j, i = X // 2, Y // 2
    col = _cls(j, sw)
    row = _cls(i, sh)
    if row == "end":
        row = "last"
    mode = "none" if not wrap else ("even" if dw % 2 == 0 else "odd")
    vals = EXPAND_SCALES[(row, col, mode)]
    return vals[(X & 1) + 2 * (Y & 1)]


def expand_sum(x, X, Y, wrap, dw):
    """Exact closed-form numerator of expand at destination (X, Y): in-image
    polyphase taps.  With wraparound the x taps beyond either edge fold to the
    opposite edge, with weight 4 instead of 1 when the destination width is
    odd (src/pyramid.h:1154-1158, :1184, :1229).  For a 1-pixel-wide source
    under wraparound the reference folds both edge taps only in the first
    source row and only the left tap in later rows (its src_w == 1 branch
    takes the no-wraparound column macro, :1245, :1329)."""
    sh, sw = x.shape
    if X % 2 == 0:
        j = X // 2
        htaps = [(j - 1, 1), (j, 6), (j + 1, 1)]
    else:
        j = (X - 1) // 2
        htaps = [(j, 4), (j + 1, 4)]
    if Y % 2 == 0:
        i = Y // 2
        vtaps = [(i - 1, 1), (i, 6), (i + 1, 1)]
    else:
        i = (Y - 1) // 2
        vtaps = [(i, 1), (i + 1, 1)]
    total = 0
    for (r, wv) in vtaps:
        if r < 0 or r >= sh:
            continue
        for (c, wh) in htaps:
            if 0 <= c < sw:
                total += wv * wh * int(x[r, c])
                continue
            if not wrap:
                continue
            if sw == 1 and r > 0 and c >= sw:
                continue  # rows after the first lose the right fold
            odd_dst = dw % 2 == 1
            if X % 2 == 0:
                wf = 4 if odd_dst else 1
            else:
                wf = wh
            cc = c % sw
            total += wv * wf * int(x[r, cc])
    return total
