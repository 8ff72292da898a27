"""Textbook restatements used to pin the oracle (tests only).

Each function computes the same quantity as an oracle function by a different
method: matrix products instead of butterflies, scipy separable correlation
instead of tile walks, the ITU-T H.264 sample table (Table 8-12 / eq. 8-250..
8-261) instead of the reference's plane operand table, numpy sliding windows
instead of loops.
"""
import numpy as np
from scipy import ndimage

TAPS = np.array([1, -5, 20, 20, -5, 1], dtype=np.float64)
# Sylvester-ordered 4-point Hadamard (the matrix form of hadamard_1d)
H4 = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]], dtype=np.int64)


# order-8 Sylvester Hadamard, H2 (x) H2 (x) H2 (the build-defined 8x8 SATD)
H8 = np.kron(np.kron([[1, 1], [1, -1]], [[1, 1], [1, -1]]), [[1, 1], [1, -1]]).astype(np.int64)


def satd8_sum(a, b):
    """Sum over 8x8 blocks of sum |H8 d H8^T| / 2 (matrix form)."""
    d = a.astype(np.int64) - b.astype(np.int64)
    h, w = d.shape
    s = 0
    for y in range(0, h, 8):
        for x in range(0, w, 8):
            t = int(np.abs(H8 @ d[y:y + 8, x:x + 8] @ H8.T).sum())
            assert t % 2 == 0
            s += t // 2
    return s


def sad(a, b):
    return int(np.abs(a.astype(np.int64) - b.astype(np.int64)).sum())


def hadamard(d4):
    return H4 @ np.asarray(d4, np.int64) @ H4.T


def satd_sum(a, b):
    """Encoder-domain sum over 4x4 blocks of sum |H d| (matrix form)."""
    d = a.astype(np.int64) - b.astype(np.int64)
    h, w = d.shape
    s = 0
    for y in range(0, h, 4):
        for x in range(0, w, 4):
            s += int(np.abs(hadamard(d[y:y + 4, x:x + 4])).sum())
    return s


def satd_wels(a, b):
    d = a.astype(np.int64) - b.astype(np.int64)
    h, w = d.shape
    s = 0
    for y in range(0, h, 4):
        for x in range(0, w, 4):
            s += (int(np.abs(hadamard(d[y:y + 4, x:x + 4])).sum()) + 1) >> 1
    return s


def _corr(a, axis):
    # taps over offsets -2..+3 around each sample: centre at tap index 2 -> origin -1
    return ndimage.correlate1d(a.astype(np.float64), TAPS, axis=axis, mode="nearest", origin=-1)


def planes(y, pad):
    """F, H, V, C of the padded frame by separable correlation with edge replication."""
    f = np.pad(y, pad, mode="edge").astype(np.int64)
    hs = np.rint(_corr(f, 1)).astype(np.int64)
    vs = np.rint(_corr(f, 0)).astype(np.int64)
    cs = np.rint(_corr(_corr(f, 0), 1)).astype(np.int64)
    h = np.clip((hs + 16) >> 5, 0, 255)
    v = np.clip((vs + 16) >> 5, 0, 255)
    c = np.clip((cs + 512) >> 10, 0, 255)
    return np.stack([f, h, v, c]).astype(np.uint8)


def spec_samples(y, pad):
    """All 16 quarter-pel sample planes at every padded position, named per the
    H.264 spec (G, b, h, j, and the neighbours m = h at x+1, s = b at y+1),
    averaged per Table 8-12's equations.  Index [fx][fy] -> plane (padded grid,
    last row/column dropped so the +1 neighbours exist)."""
    F, b, h, j = (p.astype(np.int64) for p in planes(y, pad))
    G = F[:-1, :-1]
    bb, hh, jj = b[:-1, :-1], h[:-1, :-1], j[:-1, :-1]
    m = h[:-1, 1:]   # vertical half-pel one column right
    s = b[1:, :-1]   # horizontal half-pel one row down
    Gx = F[:-1, 1:]  # full-pel one column right ("H" in the spec's figure)
    Gy = F[1:, :-1]  # full-pel one row down ("M")
    avg = lambda p, q: (p + q + 1) >> 1
    t = {
        (0, 0): G, (1, 0): avg(G, bb), (2, 0): bb, (3, 0): avg(Gx, bb),
        (0, 1): avg(G, hh), (1, 1): avg(bb, hh), (2, 1): avg(bb, jj), (3, 1): avg(bb, m),
        (0, 2): hh, (1, 2): avg(hh, jj), (2, 2): jj, (3, 2): avg(jj, m),
        (0, 3): avg(Gy, hh), (1, 3): avg(hh, s), (2, 3): avg(jj, s), (3, 3): avg(m, s),
    }
    return t


def mvbits(d):
    """Length of the signed Exp-Golomb code se(d): codeNum k -> 2*floor(log2(k+1))+1."""
    k = 2 * d - 1 if d > 0 else -2 * d
    return 2 * (int(k + 1).bit_length() - 1) + 1


def full_search(cur, F, pad, lx, ly, bw, bh, R, cx, cy, mvp, lr_fp):
    """Brute force over all (2R+1)^2 candidates with numpy sliding windows;
    np.argmin returns the first minimum in raster (dy-major) order."""
    ph, pw = F.shape
    xs = lx + (cx >> 2) + pad + np.arange(-R, R + 1)
    ys = ly + (cy >> 2) + pad + np.arange(-R, R + 1)
    src = cur[ly:ly + bh, lx:lx + bw].astype(np.int64)
    costs = np.empty((2 * R + 1, 2 * R + 1), np.int64)
    for iy, yy in enumerate(ys):
        for ix, xx in enumerate(xs):
            rows = np.clip(np.arange(yy, yy + bh), 0, ph - 1)
            cols = np.clip(np.arange(xx, xx + bw), 0, pw - 1)
            blk = F[np.ix_(rows, cols)].astype(np.int64)
            mvx, mvy = cx + 4 * (ix - R), cy + 4 * (iy - R)
            costs[iy, ix] = np.abs(src - blk).sum() + lr_fp[mvbits(mvx - mvp[0]) + mvbits(mvy - mvp[1])]
    k = int(np.argmin(costs))
    iy, ix = divmod(k, 2 * R + 1)
    return (cx + 4 * (ix - R), cy + 4 * (iy - R)), int(costs.flat[k])
