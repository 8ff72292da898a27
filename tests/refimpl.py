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


# ---------------------------------------------------------------- search walkers
#
# Independent step-by-step restatements of the reference's two search drivers,
# written from the reference source, over textbook costs: the prediction comes
# from the spec's sample equations on the unpadded frame (spec_pred below: no
# plane sets, no operand table, no padding), the SATD from the matrix Hadamard,
# the rate from the Exp-Golomb length or a per-|d| table.  Each priced candidate
# is recorded, so the oracle's evaluation order is compared, not only its result.

def spec_pred(ref, x0, y0, bw, bh, mvx, mvy):
    """Luma prediction of the bw x bh block at (x0, y0) displaced by the quarter-pel
    vector, from ITU-T H.264 8.4.2.2.1 (eq. 8-241..8-261, Table 8-12) with every
    full-sample coordinate clamped into the frame."""
    H, W = ref.shape
    ix, iy = x0 + (mvx >> 2), y0 + (mvy >> 2)
    fx, fy = mvx & 3, mvy & 3
    rows = np.clip(np.arange(iy - 2, iy + bh + 4), 0, H - 1)
    cols = np.clip(np.arange(ix - 2, ix + bw + 4), 0, W - 1)
    G = ref[np.ix_(rows, cols)].astype(np.int64)          # G[r, c] = sample (ix-2+c, iy-2+r)
    T = [1, -5, 20, 20, -5, 1]
    B1 = sum(T[k] * G[:, k:k + bw + 1] for k in range(6))       # horizontal intermediates b1
    H1 = sum(T[k] * G[k:k + bh + 1, :] for k in range(6))       # vertical intermediates h1
    J1 = sum(T[k] * B1[k:k + bh + 1, :] for k in range(6))      # centre intermediates j1
    clip = lambda v: np.clip(v, 0, 255)
    g = G[2:2 + bh, 2:2 + bw]
    g_right, g_down = G[2:2 + bh, 3:3 + bw], G[3:3 + bh, 2:2 + bw]
    b = clip((B1[2:2 + bh, :bw] + 16) >> 5)
    s = clip((B1[3:3 + bh, :bw] + 16) >> 5)                     # b one row down
    h = clip((H1[:bh, 2:2 + bw] + 16) >> 5)
    m = clip((H1[:bh, 3:3 + bw] + 16) >> 5)                     # h one column right
    j = clip((J1[:bh, :bw] + 512) >> 10)
    av = lambda p, q: (p + q + 1) >> 1
    table = {(0, 0): g, (1, 0): av(g, b), (2, 0): b, (3, 0): av(g_right, b),
             (0, 1): av(g, h), (1, 1): av(b, h), (2, 1): av(b, j), (3, 1): av(b, m),
             (0, 2): h, (1, 2): av(h, j), (2, 2): j, (3, 2): av(j, m),
             (0, 3): av(g_down, h), (1, 3): av(h, s), (2, 3): av(j, s), (3, 3): av(m, s)}
    return table[(fx, fy)].astype(np.uint8)


def smooth_mv_table():
    """The smooth mvd rate of mb16.rs:507-516: round(4 * (2 log2(a+1) + 0.718 + [a != 0]))."""
    import math
    return np.array([int(math.floor(4.0 * (2.0 * math.log2(a + 1) + 0.718 + (1.0 if a else 0.0)) + 0.5))
                     for a in range(4096)], np.uint16)


def _trunc_div(a, b):
    q = abs(a) // b
    return -q if a < 0 else q


def motion_search_walk(cur, ref, lx, ly, bw, bh, preds, lam, lam_fp=None, steps=(16, 8, 4), sadfp=False,
                       snap=True, fc=False, sad_only=False, mv_tab=None, rescue=False, rescue_range=24,
                       rescue_thresh=3, rescue_fast=True, rescue_var=800, subpel_pat=0):
    """motion_search, mb16.rs:2433-3302, walked over textbook costs.
    Returns (mv, cost, (rescue ran, rescue paid), trace[(phase, mvx, mvy, cost)])."""
    H, W = cur.shape
    lam_fp = lam * 0.5 if lam_fp is None else lam_fp
    src = cur[ly:ly + bh, lx:lx + bw]
    center = tuple(preds[0]) if len(preds) else (0, 0)
    bits = (lambda d: mvbits(d)) if mv_tab is None else (lambda d: int(mv_tab[min(abs(d), 4095)]))
    rate = lambda mv: bits(mv[0] - center[0]) + bits(mv[1] - center[1])
    price = lambda l, mv: int(l * float(rate(mv)))        # (lambda * rate as f64) as i64
    trace = []

    def dist(mv, use_sad):
        pred = spec_pred(ref, lx, ly, bw, bh, *mv)
        return sad(src, pred) if use_sad else satd_sum(src, pred)

    def priced(phase, mv, d, l):
        c = d + price(l, mv)
        trace.append((phase, mv[0], mv[1], c))
        return c

    lam_r = lam * 0.25 if mv_tab is not None else lam      # mb16.rs:2553
    cost = lambda ph, mv: priced(ph, mv, dist(mv, sad_only), lam_r)
    cost_fp = lambda ph, mv: priced(ph, mv, dist(mv, True), lam_fp) if sadfp else cost(ph, mv)
    # seeds (0,0) + predictors
    best = (0, 0)
    best_c = cost_fp(0, best)
    for p in preds:
        p = (int(p[0]), int(p[1]))
        c = cost_fp(0, p)
        if c < best_c:
            best, best_c = p, c
    seed, seed_c = best, best_c
    if snap and (best[0] & 3 or best[1] & 3):
        best = (((best[0] + 2) // 4) * 4, ((best[1] + 2) // 4) * 4)   # div_euclid by 4 = floor
        best_c = cost_fp(1, best)
    use_fc = fc and not sad_only and (bw, bh) in ((16, 16), (16, 8), (8, 16), (8, 8))
    for step in steps:
        ring = [(step, 0), (-step, 0), (0, step), (0, -step)]
        while True:
            if use_fc and best[0] % 4 == 0 and best[1] % 4 == 0:
                s = step >> 2
                px, py = lx + (best[0] >> 2), ly + (best[1] >> 2)
                if px - s >= 0 and py - s >= 0 and px + s + bw <= W and py + s + bh <= H:
                    cands = [(best[0] + dx, best[1] + dy) for dx, dy in ring]
                    if bw == 16:   # the x4 batch: SAD or SATD + lam_fp (lambda outside B2)
                        cs = [priced(3, mv, dist(mv, sadfp), lam_fp if sadfp else lam) for mv in cands]
                    else:
                        cs = [cost_fp(3, mv) for mv in cands]
                    k = min(range(4), key=lambda i: (cs[i], i))
                    if cs[k] >= best_c:
                        break
                    best, best_c = cands[k], cs[k]
                    continue
            moved = False
            for dx, dy in ring:
                mv = (best[0] + dx, best[1] + dy)
                c = cost_fp(2, mv)
                if c < best_c:
                    best, best_c, moved = mv, c, True
            if not moved:
                break
    if sadfp:
        best_c = cost(4, best)
        seed_c = cost(4, seed)
    events = (0, 0)
    if rescue and not sad_only:
        flat = False
        if not rescue_fast:
            v = src.astype(np.int64)
            n = bw * bh
            flat = (int((v * v).sum()) - int(v.sum()) ** 2 // n) // n < rescue_var
        if rescue_fast or flat:
            d = best_c - price(lam, best)
            if _trunc_div(d, bw * bh) > rescue_thresh:
                pre = best_c
                gx, gy = ((best[0] + 2) // 4) * 4, ((best[1] + 2) // 4) * 4
                r = rescue_range
                px, py = lx + gx // 4, ly + gy // 4
                grid_batch = (bw, bh) == (16, 16) and px >= r and px + r + 16 <= W and py >= r and py + r + 16 <= H
                gb = best
                for dy in range(-r, r + 1, 2):
                    for dx in range(-r, r + 1, 2):
                        mv = (gx + 4 * dx, gy + 4 * dy)
                        c = priced(5, mv, dist(mv, False), lam) if grid_batch else cost(5, mv)
                        if c < best_c:
                            best_c, gb = c, mv
                best = gb
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        mv = (best[0] + 4 * dx, best[1] + 4 * dy)
                        c = cost(6, mv)
                        if c < best_c:
                            best, best_c = mv, c
                events = (1, int(best_c * 16 <= pre * 15))
    if subpel_pat == 4:
        s0 = best
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                if dx or dy:
                    mv = (s0[0] + dx, s0[1] + dy)
                    c = cost(7, mv)
                    if c < best_c:
                        best, best_c = mv, c
    elif subpel_pat >= 0:
        for st in (2, 1):
            ring = [(st, 0), (-st, 0), (0, st), (0, -st), (st, st), (-st, -st), (st, -st), (-st, st)]
            ring = ring[:4] if subpel_pat & 1 else ring
            while True:
                moved = False
                for dx, dy in ring:
                    mv = (best[0] + dx, best[1] + dy)
                    c = cost(7, mv)
                    if c < best_c:
                        best, best_c, moved = mv, c, True
                if not moved or subpel_pat & 2:
                    break
    if snap and seed_c < best_c:
        best, best_c = seed, seed_c
    return best, best_c, events, trace


def inter_cost_walk(cur, ref, bx0, by0, bs, seed, max_step=8):
    """inter_cost, mbtree.rs:154-197, over textbook SATDs (mc_satd's value,
    mbtree.rs:102-148, is the encoder-domain SATD of the clamped prediction).
    Returns (cost, mv, trace[(mvx, mvy, satd)])."""
    src = cur[by0:by0 + bs, bx0:bx0 + bs]
    trace = []

    def satd_at(mv):
        v = satd_sum(src, spec_pred(ref, bx0, by0, bs, bs, *mv))
        trace.append((mv[0], mv[1], v))
        return v

    best_mv = (0, 0)
    best = satd_at(best_mv)
    if best == 0:
        return 0, best_mv, trace
    if tuple(seed) != (0, 0):
        s = satd_at(tuple(seed))
        if s < best:
            best, best_mv = s, tuple(seed)
    step = max_step
    while step >= 1:
        while True:
            moved = False
            for dx, dy in ((step, 0), (-step, 0), (0, step), (0, -step)):
                mv = (best_mv[0] + 4 * dx, best_mv[1] + 4 * dy)
                s = satd_at(mv)
                if s < best:
                    best, best_mv, moved = s, mv, True
            if not moved:
                break
        step >>= 1
    return min(best, 2 ** 31 - 1), best_mv, trace


def frame_costs_half_walk(half, ref_half, mb_w, mb_h):
    """frame_costs, mbtree.rs:208-263, HalfRes: raster order, the seed is the left
    macroblock's full-res vector (else the one above) halved with truncation; the
    intra cost is the AC sum of the 4x4 Hadamards (coefficient 0 excluded), >= 1."""
    intra = np.zeros((mb_h, mb_w), np.int64)
    inter = np.zeros((mb_h, mb_w), np.int64)
    mv = np.zeros((mb_h, mb_w, 2), np.int64)
    for y in range(mb_h):
        for x in range(mb_w):
            blk = half[8 * y:8 * y + 8, 8 * x:8 * x + 8].astype(np.int64)
            ac = 0
            for yy in (0, 4):
                for xx in (0, 4):
                    hh = hadamard(blk[yy:yy + 4, xx:xx + 4])
                    ac += int(np.abs(hh).sum() - abs(hh[0, 0]))
            intra[y, x] = max(1, min(ac, 2 ** 31 - 1))
            if ref_half is None:
                inter[y, x] = intra[y, x]
                continue
            sf = tuple(mv[y, x - 1]) if x > 0 else tuple(mv[y - 1, x]) if y > 0 else (0, 0)
            seed = (int(np.trunc(sf[0] / 2)), int(np.trunc(sf[1] / 2)))
            c, m, _ = inter_cost_walk(half, ref_half, 8 * x, 8 * y, 8, seed)
            inter[y, x] = min(c, intra[y, x])
            mv[y, x] = (2 * m[0], 2 * m[1])
    return intra, inter, mv
