"""Pins the oracle's two search drivers against independent walkers.

orc_motion_search_ex (motion_search, mb16.rs:2433-3302) and orc_inter_cost_ex
(inter_cost, mbtree.rs:154-197) are compared, candidate by candidate, with the
walkers of tests/refimpl.py, which price every candidate from the spec's sample
equations on the unpadded frame (no plane sets, no operand table) and the matrix
Hadamard.  The traces, not only the results, must agree: the order in which the
candidates are priced is the search.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from tests import refimpl

O = Oracle()


def _clip(cw, ch, seed, shift):
    """A smooth random texture and a copy moved by `shift` (dy, dx) with noise, so
    the diamond walks several steps and the sub-pel stage finds fractional optima."""
    rng = np.random.default_rng(seed)
    m = 16
    base = rng.integers(0, 256, (ch + 2 * m, cw + 2 * m)).astype(np.float64)
    for _ in range(3):
        base = (base + np.roll(base, 1, 0) + np.roll(base, 1, 1) + np.roll(base, (1, 1), (0, 1))) / 4
    ref = np.clip(base[m:m + ch, m:m + cw] * 1.6 - 80, 0, 255).astype(np.uint8)
    dy, dx = shift
    cur = base[m - dy:m - dy + ch, m - dx:m - dx + cw] * 1.6 - 80 + rng.normal(0, 2, (ch, cw))
    return np.clip(cur, 0, 255).astype(np.uint8), ref


def test_spec_pred_is_mc_luma():
    """The walkers' prediction (spec equations, clamped samples) equals mc_luma, which
    is pinned against the per-sample spec table elsewhere: every phase, every shape,
    vectors far outside the frame."""
    cw, ch = 48, 32
    _, ref = _clip(cw, ch, 1, (0, 0))
    rng = np.random.default_rng(2)
    for bw, bh in ((16, 16), (16, 8), (8, 16), (8, 8), (8, 4), (4, 8), (4, 4)):
        for _ in range(40):
            x0, y0 = int(rng.integers(0, cw - bw + 1)), int(rng.integers(0, ch - bh + 1))
            mvx, mvy = (int(v) for v in rng.integers(-240, 241, 2))
            assert np.array_equal(refimpl.spec_pred(ref, x0, y0, bw, bh, mvx, mvy),
                                  O.mc_luma(ref, cw, ch, x0, y0, bw, bh, mvx, mvy)), (bw, bh, mvx, mvy)


def test_smooth_table_is_rounding_robust():
    """The smooth rate table (mb16.rs:507-516) rounds 4 * (2 log2(a+1) + 0.718 + [a>0]);
    no entry lies within 1e-6 of a rounding boundary, so any log2 accurate to a few
    ulps (the reference calls the platform's) yields the same table."""
    import math
    for a in range(4096):
        v = 4.0 * (2.0 * math.log2(a + 1) + 0.718 + (1.0 if a else 0.0))
        assert abs((v - math.floor(v)) - 0.5) > 1e-6, a
    t = refimpl.smooth_mv_table()
    assert t[0] == 3 and t[1] == 15 and t[4095] == 103   # 2.872 -> 3; 4 * 3.718; 4 * 25.718
    assert (np.diff(t.astype(int)) >= 0).all()


SMOOTH = refimpl.smooth_mv_table()

# (name, shape, search knobs, sub-pel pattern)
CASES = [
    ("cascade", (16, 16), dict(), 0),
    ("fc", (16, 16), dict(fc=True), 0),
    ("fc-b2", (16, 16), dict(fc=True, sadfp=True), 0),
    ("fc-smooth", (16, 16), dict(fc=True, mv_tab=SMOOTH), 0),
    ("fc-b2-smooth", (16, 16), dict(fc=True, sadfp=True, mv_tab=SMOOTH), 2),
    ("fc-16x8", (16, 8), dict(fc=True, mv_tab=SMOOTH), 0),
    ("fc-8x16", (8, 16), dict(fc=True, sadfp=True), 1),
    ("fc-8x8-smooth", (8, 8), dict(fc=True, mv_tab=SMOOTH), 3),
    ("sub8-8x4", (8, 4), dict(fc=True, steps=(4,)), 2),
    ("sub8-4x8", (4, 8), dict(steps=(4,), mv_tab=SMOOTH), 2),
    ("sub8-4x4", (4, 4), dict(fc=True, steps=(4,)), 2),
    ("fast", (16, 16), dict(sad_only=True, steps=(16, 4)), -1),
    ("fast-subpel", (16, 16), dict(sad_only=True, steps=(16, 4), mv_tab=SMOOTH), 2),
    ("window", (16, 16), dict(fc=True), 4),
    ("no-snap", (16, 16), dict(snap=False, fc=True), 0),
    ("wide-rungs", (16, 16), dict(fc=True, steps=(64, 32, 16, 8, 4)), 0),
    ("rescue", (16, 16), dict(fc=True, rescue=True, rescue_range=10, rescue_thresh=0), 0),
    ("rescue-smooth", (16, 16), dict(fc=True, rescue=True, rescue_range=8, rescue_thresh=-2, mv_tab=SMOOTH), 2),
    ("rescue-flat", (16, 8), dict(rescue=True, rescue_range=6, rescue_thresh=0, rescue_fast=False,
                                  rescue_var=900), 0),
    ("rescue-b2", (16, 16), dict(fc=True, sadfp=True, rescue=True, rescue_range=8, rescue_thresh=1), 0),
]


@pytest.mark.parametrize("name,shape,kw,pat", CASES, ids=[c[0] for c in CASES])
def test_motion_search_trace(name, shape, kw, pat):
    cw, ch = 64, 48
    bw, bh = shape
    cur, ref = _clip(cw, ch, 7 + len(name), (3, -5))
    pad = 32
    pl = O.planes(ref, cw, ch, pad)
    lam = 4.65
    cfg = O.cfg(cw, ch, pad, bw, bh, 16, pat, lam)
    rng = np.random.default_rng(len(name))
    blocks = [(x, y) for y in range(0, ch, bh) for x in range(0, cw, bw)]
    idx = rng.choice(len(blocks), min(len(blocks), 6), replace=False)
    fired = 0
    for k, i in enumerate(idx):
        lx, ly = blocks[i]
        npreds = k % 4
        preds = [tuple(int(v) for v in rng.integers(-40, 41, 2)) for _ in range(npreds)]
        if npreds and k % 2:
            preds[0] = (4 * int(rng.integers(-6, 7)), 4 * int(rng.integers(-6, 7)))
        mv, c, ev, tr = O.motion_search_trace(cfg, cur, pl, lx, ly, preds, **kw)
        wmv, wc, wev, wtr = refimpl.motion_search_walk(cur, ref, lx, ly, bw, bh, preds, lam, lam * 0.5,
                                                       subpel_pat=pat, **kw)
        assert tr == wtr, (name, lx, ly, preds, next(j for j, (a, b) in enumerate(zip(tr + [None], wtr + [None]))
                                                     if a != b))
        assert (mv, c, ev) == (wmv, wc, wev)
        fired += ev[0]
    if kw.get("rescue"):
        assert fired > 0, "the rescue grid never ran: the case pins nothing"


def tie_corpus():
    """Searches on a texture of period 4 in x and 8 in y at lambda 0, where distinct
    vectors price exactly alike: (frame, ref, lx, ly, preds, knobs, pattern)."""
    cw, ch = 64, 48
    yy, xx = np.mgrid[0:ch, 0:cw]
    tex = (60 + 40 * (xx % 4) + 17 * (yy % 8)).astype(np.uint8)
    out = []
    for kw, pat in ((dict(), 0), (dict(fc=True), 0), (dict(fc=True, sadfp=True), 2), (dict(sad_only=True), -1),
                    (dict(fc=True, rescue=True, rescue_range=6, rescue_thresh=-1), 3)):
        for lx, ly, preds in ((16, 16, [(16, 0)]), (32, 16, [(-16, 32), (16, 0)]), (16, 0, [(0, 32)]),
                              (0, 16, [(4, 2), (16, 32)])):
            out.append((tex, tex, lx, ly, preds, kw, pat))
    return out


def test_motion_search_ties():
    """Ties at every phase (seeds, fixed-centre and cascade rings, grid, sub-pel) keep
    the earlier candidate, as the reference's strict comparisons do."""
    for cur, ref, lx, ly, preds, kw, pat in tie_corpus():
        cw, ch = cur.shape[1], cur.shape[0]
        cfg = O.cfg(cw, ch, 32, 16, 16, 16, pat, 0.0, lam_fp=0.0)
        got = O.motion_search_trace(cfg, cur, O.planes(ref, cw, ch, 32), lx, ly, preds, **kw)
        assert got == refimpl.motion_search_walk(cur, ref, lx, ly, 16, 16, preds, 0.0, 0.0, subpel_pat=pat, **kw)


def test_motion_search_edges_and_clamped_fc():
    """Blocks on every frame edge (the fixed-centre box test and the clamped reads),
    predictors pointing outside the frame."""
    cw, ch = 48, 48
    cur, ref = _clip(cw, ch, 33, (-2, 6))
    pl = O.planes(ref, cw, ch, 32)
    cfg = O.cfg(cw, ch, 32, 16, 16, 16, 0, 3.0)
    for lx, ly in ((0, 0), (32, 0), (0, 32), (32, 32), (16, 16)):
        for preds in ([(-90, 13)], [(140, -77), (4, 4)], []):
            kw = dict(fc=True, sadfp=True, mv_tab=SMOOTH)
            mv, c, ev, tr = O.motion_search_trace(cfg, cur, pl, lx, ly, preds, **kw)
            assert (mv, c, ev, tr) == refimpl.motion_search_walk(cur, ref, lx, ly, 16, 16, preds, 3.0, 1.5, **kw)


@pytest.mark.parametrize("seed,shift", [(1, (2, 1)), (2, (-5, 7)), (3, (11, -9))])
def test_inter_cost_trace(seed, shift):
    cw, ch = 48, 40
    cur, ref = _clip(cw, ch, 50 + seed, shift)
    rng = np.random.default_rng(seed)
    for by in range(0, ch, 8):
        for bx in range(0, cw, 8):
            seed_mv = tuple(4 * int(v) for v in rng.integers(-5, 6, 2)) if (bx + by) % 16 else (0, 0)
            c, mv, tr = O.inter_cost_trace(cur, cw, ch, ref, bx, by, 8, seed_mv)
            wc, wmv, wtr = refimpl.inter_cost_walk(cur, ref, bx, by, 8, seed_mv)
            assert (c, mv, tr) == (wc, wmv, wtr), (bx, by)
    # static block: the early out prices (0, 0) only
    c, mv, tr = O.inter_cost_trace(ref, cw, ch, ref, 8, 8, 8, (12, -4))
    assert (c, mv, len(tr)) == (0, (0, 0), 1)


def test_frame_costs_half_walk():
    """frame_costs (HalfRes) == the raster walk with left/above seeds halved."""
    mb_w, mb_h = 6, 4
    cur, ref = _clip(16 * mb_w, 16 * mb_h, 77, (4, -6))
    hc, hr = O.downsample2x(cur, 16 * mb_w, 16 * mb_h), O.downsample2x(ref, 16 * mb_w, 16 * mb_h)
    a, b, m = O.frame_costs_half(hc, hr, mb_w, mb_h)
    wa, wb, wm = refimpl.frame_costs_half_walk(hc, hr, mb_w, mb_h)
    assert np.array_equal(a.reshape(mb_h, mb_w), wa)
    assert np.array_equal(b.reshape(mb_h, mb_w), wb)
    assert np.array_equal(m.reshape(mb_h, mb_w, 2), wm)
    assert (wm != 0).any()


# ------------------------------------------------ the partition decision's frame walk

def _predict_mv(a, b, c, cur_ref):
    """predict_mv, inter.rs:36-73 (spec 8.4.1.3.1); neighbours (available, mv, ref)."""
    res = lambda n: (n[1], n[2]) if n[0] and n[2] >= 0 else ((0, 0), -1)
    (ma, ra), (mb, rb), (mc, rc) = res(a), res(b), res(c)
    if not b[0] and not c[0] and a[0]:
        mb, rb, mc, rc = ma, ra, ma, ra
    hits = [m for m, r in ((ma, ra), (mb, rb), (mc, rc)) if r == cur_ref]
    if len(hits) == 1:
        return hits[0]
    return tuple(sorted(v)[1] for v in zip(ma, mb, mc))


def partition_decision_walk(cur, planes, cfg, lme, gate, p8x8, kw):
    """mb16.rs:8849-9010 in the SATD currency over a 4x4 motion grid: each decided
    macroblock writes its partitions' (mv, ref) into the 4x4 blocks they cover; the
    next macroblocks read A = (4x-1, 4y), B = (4x, 4y-1), C = (4x+4, 4y-1) or, when
    C is unavailable, D = (4x-1, 4y-1) (mv_neighbors_block, mb16.rs:1871-1888)."""
    ch, cw = cur.shape
    nbx, nby = cw // 16, ch // 16
    W4, H4 = 4 * nbx, 4 * nby
    grid_mv = np.zeros((H4, W4, 2), np.int64)
    grid_ref = np.full((H4, W4), -1, np.int64)
    coded = np.zeros((H4, W4), bool)

    def get(bx, by):
        if bx < 0 or by < 0 or bx >= W4 or by >= H4 or not coded[by, bx]:
            return (False, (0, 0), -1)
        return (True, tuple(int(v) for v in grid_mv[by, bx]), int(grid_ref[by, bx]))

    refbits = lambda r, n: 0 if n <= 1 else 1 if n == 2 else 2 * (r + 1).bit_length() - 1
    # the rescue's learning window over the frame's searches in call order
    kw = dict(kw)
    learn, payoff = kw.pop("rescue_learn", 40), kw.pop("rescue_payoff", 15)
    rescue = kw.pop("rescue", False)
    win = {"n": 0, "big": 0, "off": False}

    def best_part(nb, lx, ly, bw, bh, extra):
        c2 = O.cfg(cw, ch, 32, bw, bh, cfg.range, cfg.subpel_pat, cfg._lam, lam_fp=cfg._lam_fp)
        best = None
        for r, pl in enumerate(planes):
            seeds = [_predict_mv(*nb, r)] + extra
            mv, c, (ran, paid), _ = O.motion_search_trace(c2, cur, pl, lx, ly, seeds, cap=0,
                                                          rescue=rescue and not win["off"], **kw)
            if ran and win["n"] < learn:
                win["n"] += 1
                win["big"] += bool(paid)
                if win["n"] == learn and win["big"] * 100 < learn * payoff:
                    win["off"] = True
            c += int(lme * refbits(r, len(planes)))
            if best is None or c < best[0]:
                best = (c, r, mv)
        return best

    rec = np.zeros((nbx * nby, 14), np.int64)
    for y in range(nby):
        for x in range(nbx):
            a, b = get(4 * x - 1, 4 * y), get(4 * x, 4 * y - 1)
            c = get(4 * x + 4, 4 * y - 1)
            if not c[0]:
                c = get(4 * x - 1, 4 * y - 1)
            nb = (a, b, c)
            c16, r16, mv16 = best_part(nb, 16 * x, 16 * y, 16, 16, [])
            mode, best = 0, c16
            parts = [(0, 0, 16, 16, r16, mv16)]
            if c16 > gate:
                t = best_part(nb, 16 * x, 16 * y, 16, 8, [mv16])
                bo = best_part(nb, 16 * x, 16 * y + 8, 16, 8, [mv16])
                le = best_part(nb, 16 * x, 16 * y, 8, 16, [mv16])
                ri = best_part(nb, 16 * x + 8, 16 * y, 8, 16, [mv16])
                if t[0] + bo[0] < best:
                    mode, best = 1, t[0] + bo[0]
                    parts = [(0, 0, 16, 8, t[1], t[2]), (0, 8, 16, 8, bo[1], bo[2])]
                if le[0] + ri[0] < best:
                    mode, best = 2, le[0] + ri[0]
                    parts = [(0, 0, 8, 16, le[1], le[2]), (8, 0, 8, 16, ri[1], ri[2])]
                if p8x8:
                    qs = [best_part(nb, 16 * x + qx, 16 * y + qy, 8, 8, [mv16]) for qx, qy in
                          ((0, 0), (8, 0), (0, 8), (8, 8))]
                    c8 = int(lme * 4.0) + sum(q[0] for q in qs)
                    if c8 < best:
                        mode, best = 3, c8
                        parts = [(qx, qy, 8, 8, q[1], q[2]) for (qx, qy), q in
                                 zip(((0, 0), (8, 0), (0, 8), (8, 8)), qs)]
            for px, py, pw, ph, r, mv in parts:
                by0, bx0 = 4 * y + py // 4, 4 * x + px // 4
                grid_mv[by0:by0 + ph // 4, bx0:bx0 + pw // 4] = mv
                grid_ref[by0:by0 + ph // 4, bx0:bx0 + pw // 4] = r
            coded[4 * y:4 * y + 4, 4 * x:4 * x + 4] = True
            o = rec[y * nbx + x]
            o[0], o[1] = mode, best
            for q, (qx, qy) in enumerate(((0, 0), (3, 0), (0, 3), (3, 3))):
                o[2 + 3 * q] = grid_ref[4 * y + qy, 4 * x + qx]
                o[3 + 3 * q: 5 + 3 * q] = grid_mv[4 * y + qy, 4 * x + qx]
    return rec


def partition_clip(cw, ch, nrefs, seed):
    """References of independent smooth textures and a current frame whose macroblock
    quadrants copy them at chosen displacements: macroblock (x, y) is uniform, split
    top/bottom, left/right or into quadrants by (x + 2y) % 4, each piece taking its
    own reference and motion, so every partition mode and reference has a reason to
    win."""
    rng = np.random.default_rng(seed)
    refs = [_clip(cw, ch, seed * 7 + k, (0, 0))[1] for k in range(nrefs)]
    cur = np.zeros((ch, cw), np.float64)
    yy, xx = np.mgrid[0:8, 0:8]
    for y in range(ch // 16):
        for x in range(cw // 16):
            kind = (x + 2 * y) % 4
            pieces = [(int(rng.integers(0, nrefs)), int(rng.integers(-6, 7)), int(rng.integers(-6, 7)))
                      for _ in range(4)]
            group = {0: (0, 0, 0, 0), 1: (0, 0, 1, 1), 2: (0, 1, 0, 1), 3: (0, 1, 2, 3)}[kind]
            for q, (qx, qy) in enumerate(((0, 0), (8, 0), (0, 8), (8, 8))):
                k, dx, dy = pieces[group[q]]
                ry = np.clip(16 * y + qy + yy + dy, 0, ch - 1)
                rx = np.clip(16 * x + qx + xx + dx, 0, cw - 1)
                cur[16 * y + qy:16 * y + qy + 8, 16 * x + qx:16 * x + qx + 8] = refs[k][ry, rx]
    cur = np.clip(cur + rng.normal(0, 1.5, cur.shape), 0, 255).astype(np.uint8)
    return cur, refs


@pytest.mark.parametrize("nrefs,qp,p8x8,kw,pat", [(1, 26, True, dict(fc=True), 0),
                                                   (2, 20, True, dict(fc=True, mv_tab=SMOOTH), 2),
                                                   (1, 14, False, dict(sadfp=True), 0),
                                                   (3, 26, True, dict(fc=True, sadfp=True), 4)])
def test_partition_decision_walk(nrefs, qp, p8x8, kw, pat):
    """The oracle's raster-order partition decision == the walk over the 4x4 motion
    grid (every mode occurs)."""
    cw, ch = 80, 64
    cur, frames = partition_clip(cw, ch, nrefs, qp + nrefs)
    planes = [O.planes(f, cw, ch, 32) for f in frames]
    lam = float(np.sqrt(0.85 * 2.0 ** ((qp - 12) / 3.0)))
    cfg = O.cfg(cw, ch, 32, 16, 16, 16, pat, lam)
    gate = O.split_gate(qp)
    got = O.partition_decision_frame(cfg, cur, planes, lam, gate, p8x8, **kw)
    want = partition_decision_walk(cur, planes, cfg, lam, gate, p8x8, kw)
    assert np.array_equal(got, want), np.argwhere((got != want).any(1))[:5]
    assert set(got[:, 0].tolist()) == ({0, 1, 2, 3} if p8x8 else {0, 1, 2}), np.bincount(got[:, 0])
    assert set(got[:, 2::3].ravel().tolist()) == set(range(nrefs))


@pytest.mark.parametrize("learn,payoff,nrefs", [(6, 50, 1), (20, 40, 1), (1000, 15, 2), (8, 60, 2)])
def test_partition_decision_rescue_window(learn, payoff, nrefs):
    """The partition decision with the rescue: the window counts every search in call
    order (the 16x16 references, then 16x8, 8x16 and the 8x8 quadrants) == the walk."""
    cw, ch, qp = 96, 64, 20
    cur, ref0 = latch_clip(cw, ch, 3, (10, -13))
    rng = np.random.default_rng(nrefs)
    frames = [ref0] + [np.clip(ref0.astype(int) + rng.integers(-6, 7, ref0.shape), 0, 255).astype(np.uint8)
                       for _ in range(nrefs - 1)]
    planes = [O.planes(f, cw, ch, 32) for f in frames]
    lam = 4.0
    cfg = O.cfg(cw, ch, 32, 16, 16, 16, 0, lam)
    gate = O.split_gate(qp)
    kw = dict(fc=True, rescue=True, rescue_range=12, rescue_thresh=0, rescue_learn=learn, rescue_payoff=payoff)
    got = O.partition_decision_frame(cfg, cur, planes, lam, gate, True, **kw)
    want = partition_decision_walk(cur, planes, cfg, lam, gate, True, kw)
    assert np.array_equal(got, want), np.argwhere((got != want).any(1))[:5]
    assert (got[:, 0] > 0).any()  # some macroblocks split: their searches are in the window
    if learn < 1000:  # the window closes on this clip and changes decisions
        assert (got != O.partition_decision_frame(cfg, cur, planes, lam, gate, True,
                                                  **dict(kw, rescue_learn=1 << 30))).any()


# ---- the frame decision's me_wide learning window (mb16.rs:2868-2886) ----
def decision_walk(cur, refs, lam, lambda_ref, learn, payoff, kw):
    """The 16x16 frame decision in raster order, each macroblock's searches in
    reference order, composed from the independent search walker and predict_mv; the
    rescue learning window restated: the first `learn` searches whose grid ran are
    tallied (those that paid counted), and a window that fills with fewer than
    `payoff` % paid turns the rescue off for the rest of the frame."""
    ch, cw = cur.shape
    nbx, nby = cw // 16, ch // 16
    dec = {}
    n = big = 0
    off = False
    for y in range(nby):
        for x in range(nbx):
            def nb(xx, yy):
                if 0 <= xx < nbx and 0 <= yy < nby and (xx, yy) in dec:
                    r_, mv_ = dec[(xx, yy)][0], dec[(xx, yy)][1]
                    return (True, mv_, r_)
                return (False, (0, 0), -1)
            a, b = nb(x - 1, y), nb(x, y - 1)
            c = nb(x + 1, y - 1) if (y > 0 and x + 1 < nbx) else nb(x - 1, y - 1)
            best = None
            for r, ref in enumerate(refs):
                pr = _predict_mv(a, b, c, r)
                mv, cost, (ran, paid), _ = refimpl.motion_search_walk(cur, ref, 16 * x, 16 * y, 16, 16, [pr], lam,
                                                                      rescue=kw["rescue"] and not off, **{
                                                                          k: v for k, v in kw.items() if k != "rescue"})
                if ran and n < learn:
                    n += 1
                    big += bool(paid)
                    if n == learn and big * 100 < learn * payoff:
                        off = True
                cost += int(lambda_ref * float(2 * int(np.floor(np.log2(r + 1))) + 1)) if len(refs) > 2 else \
                    int(lambda_ref * float(1 if len(refs) == 2 else 0))
                if best is None or cost < best[2]:
                    best = (r, mv, cost)
            dec[(x, y)] = best
    return dec, (n, big, off)


def latch_clip(cw, ch, seed, shift):
    """Top half: the reference, plus heavy noise in the left half (every rescue there
    runs and none pays; the clean right half runs none); bottom half: a textured field
    moved by `shift`, more than the diamond finds from (0, 0), so the rescue grid pays
    there -- the learning window decides which of the bottom macroblocks still get it."""
    rng = np.random.default_rng(seed)
    m = 24
    base = rng.integers(0, 256, (ch + 2 * m, cw + 2 * m)).astype(np.float64)
    for _ in range(2):
        base = (base + np.roll(base, 1, 0) + np.roll(base, 1, 1) + np.roll(base, (1, 1), (0, 1))) / 4
    base = base * 2.0 - 128
    ref = np.clip(base[m:m + ch, m:m + cw], 0, 255)
    dy, dx = shift
    cur = base[m - dy:m - dy + ch, m - dx:m - dx + cw].copy()
    top = ch // 2
    cur[:top] = ref[:top]
    cur[:top, :cw // 2] += rng.normal(0, 10, (top, cw // 2))
    return np.clip(cur, 0, 255).astype(np.uint8), ref.astype(np.uint8)


@pytest.mark.parametrize("learn,payoff,nrefs", [(4, 50, 1), (14, 50, 1), (13, 10, 1), (4, 0, 1), (10, 50, 2),
                                                (1000, 15, 1)])
def test_decision_rescue_window(learn, payoff, nrefs):
    """orc_motion_search_frame with the rescue == the walk above, on a clip where the
    window matters: it closes on the top half's unpaid rescues (learn 4), after the
    first paid ones of the bottom half (learn 14 and 13: 1 paid, under 50 % and 10 %),
    never closes (payoff 0) or never fills (learn 1000); with two references both
    searches of every macroblock count."""
    rng = np.random.default_rng(learn + payoff)
    cw, ch = 96, 64
    cur, ref0 = latch_clip(cw, ch, 1, (10, -13))
    refs = [ref0] + [np.clip(ref0.astype(int) + rng.integers(-6, 7, ref0.shape), 0, 255).astype(np.uint8)
                     for _ in range(nrefs - 1)]
    lam = 4.0
    kw = dict(fc=True, rescue=True, rescue_range=12, rescue_thresh=0)
    dec, (n, big, off) = decision_walk(cur, refs, lam, 0.0, learn, payoff, kw)
    cfg = O.cfg(cw, ch, 32, 16, 16, 32, 0, lam)
    pls = [O.planes(r, cw, ch, 32) for r in refs]
    ref, mv, cost = O.motion_search_frame(cfg, cur, pls, lambda_ref=0.0, rescue_learn=learn, rescue_payoff=payoff, **kw)
    nbx = cw // 16
    for (x, y), (r, m, c) in dec.items():
        i = y * nbx + x
        assert (int(ref[i]), (int(mv[i, 0]), int(mv[i, 1])), int(cost[i])) == (r, tuple(m), c), (x, y)
    # the window's outcome, and that it shows in the vectors
    _, mv_open, _ = O.motion_search_frame(cfg, cur, pls, lambda_ref=0.0, rescue_learn=1 << 30, **kw)
    print("window", learn, payoff, nrefs, "ran", n, "paid", big, "off", off, flush=True)
    if off:
        assert (mv != mv_open).any(), "a closed window must change the decision on this clip"
    else:
        assert (mv == mv_open).all()
