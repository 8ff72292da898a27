"""GPU parity of the reference-mode motion search with every knob: the fixed-centre
diamond passes, the mv-cost rate tables, the Fast preset's SAD cost, every partition
shape, the me_wide rescue grid and its events (rh264_motion_search against the
oracle's orc_motion_search_ex, itself pinned by the trace walkers), and the
encoder-order decision under the reference's default knobs."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rusty_h264_amd import accel, synth

pytestmark = pytest.mark.gpu
O = Oracle()
LAM26 = accel.lambda_me(26)


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    print("device:", d.name(), flush=True)
    yield d
    d.close()


SMOOTH = None


def _smooth():
    global SMOOTH
    if SMOOTH is None:
        SMOOTH = accel.mv_cost_smooth()
    return SMOOTH


def test_smooth_table_matches_the_textbook_form():
    """The library's smooth rate table == round(4 (2 log2(a+1) + 0.718 + [a>0])) (host code)."""
    from tests import refimpl
    assert np.array_equal(accel.mv_cost_smooth(), refimpl.smooth_mv_table())


# (name, shape, knobs, sub-pel pattern, predictors per block)
CASES = [
    ("default-quality", (16, 16), dict(fc=True), 0, 1),
    ("fc-b2", (16, 16), dict(fc=True, sadfp=True), 0, 3),
    ("fc-smooth", (16, 16), dict(fc=True, mv_tab="smooth"), 0, 1),
    ("fc-b2-smooth", (16, 16), dict(fc=True, sadfp=True, mv_tab="smooth"), 2, 2),
    ("16x8", (16, 8), dict(fc=True, mv_tab="smooth"), 0, 2),
    ("8x16", (8, 16), dict(fc=True, sadfp=True), 1, 2),
    ("8x8", (8, 8), dict(fc=True), 0, 3),
    ("8x4", (8, 4), dict(fc=True, steps=(4,)), 2, 3),
    ("4x8", (4, 8), dict(steps=(4,), mv_tab="smooth"), 2, 2),
    ("4x4", (4, 4), dict(fc=True, steps=(4,)), 2, 3),
    ("fast", (16, 16), dict(sad_only=True, steps=(16, 4)), -1, 1),
    ("fast-subpel-8x8", (8, 8), dict(sad_only=True, steps=(16, 4), mv_tab="smooth"), 2, 1),
    ("window", (16, 16), dict(fc=True), 4, 2),
    ("rescue", (16, 16), dict(fc=True, rescue=True, rescue_thresh=1), 0, 1),
    ("rescue-smooth-b2", (16, 16), dict(fc=True, sadfp=True, rescue=True, rescue_thresh=0, mv_tab="smooth"), 2, 1),
    ("rescue-flat-16x8", (16, 8), dict(rescue=True, rescue_range=12, rescue_thresh=0, rescue_fast=False,
                                       rescue_var=600), 0, 2),
    ("rescue-8x8", (8, 8), dict(fc=True, rescue=True, rescue_range=16, rescue_thresh=2), 3, 1),
]


@pytest.mark.parametrize("name,shape,kw,pat,npreds", CASES, ids=[c[0] for c in CASES])
def test_motion_search_knobs(dev, name, shape, kw, pat, npreds):
    """Every block of a CIF frame, fractional and far predictors, bit-exact (vector,
    cost, rescue events) against the oracle."""
    cw, ch, pad = 352, 288, 32
    bw, bh = shape
    kw = dict(kw)
    if kw.get("mv_tab") == "smooth":
        kw["mv_tab"] = _smooth()
    frames = synth.cif_clip(2, seed=len(name) + 3 * pat, w=cw, h=ch)
    cur, ref = dev.upload(frames[1]), dev.upload(frames[0])
    pl = accel.build_planes(dev, ref, 1, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, bw, bh, 16, pat, LAM26)
    nb = (cw // bw) * (ch // bh)
    rng = np.random.default_rng(len(name))
    preds = rng.integers(-60, 61, (nb, npreds, 2)).astype(np.int32)
    preds[::3, 0] &= ~3  # a third of the rate predictors full-pel
    dp = dev.upload(preds)
    print("case", name, shape, flush=True)
    mv, cost, ev = accel.motion_search_ref(dev, p, 1, cur, pl, dp, npreds, events=True, lam=LAM26,
                                           lam_fp=LAM26 * 0.5, **kw)
    cfg = O.cfg(cw, ch, pad, bw, bh, 16, pat, LAM26)
    opl = O.planes(frames[0], cw, ch, pad)
    fired = 0
    for b in range(nb):
        bx, by = (b % (cw // bw)) * bw, (b // (cw // bw)) * bh
        wmv, wc, wev, _ = O.motion_search_trace(cfg, frames[1], opl, bx, by, preds[b], cap=0, **kw)
        assert (tuple(mv[b]), int(cost[b]), int(ev[b])) == (wmv, wc, wev[0] | (wev[1] << 1)), (b, preds[b])
        fired += wev[0]
    if kw.get("rescue"):
        assert fired > 0
    for x in (cur, ref, pl, dp):
        x.free()


def test_motion_search_declines_and_rejects(dev):
    cw, ch, pad = 64, 48, 32
    frames = synth.cif_clip(2, seed=5, w=cw, h=ch)
    cur, ref = dev.upload(frames[1]), dev.upload(frames[0])
    pl = accel.build_planes(dev, ref, 1, cw, ch, pad)
    # the 8x8 transform covers 16x16 only
    with pytest.raises(accel.Declined):
        accel.motion_search_ref(dev, accel.search_params(cw, ch, pad, 8, 8, 16, 0, LAM26, hadamard=8), 1, cur, pl)
    # a table without lambdas in range, a flag that is not 0/1
    with pytest.raises(RuntimeError):
        accel.motion_search_ref(dev, accel.search_params(cw, ch, pad, 16, 16, 16, 0, LAM26), 1, cur, pl,
                                mv_tab=_smooth(), lam=-1.0)
    with pytest.raises(RuntimeError):
        accel.motion_search_ref(dev, accel.search_params(cw, ch, pad, 16, 16, 16, 0, LAM26), 1, cur, pl, fc=2)
    # a learning window's payoff is a percentage
    with pytest.raises(RuntimeError):
        accel.motion_search_wavefront(dev, accel.search_params(cw, ch, pad, 16, 16, 16, 0, LAM26), 1, cur, pl,
                                      rescue=True, rescue_payoff=101)
    for x in (cur, ref, pl):
        x.free()


@pytest.mark.parametrize("nrefs,kw,pat,mode", [
    (1, dict(fc=True), 0, 2),
    (1, dict(fc=True, sadfp=True, mv_tab="smooth"), 0, 0),
    (2, dict(fc=True, mv_tab="smooth"), 2, 2),
    (1, dict(sad_only=True, steps=(16, 4)), -1, 1),
    (3, dict(fc=True, sadfp=True), 4, 2)])
def test_decision_default_knobs(dev, nrefs, kw, pat, mode):
    """The encoder-order 16x16 decision with the reference's default search (fixed-
    centre passes; B2 and the smooth rate table as the per-frame dispatch selects
    them; the Fast preset's SAD search) == the oracle's raster-order decision, under
    both schedules."""
    cw, ch, pad, nf = 176, 144, 32, 2
    kw = dict(kw)
    if kw.get("mv_tab") == "smooth":
        kw["mv_tab"] = _smooth()
    clip = synth.cif_clip(nf + nrefs, seed=11 + nrefs + pat, w=cw, h=ch)
    cur, refs = dev.upload(clip[nrefs:]), dev.upload(clip[:-1])
    pl = accel.build_planes(dev, refs, nf + nrefs - 1, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, 16, 16, 16, pat, LAM26)
    dev.tune("decide_mode", mode)
    lam_ref = 3.2 if nrefs > 1 else 0.0
    out = accel.motion_search_wavefront(dev, p, nf, cur, pl, nrefs=nrefs, lambda_ref=lam_ref, lam=LAM26,
                                        lam_fp=LAM26 * 0.5, **kw)
    nb = (cw // 16) * (ch // 16)
    ref = dev.download(out[0], (nf, nb), np.int32)
    mv = dev.download(out[1], (nf, nb, 2), np.int32)
    cost = dev.download(out[2], (nf, nb), np.int32)
    dev.tune("decide_mode", 2)
    cfg = O.cfg(cw, ch, pad, 16, 16, 16, pat, LAM26)
    opl = [O.planes(clip[i], cw, ch, pad) for i in range(nf + nrefs - 1)]
    for f in range(nf):
        pls = [opl[f + nrefs - 1 - r] for r in range(nrefs)]
        wr, wm, wc = O.motion_search_frame(cfg, clip[nrefs + f], pls, lambda_ref=lam_ref, **kw)
        assert np.array_equal(ref[f], wr) and np.array_equal(mv[f], wm) and np.array_equal(cost[f], wc), f
    for x in tuple(out) + (cur, refs, pl):
        x.free()


@pytest.mark.parametrize("learn,payoff,nrefs", [(4, 50, 1), (14, 50, 1), (9, 20, 1), (10, 30, 1), (1000, 15, 1),
                                                (10, 50, 2), (6, 15, 3)])
def test_decision_rescue_window(dev, learn, payoff, nrefs):
    """The encoder-order decision with the me_wide rescue and its learning window ==
    the oracle's raster decision, on two frames of the designed clip (a noisy static
    top half whose rescues do not pay, a far-moving bottom half where they do) in one
    call, each frame with its own window; from 1 to 3 references."""
    from tests.test_oracle_walkers import latch_clip
    cw, ch, pad, nf = 96, 64, 32, 2
    kw = dict(fc=True, rescue=True, rescue_range=12, rescue_thresh=0, rescue_learn=learn, rescue_payoff=payoff)
    rng = np.random.default_rng(nrefs)
    curs, frames = [], []
    for f in range(nf):
        cur, ref0 = latch_clip(cw, ch, 1 + f, (10, -13) if f == 0 else (-11, 9))
        older = [np.clip(ref0.astype(int) + rng.integers(-6, 7, ref0.shape), 0, 255).astype(np.uint8)
                 for _ in range(nrefs - 1)]
        curs.append(cur)
        frames.append(older[::-1] + [ref0])  # oldest first: reference r = frames[f][nrefs - 1 - r]
    # frame f's plane sets f .. f + nrefs - 1 must be its references: one plane set per
    # (frame, reference) slot, laid out so that set f + nrefs - 1 - r is reference r
    print("rescue window", learn, payoff, nrefs, flush=True)
    lam = 4.0
    p = accel.search_params(cw, ch, pad, 16, 16, 32, 0, lam)
    cfg = O.cfg(cw, ch, pad, 16, 16, 32, 0, lam)
    nb = cw // 16 * (ch // 16)
    for f in range(nf):  # one frame per call keeps the plane layout simple; then both at once below
        planes = np.stack(frames[f])
        pl_in = dev.upload(planes)
        pl = accel.build_planes(dev, pl_in, nrefs, cw, ch, pad)
        cur = dev.upload(curs[f])
        out = accel.motion_search_wavefront(dev, p, 1, cur, pl, nrefs=nrefs, lambda_ref=0.0, **kw)
        ref = dev.download(out[0], (nb,), np.int32)
        mv = dev.download(out[1], (nb, 2), np.int32)
        cost = dev.download(out[2], (nb,), np.int32)
        pls = [O.planes(frames[f][nrefs - 1 - r], cw, ch, pad) for r in range(nrefs)]
        wr, wm, wc = O.motion_search_frame(cfg, curs[f], pls, lambda_ref=0.0, **kw)
        assert np.array_equal(ref, wr) and np.array_equal(mv, wm) and np.array_equal(cost, wc), f
        for x in tuple(out) + (cur, pl, pl_in):
            x.free()
    # both frames in one call (one reference): a consecutive clip whose frame 1 is
    # frame 0's current picture, so each frame keeps its own window
    if nrefs == 1:
        clip = np.stack([frames[0][0], curs[0], curs[1]])
        cur = dev.upload(clip[1:])
        refs = dev.upload(clip[:-1])
        pl = accel.build_planes(dev, refs, 2, cw, ch, pad)
        out = accel.motion_search_wavefront(dev, p, 2, cur, pl, nrefs=1, lambda_ref=0.0, **kw)
        mv = dev.download(out[1], (2, nb, 2), np.int32)
        cost = dev.download(out[2], (2, nb), np.int32)
        for f in range(2):
            _, wm, wc = O.motion_search_frame(cfg, clip[f + 1], [O.planes(clip[f], cw, ch, pad)], **kw)
            assert np.array_equal(mv[f], wm) and np.array_equal(cost[f], wc), ("batch", f)
        for x in tuple(out) + (cur, refs, pl):
            x.free()


def test_decision_rescue_cif(dev):
    """The decision with the reference's Quality-preset rescue knobs (me_range 24,
    me_rescue 3, me_fast, a 40-search window at 15 %) on a CIF clip, two references."""
    cw, ch, pad, nf, nrefs = 352, 288, 32, 2, 2
    clip = synth.cif_clip(nf + nrefs, seed=77, w=cw, h=ch)
    cur, refs = dev.upload(clip[nrefs:]), dev.upload(clip[:-1])
    pl = accel.build_planes(dev, refs, nf + nrefs - 1, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, 16, 16, 16, 0, LAM26)
    kw = dict(fc=True, rescue=True, lam=LAM26, lam_fp=LAM26 * 0.5)
    print("rescue cif", flush=True)
    out = accel.motion_search_wavefront(dev, p, nf, cur, pl, nrefs=nrefs, lambda_ref=3.2, **kw)
    nb = (cw // 16) * (ch // 16)
    ref = dev.download(out[0], (nf, nb), np.int32)
    mv = dev.download(out[1], (nf, nb, 2), np.int32)
    cost = dev.download(out[2], (nf, nb), np.int32)
    cfg = O.cfg(cw, ch, pad, 16, 16, 16, 0, LAM26)
    opl = [O.planes(clip[i], cw, ch, pad) for i in range(nf + nrefs - 1)]
    for f in range(nf):
        pls = [opl[f + nrefs - 1 - r] for r in range(nrefs)]
        wr, wm, wc = O.motion_search_frame(cfg, clip[nrefs + f], pls, lambda_ref=3.2, fc=True, rescue=True)
        assert np.array_equal(ref[f], wr) and np.array_equal(mv[f], wm) and np.array_equal(cost[f], wc), f
    for x in tuple(out) + (cur, refs, pl):
        x.free()
