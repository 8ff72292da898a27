"""GPU parity of the P macroblock decision with partitions in macroblock order
(rh264_partition_decision) against the oracle's raster-order decision
(orc_partition_decision_frame, pinned by the 4x4 motion-grid walk)."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rusty_h264_amd import accel, synth
from tests.test_oracle_walkers import partition_clip

pytestmark = pytest.mark.gpu
O = Oracle(fast=True)
PAD = 32


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    print("device:", d.name(), flush=True)
    yield d
    d.close()


def _lam(qp):
    return accel.lambda_me(qp)


@pytest.mark.parametrize("nrefs,qp,p8x8,kw,pat", [
    (1, 26, True, dict(fc=True), 0),
    (1, 20, False, dict(fc=True, sadfp=True), 2),
    (2, 20, True, dict(fc=True, mv_tab="smooth"), 0),
    (3, 26, True, dict(fc=True, sadfp=True), 4),
    (4, 14, True, dict(sad_only=True, steps=(16, 4)), -1)])
def test_partition_decision_designed_clip(dev, nrefs, qp, p8x8, kw, pat):
    cw, ch = 176, 144
    kw = dict(kw)
    if kw.get("mv_tab") == "smooth":
        kw["mv_tab"] = accel.mv_cost_smooth()
    cur, refs = partition_clip(cw, ch, nrefs, 200 + qp + nrefs)
    # the layout of planes: nrefs plane sets oldest first, frame 0's reference r = set nrefs-1-r
    dref = dev.upload(np.stack(refs[::-1]))
    pl = accel.build_planes(dev, dref, nrefs, cw, ch, PAD)
    dcur = dev.upload(cur)
    lam = _lam(qp)
    gate = accel.split_gate(qp)
    assert gate == O.split_gate(qp)
    p = accel.search_params(cw, ch, PAD, 16, 16, 16, pat, lam)
    print("case", nrefs, qp, p8x8, pat, flush=True)
    out = accel.partition_decision(dev, p, 1, dcur, pl, nrefs=nrefs, lambda_ref=lam, gate=gate, p8x8=p8x8,
                                   lam=lam, lam_fp=lam * 0.5, **kw)
    got = dev.download(out, ((cw // 16) * (ch // 16), 14), np.int32)
    cfg = O.cfg(cw, ch, PAD, 16, 16, 16, pat, lam)
    want = O.partition_decision_frame(cfg, cur, [O.planes(r, cw, ch, PAD) for r in refs], lam, gate, p8x8, **kw)
    bad = np.argwhere((got != want).any(1)).ravel()
    assert bad.size == 0, (bad[:5], got[bad[:2]], want[bad[:2]])
    assert len(set(got[:, 0].tolist())) >= 3
    for x in (dref, pl, dcur, out):
        x.free()


def test_partition_decision_batch_of_frames(dev):
    """Three P frames of a synthetic clip decided together (2 references each) ==
    each frame decided by the oracle."""
    cw, ch, nf, nrefs, qp = 352, 288, 3, 2, 26
    clip = synth.hd_clip(nf + nrefs, seed=77, w=cw, h=ch, nrects=12, vmax=6.0)
    dcur = dev.upload(clip[nrefs:])
    drefs = dev.upload(clip[:-1])
    pl = accel.build_planes(dev, drefs, nf + nrefs - 1, cw, ch, PAD)
    lam, gate = _lam(qp), accel.split_gate(qp)
    p = accel.search_params(cw, ch, PAD, 16, 16, 16, 0, lam)
    out = accel.partition_decision(dev, p, nf, dcur, pl, nrefs=nrefs, lambda_ref=lam, gate=gate, fc=True)
    nb = (cw // 16) * (ch // 16)
    got = dev.download(out, (nf, nb, 14), np.int32)
    cfg = O.cfg(cw, ch, PAD, 16, 16, 16, 0, lam)
    opl = [O.planes(clip[i], cw, ch, PAD) for i in range(nf + nrefs - 1)]
    for f in range(nf):
        pls = [opl[f + nrefs - 1 - r] for r in range(nrefs)]
        want = O.partition_decision_frame(cfg, clip[nrefs + f], pls, lam, gate, True, fc=True)
        assert np.array_equal(got[f], want), (f, np.argwhere((got[f] != want).any(1)).ravel()[:5])
    print("modes", np.bincount(got[..., 0].ravel(), minlength=4), flush=True)
    for x in (dcur, drefs, pl, out):
        x.free()


@pytest.mark.parametrize("learn,payoff,nrefs", [(6, 50, 1), (20, 40, 1), (8, 60, 2), (1000, 15, 2)])
def test_partition_decision_rescue_window(dev, learn, payoff, nrefs):
    """The decision with partitions and the me_wide rescue under the frame's learning
    window (every search in call order) == the oracle, on the designed clip where the
    window changes the decisions."""
    from tests.test_oracle_walkers import latch_clip
    cw, ch, qp = 96, 64, 20
    cur, ref0 = latch_clip(cw, ch, 3, (10, -13))
    rng = np.random.default_rng(nrefs)
    refs = [ref0] + [np.clip(ref0.astype(int) + rng.integers(-6, 7, ref0.shape), 0, 255).astype(np.uint8)
                     for _ in range(nrefs - 1)]
    dref = dev.upload(np.stack(refs[::-1]))
    pl = accel.build_planes(dev, dref, nrefs, cw, ch, PAD)
    dcur = dev.upload(cur)
    lam, gate = 4.0, accel.split_gate(qp)
    p = accel.search_params(cw, ch, PAD, 16, 16, 16, 0, lam)
    kw = dict(fc=True, rescue=True, rescue_range=12, rescue_thresh=0, rescue_learn=learn, rescue_payoff=payoff)
    print("partition rescue window", learn, payoff, nrefs, flush=True)
    out = accel.partition_decision(dev, p, 1, dcur, pl, nrefs=nrefs, lambda_ref=lam, gate=gate, p8x8=True, **kw)
    got = dev.download(out, ((cw // 16) * (ch // 16), 14), np.int32)
    cfg = O.cfg(cw, ch, PAD, 16, 16, 16, 0, lam)
    want = O.partition_decision_frame(cfg, cur, [O.planes(r, cw, ch, PAD) for r in refs], lam, gate, True, **kw)
    bad = np.argwhere((got != want).any(1)).ravel()
    assert bad.size == 0, (bad[:5], got[bad[:2]], want[bad[:2]])
    for x in (dref, pl, dcur, out):
        x.free()
