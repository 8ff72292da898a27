"""GPU kernels of the learned proposal vs the numpy reference (BF16 rounding at
the same points; tolerances stated per check), and the learned search's exact
re-costing vs the oracle."""
import os

import numpy as np
import pytest

from oracle import mvnet_ref as R
from oracle.oracle import Oracle
from rusty_h264_amd import accel, mvnet, synth

pytestmark = pytest.mark.gpu
O = Oracle()
NT = min(os.cpu_count() or 1, 16)
LAM = accel.lambda_me(26)


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    yield d
    d.close()


def _pair_batch(n, w, h, seed):
    frames = [synth.hd_clip(2, seed=seed + i, w=w, h=h, nrects=2) for i in range(n)]
    cur = np.stack([f[1] for f in frames])
    ref = np.stack([f[0] for f in frames])
    return cur, ref


def _params(seed):
    p = R.init_params(seed)
    P = R.unpack(p.astype(np.float64))
    for name in ("b1", "b2", "b3", "b4"):
        P[name][:] = 0.05
    return R.pack(P).astype(np.float32)


@pytest.mark.parametrize("w,h", [(128, 96), (208, 160), (400, 432)])
@pytest.mark.parametrize("fused", [2, 1, 0])
def test_forward_logits_and_top8(dev, fused, w, h):
    """Every inference path: the register-resident fused kernel (tiles of 12x24 cells;
    208x160 leaves partial tiles on both axes), the LDS-tile fused kernel (tiles of 8x8
    cells) and the layer-by-layer kernels the training step uses."""
    n = 2
    cur, ref = _pair_batch(n, w, h, 10)
    p = _params(1)
    dev.tune("mvnet_fused", fused)
    try:
        net = mvnet.MVNet(dev, n, w, h, p)
        both = dev.upload(np.stack([cur, ref]))       # [cur frames..., ref frames...]
        ncell = n * (h // 16) * (w // 16)
        lg = dev.alloc(4 * ncell * R.K)
        t8 = dev.alloc(4 * ncell * 8)
        net.forward(int(both), int(both) + n * w * h, w * h, n, w, h, lg, t8)
        got = dev.download(lg, (ncell, R.K), np.float32)
        top = dev.download(t8, (ncell, 8), np.int32)
        net.close()
    finally:
        dev.tune("mvnet_fused", 2)
    want = R.forward(p, cur, ref, bf=True)["logits"].reshape(ncell, R.K)
    # BF16 activations: summation order can flip a rounding; |err| <= 0.02 + 2% of |logit|
    err = np.abs(got - want)
    assert (err <= 0.02 + 0.02 * np.abs(want)).all(), float(err.max())
    # top-8 = bins of the eight largest ranking keys, in order (on the GPU's own logits)
    assert np.array_equal(top, R.top8(got))


def test_backward_gradients(dev):
    n, w, h = 2, 64, 64
    cur, ref = _pair_batch(n, w, h, 20)
    p = _params(2)
    labels = np.random.default_rng(3).integers(0, R.K, (n, h // 16, w // 16)).astype(np.int32)
    net = mvnet.MVNet(dev, n, w, h, p)
    both = dev.upload(np.stack([cur, ref]))
    lab = dev.upload(labels)
    loss = net.backward(int(both), int(both) + n * w * h, w * h, n, w, h, lab)
    g = net.grads()
    lref, gref = R.loss_and_grads(p, cur, ref, labels, bf=True)
    assert abs(loss - lref) <= 1e-3 * abs(lref) + 1e-3
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in R.SIZES])
    for i, (name, _) in enumerate(R.SIZES):
        a, b = g[offs[i]:offs[i + 1]], gref[offs[i]:offs[i + 1]]
        # relative error of the whole tensor (BF16 data gradients): <= 3%
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel <= 0.03, (name, rel)
    # one Adam step changes the parameters in the direction opposite to the gradient sign
    net.adam(1e-3)
    p2 = net.params()
    moved = p2 - p
    big = np.abs(g) > 1e-6
    assert (np.sign(moved[big]) == -np.sign(g[big])).mean() > 0.99
    net.close()


@pytest.mark.parametrize("Rl,cand_quad,fallback_v,pat", [(2, 1, 1, 4), (2, 1, 1, 2), (1, 1, 1, 4), (3, 1, 0, 2),
                                                          (4, 1, 1, 4), (5, 1, 1, 2), (2, 0, 0, 4), (2, 2, 1, 4),
                                                          (1, 2, 1, 2), (3, 2, 1, 4), (4, 2, 0, 4), (2, 2, 2, 4)])
def test_learned_search_recosting_is_exact(dev, Rl, cand_quad, fallback_v, pat):
    """Every step after the network is exact: oracle local searches around the GPU's
    eight centres, first minimum, fall-back full search, then the refinement (the
    +-2 quarter-pel window, pattern 4, or a ring walk).  Run once without fall-back
    and once with the threshold at the median candidate cost so both paths are
    exercised.  Local ranges up to 3 take the quad-SAD candidate kernel (cand_quad
    1) or the v_sad_u8 one, 4 the latter, larger ones the per-candidate integer
    searches; cand_quad 2 is the lane-per-candidate kernel (ranges 1..4); every
    fall-back search variant is covered."""
    dev.tune("cand_quad", cand_quad)
    dev.tune("fallback_v", fallback_v)
    try:
        _learned_recost_case(dev, Rl, pat)
    finally:
        dev.tune("cand_quad", 2)
        dev.tune("fallback_v", 1)


@pytest.mark.parametrize("fallback_v,wgs,Rf", [(1, 1, 16), (2, 3, 16), (1, 2, 32), (0, 2, 16)])
def test_learned_search_fallback_grid_loops(dev, fallback_v, wgs, Rf):
    """fallback_wgs caps the fall-back search's grid: with 1-3 workgroups each one
    loops over several parts of the list (half the blocks at the median threshold);
    the results stay the oracle's."""
    dev.tune("fallback_v", fallback_v)
    dev.tune("fallback_wgs", wgs)
    try:
        _learned_recost_case(dev, 2, 4, Rf=Rf)
    finally:
        dev.tune("fallback_v", 1)
        dev.tune("fallback_wgs", 2048)


def test_learned_search_fallback_beside_refinement(dev):
    """fallback_overlap 1: the fall-back search runs on the context's second stream
    beside the refinement of the other blocks, and its own blocks are refined after
    the join (by default it runs on the context's stream before one refinement of
    every block); the same (oracle) results."""
    dev.tune("fallback_overlap", 1)
    try:
        _learned_recost_case(dev, 2, 4)
    finally:
        dev.tune("fallback_overlap", 0)


def test_learned_search_source_in_registers(dev):
    """cand_lds 0: the lane-per-candidate kernel holds the source rows in registers
    instead of reading them from LDS; the same (oracle) results."""
    dev.tune("cand_lds", 0)
    try:
        _learned_recost_case(dev, 2, 4)
    finally:
        dev.tune("cand_lds", 1)


@pytest.mark.parametrize("fallback_v", [0, 1])
def test_learned_search_windows_leave_the_planes(dev, fallback_v):
    """pad 8 < the +-16 fall-back radius: edge blocks' fall-back windows (and the
    network's centres up to 32 pels out) leave the padded planes and take the clamped
    reads; the results stay the oracle's."""
    dev.tune("fallback_v", fallback_v)
    try:
        _learned_recost_case(dev, 2, 4, pad=8)
    finally:
        dev.tune("fallback_v", 1)


def _learned_recost_case(dev, Rl, pat, pad=32, Rf=16):
    cw, ch = 176, 144
    frames = synth.hd_clip(3, seed=31, w=cw, h=ch, nrects=3)
    nf = 2
    clip = dev.upload(frames)
    planes = accel.build_planes(dev, clip, nf, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, 16, 16, Rf, pat, LAM)
    net = mvnet.MVNet(dev, nf, cw, ch, _params(4))
    nblk = (cw // 16) * (ch // 16)
    t8 = dev.alloc(4 * nf * nblk * 8)
    net.forward(int(clip) + cw * ch, int(clip), cw * ch, nf, cw, ch, None, t8)
    top = dev.download(t8, (nf, nblk, 8), np.int32)
    loc = O.cfg(cw, ch, pad, 16, 16, Rl, pat, LAM)
    full = O.cfg(cw, ch, pad, 16, 16, Rf, pat, LAM)
    opls = [O.planes(frames[f], cw, ch, pad) for f in range(nf)]
    cand = {}
    for f in range(nf):
        for b in range(nblk):
            bx, by = (b % (cw // 16)) * 16, (b // (cw // 16)) * 16
            best = None
            for c in R.bin_centre(top[f, b]):
                m, cst = O.full_search_block(loc, frames[f + 1], opls[f], bx, by, int(c[0]), int(c[1]))
                if best is None or cst < best[1]:
                    best = (m, cst)
            cand[f, b] = best
    med = int(np.median([v[1] for v in cand.values()]))
    for fb_cost in (1 << 30, med):
        mv, cost, fb = mvnet.learned_search(dev, net, p, nf, int(clip) + cw * ch, int(clip), cw * ch, planes,
                                            accel.planes_bytes(cw, ch, pad), Rl, fb_cost)
        gmv = dev.download(mv, (nf, nblk, 2), np.int32)
        gcost = dev.download(cost, (nf, nblk), np.int32)
        gfb = dev.download(fb, (nf, nblk), np.int32)
        nfb = 0
        for f in range(nf):
            for b in range(nblk):
                bx, by = (b % (cw // 16)) * 16, (b // (cw // 16)) * 16
                best = cand[f, b]
                isfb = best[1] > fb_cost
                if isfb:
                    best = O.full_search_block(full, frames[f + 1], opls[f], bx, by)
                nfb += isfb
                m, cst, _ = O.subpel_refine_block(full, frames[f + 1], opls[f], bx, by, *best[0])
                assert gfb[f, b] == int(isfb), (fb_cost, f, b)
                assert tuple(gmv[f, b]) == m and gcost[f, b] == cst, (fb_cost, f, b)
        if fb_cost == med:
            assert 0 < nfb < nf * nblk  # both paths exercised
        else:
            assert nfb == 0
    net.close()
