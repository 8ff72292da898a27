"""GPU parity of the mb-tree lookahead (downsample, intra/inter costs with the
seeded cascade diamond, int32 propagation) against the oracle."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rusty_h264_amd import accel, synth

pytestmark = pytest.mark.gpu
O = Oracle()


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    yield d
    d.close()


def _oracle_gop(frames, cw, ch):
    mb_w, mb_h = cw // 16, ch // 16
    halves = [O.downsample2x(f, cw, ch) for f in frames]
    intra, inter, mv = [], [], []
    for i, h in enumerate(halves):
        a, b, m = O.frame_costs_half(h, halves[i - 1] if i else None, mb_w, mb_h)
        intra.append(a); inter.append(b); mv.append(m)
    intra, inter, mv = np.stack(intra), np.stack(inter), np.stack(mv)
    return np.stack(halves), intra, inter, mv, O.propagate_i32(intra, inter, mv, mb_w, mb_h)


@pytest.mark.parametrize("cw,ch,n,kind", [(64, 48, 4, "cif"), (352, 288, 6, "cif"), (176, 144, 5, "noise"),
                                          (16, 16, 3, "cif"), (640, 368, 4, "hd"), (3840, 2160, 3, "hd")])
def test_mbtree_gop_bit_exact(dev, cw, ch, n, kind):
    if kind == "cif":
        frames = synth.cif_clip(n, seed=cw + n, w=cw, h=ch)
    elif kind == "hd":
        frames = synth.hd_clip(n, seed=5, w=cw, h=ch)
    else:
        frames = np.random.default_rng(cw).integers(0, 256, (n, ch, cw), dtype=np.uint8)
    half, intra, inter, mv, prop = _oracle_gop(frames, cw, ch)
    fb = dev.upload(frames)
    mt = accel.MbTree(dev, cw, ch, n)
    mt.run(fb)
    got = mt.results()
    assert np.array_equal(got["half"], half)
    assert np.array_equal(got["intra"], intra)
    for f in range(n):
        bad = np.nonzero((got["inter"][f] != inter[f]) | (got["mv"][f] != mv[f]).any(1))[0]
        assert bad.size == 0, f"frame {f}: {bad.size} MBs differ, first {bad[:5]}"
    assert np.array_equal(got["prop"], prop)
    fb.free()


def test_mbtree_static_and_flat(dev):
    """Static content: inter 0 everywhere after the IDR; flat frames: intra 1 (the
    max(1, .) floor of intra_cost), inter 1 on the IDR (= intra) and 0 after it."""
    f = synth.cif_clip(1, seed=3, w=128, h=64)[0]
    frames = np.stack([f, f, f])
    fb = dev.upload(frames)
    mt = accel.MbTree(dev, 128, 64, 3)
    mt.run(fb)
    got = mt.results()
    assert (got["inter"][1:] == 0).all() and (got["mv"][1:] == 0).all()
    flat = np.full((2, 64, 128), 77, np.uint8)
    fb2 = dev.upload(flat)
    mt2 = accel.MbTree(dev, 128, 64, 2)
    mt2.run(fb2)
    r2 = mt2.results()
    assert (r2["intra"] == 1).all() and (r2["inter"][0] == 1).all() and (r2["inter"][1] == 0).all()


@pytest.mark.parametrize("cw,ch,n,kind", [(16, 16, 3, "noise"), (64, 48, 4, "noise"), (352, 288, 4, "cif"),
                                          (1920, 1088, 3, "hd")])
def test_rc_complexity_bit_exact(dev, cw, ch, n, kind):
    """lookahead::complexity: IDR and P scores of every frame equal the oracle's
    exactly (frames small enough that every macroblock takes the clamped path,
    and large ones where most take the aligned path)."""
    if kind == "cif":
        frames = synth.cif_clip(n + 1, seed=cw, w=cw, h=ch)
    elif kind == "hd":
        frames = synth.hd_clip(n + 1, seed=9, w=cw, h=ch)
    else:
        frames = np.random.default_rng(cw * ch).integers(0, 256, (n + 1, ch, cw), dtype=np.uint8)
    fb = dev.upload(frames)
    cur, ref = int(fb) + cw * ch, int(fb)
    got_p = accel.rc_complexity(dev, cur, ref, n, cw, ch)
    got_i = accel.rc_complexity(dev, cur, None, n, cw, ch)
    for i in range(n):
        assert got_p[i] == O.rc_complexity(frames[i + 1], frames[i], cw, ch), i
        assert got_i[i] == O.rc_complexity(frames[i + 1], None, cw, ch), i
    fb.free()


def test_rc_complexity_floor(dev):
    """Flat IDR frames and P frames equal to their reference score the floor 1."""
    f = np.full((2, 48, 64), 123, np.uint8)
    fb = dev.upload(f)
    assert (accel.rc_complexity(dev, int(fb), None, 2, 64, 48) == 1.0).all()
    assert (accel.rc_complexity(dev, int(fb) + 64 * 48, int(fb), 1, 64, 48) == 1.0).all()
    fb.free()


@pytest.mark.parametrize("cw,ch,n,nrefs,R,kind", [(64, 48, 5, 3, 8, "noise"), (176, 144, 4, 3, 16, "cif"),
                                                  (352, 288, 4, 3, 64, "cif"), (16, 16, 6, 4, 1, "noise"),
                                                  (640, 368, 3, 2, 24, "hd")])
def test_mbtree_multi_reference_bit_exact(dev, cw, ch, n, nrefs, R, kind):
    """The build-defined multi-reference form (exhaustive +-R per reference, 8x8
    Hadamard costs, cheapest reference) and its propagation == the oracle."""
    if kind == "cif":
        frames = synth.cif_clip(n, seed=cw + R, w=cw, h=ch)
    elif kind == "hd":
        frames = synth.hd_clip(n, seed=11, w=cw, h=ch)
    else:
        frames = np.random.default_rng(cw + n).integers(0, 256, (n, ch, cw), dtype=np.uint8)
    mb_w, mb_h = cw // 16, ch // 16
    halves = np.stack([O.downsample2x(f, cw, ch) for f in frames])
    intra, inter, mv, ref = O.mbtree_costs_multi(halves, mb_w, mb_h, nrefs, R, nthreads=16)
    prop = O.propagate_multi_i32(intra, inter, mv, ref, mb_w, mb_h)
    fb = dev.upload(frames)
    mt = accel.MbTree(dev, cw, ch, n)
    mt.run_multi(fb, nrefs=nrefs, search_range=R)
    got = mt.results()
    assert np.array_equal(got["intra"], intra)
    for f in range(n):
        bad = np.nonzero((got["inter"][f] != inter[f]) | (got["mv"][f] != mv[f]).any(1) | (got["ref"][f] != ref[f]))[0]
        assert bad.size == 0, f"frame {f}: {bad.size} MBs differ, first {bad[:5]}"
    assert np.array_equal(got["prop"], prop)
    fb.free()


@pytest.mark.parametrize("cw,ch,n,nrefs,R,Rl,fb,scaled", [
    (128, 96, 5, 3, 16, 2, 300, 0), (176, 144, 4, 2, 24, 1, 150, 0), (64, 48, 4, 3, 8, 4, -1, 0),
    (352, 288, 3, 3, 64, 2, 400, 0), (128, 96, 5, 3, 16, 2, 300, 1), (352, 288, 4, 3, 64, 2, 400, 1),
    (176, 144, 4, 4, 8, 3, 150, 1)])
def test_mbtree_multi_learned_bit_exact(dev, cw, ch, n, nrefs, R, Rl, fb, scaled):
    """The learned-seeded multi-reference form: the network's proposals for every
    (frame, reference) pair (taken from the GPU network itself), the oracle's proposal
    windows, fall-back and reference choice; costs, vectors, references, fall-back
    masks and the propagation bit-exact.  fb = -1 sends every block to the fall-back.
    scaled: the distance-1 proposals times the distance k, clamped to the padded
    plane's bound P - 8 - Rl (la_scaled 1)."""
    from rusty_h264_amd import mvnet
    frames = synth.hd_clip(n, seed=cw + n, w=cw, h=ch, nrects=3)
    mb_w, mb_h = cw // 16, ch // 16
    nm = mb_w * mb_h
    fbuf = dev.upload(frames)
    net = mvnet.MVNet(dev, n - 1, cw, ch, mvnet.load_weights())
    dev.tune("la_scaled", scaled)
    try:
        mt = accel.MbTree(dev, cw, ch, n)
        mt.run_multi_learned(fbuf, net, nrefs=nrefs, search_range=R, local_range=Rl, fallback_sad=fb)
        got = mt.results()
        # proposals per (frame f, reference k): the network over the pairs (j + k, j), or
        # (scaled) over the pairs (j + 1, j) times k
        cen = np.zeros((n, nrefs, nm, 8, 2), np.int32)
        cmax = ((max(R, 16 + Rl) + 8 + 3) & ~3) - 8 - Rl
        for k in range(1, min(1 if scaled else nrefs, n - 1) + 1):
            t8 = dev.alloc(4 * (n - k) * nm * 8)
            net.forward(int(fbuf) + k * cw * ch, int(fbuf), cw * ch, n - k, cw, ch, None, t8)
            top = dev.download(t8, (n - k, nm, 8), np.int32)
            t8.free()
            for kk in (range(1, min(nrefs, n - 1) + 1) if scaled else (k,)):
                s = kk if scaled else 1
                cen[kk:, kk - 1, :, :, 0] = np.clip((top[kk - k:] % 17 - 8) * 2 * s, -cmax, cmax)
                cen[kk:, kk - 1, :, :, 1] = np.clip((top[kk - k:] // 17 - 8) * 2 * s, -cmax, cmax)
        if scaled:
            assert np.abs(cen).max() > 16  # scaled proposals beyond the network's own range
    finally:
        net.close()
        dev.tune("la_scaled", 1)
    halves = got["half"]
    intra0, _, _, _ = O.mbtree_costs_multi(halves[:1], mb_w, mb_h, nrefs, R)
    assert np.array_equal(got["intra"][0], intra0[0]) and np.array_equal(got["inter"][0], intra0[0])
    fbs = 0
    for f in range(1, n):
        for m in range(nm):
            c, mv, k, mask = O.la_block_learned(halves, cw // 2, ch // 2, f, nrefs, R, Rl, fb, cen[f, :, m],
                                                (m % mb_w) * 8, (m // mb_w) * 8)
            ic = int(got["intra"][f, m])
            want = (min(c, ic), mv, k, mask)
            have = (int(got["inter"][f, m]), tuple(got["mv"][f, m]), int(got["ref"][f, m]), int(got["fb"][f, m]))
            assert have == want, (f, m, have, want)
            fbs += mask != 0
    if fb < 0:
        assert fbs == (n - 1) * nm
    assert np.array_equal(got["prop"], O.propagate_multi_i32(got["intra"], got["inter"], got["mv"], got["ref"],
                                                             mb_w, mb_h))
    fbuf.free()


@pytest.mark.parametrize("kind,strength,res_min", [("cif", 0.9, 0.03), ("hd", 0.9, 0.03), ("cif", 2.0, 0.0)])
def test_mbtree_policy_on_gpu_costs(dev, kind, strength, res_min):
    """The QP-offset policy (rh264_mbtree_qp_offsets, host f64) applied to a GOP's
    costs from the GPU == the oracle's policy on the oracle's costs."""
    cw, ch, n = (352, 288, 8) if kind == "cif" else (1920, 1088, 4)
    frames = synth.cif_clip(n, seed=21, w=cw, h=ch) if kind == "cif" else synth.hd_clip(n, seed=6, w=cw, h=ch)
    _, intra, inter, mv, _ = _oracle_gop(frames, cw, ch)
    fb = dev.upload(frames)
    mt = accel.MbTree(dev, cw, ch, n)
    mt.run(fb)
    got = mt.results()
    fb.free()
    mb_w, mb_h = cw // 16, ch // 16
    shp = (n, mb_h, mb_w)
    off, st = accel.mbtree_qp_offsets(dev.lib, got["intra"].reshape(shp), got["inter"].reshape(shp),
                                      got["mv"].reshape(shp + (2,)), strength, res_min)
    want, wst = O.mbtree_qp_offsets(np.reshape(intra, shp), np.reshape(inter, shp), np.reshape(mv, shp + (2,)),
                                    strength, res_min)
    assert np.array_equal(off, want) and st == wst
    assert off.any()  # the GOP's offsets are not latched off
