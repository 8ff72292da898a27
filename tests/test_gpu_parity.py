"""GPU parity: the HIP provider (through the C ABI) vs the CPU oracle on the same
seeded inputs.  Bit-exact for every output (all integer)."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from rusty_h264_amd import accel, synth

pytestmark = pytest.mark.gpu
O = Oracle()
NT = min(os.cpu_count() or 1, 16)
LAM26 = accel.lambda_me(26)


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    print("device:", d.name(), flush=True)
    yield d
    d.close()


# ------------------------------------------------------------ block metrics

@pytest.mark.parametrize("kind", [accel.METRIC_SAD, accel.METRIC_SATD_WELS, accel.METRIC_SATD, accel.METRIC_SATD8])
def test_block_metrics_corpus_and_lcg(dev, kind):
    shapes = accel.SHAPES if kind != accel.METRIC_SATD8 else [s for s in accel.SHAPES if s[0] % 8 == 0 and s[1] % 8 == 0]
    fn = {accel.METRIC_SAD: O.sad, accel.METRIC_SATD_WELS: O.satd_wels, accel.METRIC_SATD: O.satd_px,
          accel.METRIC_SATD8: O.satd8_px}[kind]
    # closed-form corpus (satd_sad.rs:390-399), 96 seeds, 16-stride blocks
    src, ref = [], []
    for seed in range(96):
        a, b = synth.corpus_pair(seed)
        src.append(a.reshape(-1)); ref.append(b.reshape(-1))
    # LCG blocks (satd_sad.rs:459-489), strides 24 / 20 re-laid at stride 16 windows
    src = np.concatenate(src); ref = np.concatenate(ref)
    for w, h in shapes:
        offs = np.arange(96, dtype=np.int64) * 256
        got = accel.block_metric(dev, kind, w, h, src, 16, offs, ref, 16, offs)
        for i in range(96):
            a, b = src[offs[i]:offs[i] + 256], ref[offs[i]:offs[i] + 256]
            want = fn(a, 16, b, 16, w, h)
            assert got[i] == want, (w, h, i)
    buf = synth.lcg_u8(0x777, 24 * 20 * 64)
    so = np.arange(32, dtype=np.int64) * 24 * 20
    ro = so + 32 * 24 * 20
    for w, h in shapes:
        got = accel.block_metric(dev, kind, w, h, buf, 24, so, buf, 24, ro)
        for i in range(32):
            want = fn(buf[so[i]:], 24, buf[ro[i]:], 24, w, h)
            assert got[i] == want


def test_satd_avg_x4(dev):
    rng = np.random.default_rng(12)
    plane = rng.integers(0, 256, 64 * 64, dtype=np.uint8)
    src = rng.integers(0, 256, 16 * 16, dtype=np.uint8)
    for w, h in [(16, 16), (16, 8), (8, 16), (8, 8)]:
        oa = rng.integers(0, 64 * 40, 4)
        ob = rng.integers(0, 64 * 40, 4)
        got = accel.satd_avg_x4(dev, src, 16, plane, oa, ob, 64, w, h)
        for k in range(4):
            assert got[k] == O.satd_avg(src, 16, plane[oa[k]:], plane[ob[k]:], 64, w, h)
        # satd_avg(a, a) == satd_x4 (satd_avg_compare.rs:165-167)
        assert (accel.satd_avg_x4(dev, src, 16, plane, oa, oa, 64, w, h) ==
                accel.satd_x4(dev, src, 16, plane, oa, 64, w, h)).all()
    # extremal planes (satd_avg_compare.rs:209-221)
    z, f = np.zeros(256, np.uint8), np.full(256, 255, np.uint8)
    assert accel.sad_x4(dev, z, 16, f, [0, 0, 0, 0], 16, 16, 16)[0] == 65280


# ------------------------------------------------------------ planes

def _plane_frames(kind, nf, ch, cw, seed):
    if kind == "noise":
        return np.random.default_rng(seed).integers(0, 256, (nf, ch, cw), dtype=np.uint8)
    # extremes: 0/255 checkerboards and stripes drive every filter to its range ends
    yy, xx = np.mgrid[0:ch, 0:cw]
    pats = [((xx + yy) & 1) * 255, (xx & 1) * 255, ((yy >> 1) & 1) * 255, ((xx // 3 + yy // 2) & 1) * 255]
    return np.stack([pats[i % 4] for i in range(nf)]).astype(np.uint8)


@pytest.mark.parametrize("planes_rw", [1, 0])
@pytest.mark.parametrize("cw,ch,pad,nf,kind", [(64, 48, 16, 1, "noise"), (176, 144, 32, 3, "noise"),
                                               (352, 288, 32, 2, "noise"), (48, 32, 4, 2, "noise"),
                                               (1920, 1088, 32, 1, "noise"), (96, 64, 128, 1, "noise"),
                                               (272, 80, 12, 4, "extreme"), (16, 16, 4, 1, "noise")])
def test_planes_bit_exact(dev, cw, ch, pad, nf, kind, planes_rw):
    """Both plane kernels: the register-window one (default) and the LDS-tile one."""
    frames = _plane_frames(kind, nf, ch, cw, cw + pad)
    fb = dev.upload(frames)
    dev.tune("planes_rw", planes_rw)
    try:
        pl = accel.build_planes(dev, fb, nf, cw, ch, pad)
    finally:
        dev.tune("planes_rw", 1)
    pb = accel.planes_bytes(cw, ch, pad)
    got = dev.download(pl, (nf, pb), np.uint8)
    pw, ph = cw + 2 * pad, ch + 2 * pad
    for i in range(nf):
        want = O.planes(frames[i], cw, ch, pad, fused=True)
        g = got[i, :4 * pw * ph].reshape(4, ph, pw)
        if not np.array_equal(g, want):
            msg = []
            for k in range(4):
                ys, xs = np.nonzero(g[k] != want[k])
                if ys.size:
                    msg.append(f"plane {'FHVC'[k]}: {ys.size} px differ, rows {np.unique(ys)[:8]} "
                               f"cols {np.unique(xs)[:8]}, got {g[k][ys[0], xs[0]]} want {want[k][ys[0], xs[0]]}")
            raise AssertionError(f"frame {i}: " + "; ".join(msg))
    fb.free(); pl.free()


def test_predict_blocks_any_mv(dev):
    """Plane reads with clamping == mc_luma (inter.rs:1193-1262) for MVs far outside."""
    cw, ch, pad = 64, 48, 32
    y = np.random.default_rng(13).integers(0, 256, (ch, cw), dtype=np.uint8)
    pl = dev.upload(O.planes(y, cw, ch, pad).reshape(-1))
    rng = np.random.default_rng(14)
    for bw, bh in accel.SHAPES:
        n = 300
        xy = np.stack([rng.integers(0, cw - bw + 1, n), rng.integers(0, ch - bh + 1, n)], 1)
        mv = rng.integers(-500, 501, (n, 2))
        got = accel.predict_blocks(dev, pl, cw, ch, pad, bw, bh, xy, mv)
        for i in range(n):
            want = O.mc_luma(y, cw, ch, int(xy[i, 0]), int(xy[i, 1]), bw, bh, int(mv[i, 0]), int(mv[i, 1]))
            assert np.array_equal(got[i], want), (bw, bh, i)
    pl.free()


# ------------------------------------------------------------ search

def _search_case(dev, frames, cw, ch, pad, bw, bh, R, pat, lam, centres=None, mvps=None, nthreads=NT, hadamard=4):
    nf = frames.shape[0] - 1
    cur = dev.upload(frames[1:])
    refs = dev.upload(frames[:-1])
    pl = accel.build_planes(dev, refs, nf, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, bw, bh, R, pat, lam, hadamard=hadamard)
    ms = accel.MotionSearch(dev, p)
    dc = dev.upload(centres) if centres is not None else None
    dm = dev.upload(mvps) if mvps is not None else None
    ms.run_plain(cur, pl, nf, dc, dm)
    got = ms.results(nf)
    cfg = O.cfg(cw, ch, pad, bw, bh, R, pat, lam, hadamard=hadamard)
    nb = ms.nblk
    for f in range(nf):
        opl = O.planes(frames[f], cw, ch, pad)
        ce = centres[f * nb:(f + 1) * nb] if centres is not None else None
        mp = mvps[f * nb:(f + 1) * nb] if mvps is not None else None
        want = O.plain_search_frame(cfg, frames[f + 1], opl, ce, mp, nthreads=nthreads)
        sl = slice(f * nb, (f + 1) * nb)
        for k in ("int_mv", "int_cost", "mv", "cost"):
            bad = np.nonzero((got[k][sl] != want[k]).reshape(nb, -1).any(1))[0]
            assert bad.size == 0, f"{k} frame {f}: {bad.size} blocks differ, first {bad[:5]}"
    for b in (cur, refs, pl, dc, dm):
        if b is not None:
            b.free()
    return got


@pytest.mark.parametrize("bw,bh", accel.SHAPES)
def test_plain_search_shapes_small(dev, bw, bh):
    frames = synth.cif_clip(3, seed=bw * 10 + bh, w=96, h=64)
    _search_case(dev, frames, 96, 64, 32, bw, bh, 7, 0, LAM26)


@pytest.mark.parametrize("R,pat,lam,had", [(16, 0, LAM26, 4), (16, 2, LAM26, 4), (5, 1, 0.0, 4), (13, 3, 9.0, 4),
                                           (16, -1, LAM26, 4), (16, 4, LAM26, 4), (16, 0, LAM26, 8), (16, 4, LAM26, 8),
                                           (9, 3, 0.0, 8)])
def test_plain_search_cif(dev, R, pat, lam, had):
    """Config C1 geometry: CIF 352x288, 16x16, +-16 full search, qpel SATD refine
    (the 4x4 transform, or the 8x8 one)."""
    frames = synth.cif_clip(4, seed=R + pat + had)
    _search_case(dev, frames, 352, 288, 32, 16, 16, R, pat, lam, hadamard=had)


def test_plain_search_centres_and_predictors_outside_pad(dev):
    """Windows that leave the padded planes (clamped reads) and non-zero rate
    predictors; pad 8 < R so edge blocks take the clamped path."""
    cw, ch = 128, 96
    frames = synth.cif_clip(3, seed=77, w=cw, h=ch)
    nb = (cw // 16) * (ch // 16) * 2
    rng = np.random.default_rng(15)
    centres = (4 * rng.integers(-60, 61, (nb, 2))).astype(np.int32)
    mvps = rng.integers(-300, 301, (nb, 2)).astype(np.int32)
    _search_case(dev, frames, cw, ch, 8, 16, 16, 12, 0, LAM26, centres, mvps)


def test_subpel_refine_from_given_starts(dev):
    cw, ch, pad = 176, 144, 32
    frames = synth.cif_clip(2, seed=5, w=cw, h=ch)
    cur, ref = dev.upload(frames[1]), dev.upload(frames[0])
    pl = accel.build_planes(dev, ref, 1, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, 16, 16, 8, 0, LAM26)
    ms = accel.MotionSearch(dev, p)
    rng = np.random.default_rng(16)
    starts = rng.integers(-200, 201, (ms.nblk, 2)).astype(np.int32)
    ds = dev.upload(starts)
    ms.run_refine(cur, pl, 1, ds)
    got = ms.results(1, with_evals=True)
    cfg = O.cfg(cw, ch, pad, 16, 16, 8, 0, LAM26)
    opl = O.planes(frames[0], cw, ch, pad)
    for b in range(ms.nblk):
        bx, by = (b % (cw // 16)) * 16, (b // (cw // 16)) * 16
        mv, c, ev = O.subpel_refine_block(cfg, frames[1], opl, bx, by, *starts[b])
        assert tuple(got["mv"][b]) == mv and got["cost"][b] == c and got["evals"][b] == ev, b


def _smooth_pair(cw, ch, dx, dy):
    """A smooth reference (long-wavelength sinusoid) and a current frame that is the
    reference moved by (dx, dy) pels: cur(x, y) = ref(x + dx, y + dy)."""
    yy, xx = np.mgrid[0:ch + 64, 0:cw + 64]
    big = np.rint(128 + 90 * np.sin(2 * np.pi * xx / 97.0) * np.cos(2 * np.pi * yy / 83.0)).astype(np.uint8)
    ref = np.ascontiguousarray(big[16:16 + ch, 16:16 + cw])
    cur = np.ascontiguousarray(big[16 + dy:16 + dy + ch, 16 + dx:16 + dx + cw])
    return np.stack([ref, cur])


@pytest.mark.parametrize("pat,refine_mfma,had", [(0, 1, 4), (1, 1, 4), (0, 0, 4), (0, 1, 8)])
def test_subpel_refine_long_walks(dev, pat, refine_mfma, had):
    """Ring walks that travel far from their start (smooth content, lambda 0, start at
    zero, true motion (+7, +3) pels): candidates beyond the window staged around the
    start read the planes in HBM.  Bit-exact with the oracle, evaluation counts included."""
    cw, ch, pad = 176, 144, 32
    frames = _smooth_pair(cw, ch, 7, 3)
    dev.tune("refine_mfma", refine_mfma)
    try:
        cur, ref = dev.upload(frames[1]), dev.upload(frames[0])
        pl = accel.build_planes(dev, ref, 1, cw, ch, pad)
        p = accel.search_params(cw, ch, pad, 16, 16, 8, pat, 0.0, hadamard=had)
        ms = accel.MotionSearch(dev, p)
        starts = np.zeros((ms.nblk, 2), np.int32)
        ds = dev.upload(starts)
        ms.run_refine(cur, pl, 1, ds)
        got = ms.results(1, with_evals=True)
    finally:
        dev.tune("refine_mfma", 1)
    cfg = O.cfg(cw, ch, pad, 16, 16, 8, pat, 0.0, hadamard=had)
    opl = O.planes(frames[0], cw, ch, pad)
    far = 0
    for b in range(ms.nblk):
        bx, by = (b % (cw // 16)) * 16, (b // (cw // 16)) * 16
        mv, c, ev = O.subpel_refine_block(cfg, frames[1], opl, bx, by, 0, 0)
        assert tuple(got["mv"][b]) == mv and got["cost"][b] == c and got["evals"][b] == ev, b
        far += max(abs(mv[0]), abs(mv[1])) > 16
    assert far > ms.nblk // 2  # most walks leave the +-4 pel window


@pytest.mark.parametrize("refine_mfma,had", [(1, 4), (0, 4), (1, 8)])
@pytest.mark.parametrize("bw,bh", [(16, 16), (16, 8), (8, 8), (4, 4)])
def test_qpel_window_refine(dev, bw, bh, refine_mfma, had):
    """subpel_pat 4 (the +-2 quarter-pel window) from given starts: integer-pel
    starts (the window kernel's compile-time path), fractional starts, and starts
    far outside the padded planes (clamped reads), with rate predictors.  mv, cost
    and evaluation count bit-exact with the oracle."""
    cw, ch, pad = 176, 144, 32
    if had == 8 and (bw, bh) != (16, 16):
        pytest.skip("the 8x8 transform refines 16x16 blocks only (other shapes decline)")
    frames = synth.cif_clip(2, seed=bw + bh + refine_mfma, w=cw, h=ch)
    dev.tune("refine_mfma", refine_mfma)
    try:
        cur, ref = dev.upload(frames[1]), dev.upload(frames[0])
        pl = accel.build_planes(dev, ref, 1, cw, ch, pad)
        p = accel.search_params(cw, ch, pad, bw, bh, 8, 4, LAM26, hadamard=had)
        ms = accel.MotionSearch(dev, p)
        rng = np.random.default_rng(bw * bh)
        nb = ms.nblk
        starts = 4 * rng.integers(-12, 13, (nb, 2))
        kind = rng.integers(0, 4, nb)
        starts[kind == 1] += rng.integers(-3, 4, (int((kind == 1).sum()), 2))  # fractional
        starts[kind == 2] = 4 * rng.integers(-70, 71, (int((kind == 2).sum()), 2))  # far outside the pad
        starts = starts.astype(np.int32)
        mvps = rng.integers(-60, 61, (nb, 2)).astype(np.int32)
        ds, dm = dev.upload(starts), dev.upload(mvps)
        ms.run_refine(cur, pl, 1, ds, dm)
        got = ms.results(1, with_evals=True)
    finally:
        dev.tune("refine_mfma", 1)
    cfg = O.cfg(cw, ch, pad, bw, bh, 8, 4, LAM26, hadamard=had)
    opl = O.planes(frames[0], cw, ch, pad)
    nbx = cw // bw
    for b in range(nb):
        bx, by = (b % nbx) * bw, (b // nbx) * bh
        mv, c, ev = O.subpel_refine_block(cfg, frames[1], opl, bx, by, *starts[b], *mvps[b])
        assert tuple(got["mv"][b]) == mv and got["cost"][b] == c and got["evals"][b] == ev == 25, (b, starts[b])
    for x in (cur, ref, pl, ds, dm):
        x.free()


def test_hadamard8_declines_other_shapes(dev):
    """The 8x8 transform refines 16x16 blocks only: other shapes decline (status 1,
    the caller's CPU path), and a transform other than 4 or 8 is an error."""
    cw, ch, pad = 64, 48, 16
    f = synth.cif_clip(2, seed=1, w=cw, h=ch)
    cur, ref = dev.upload(f[1]), dev.upload(f[0])
    pl = accel.build_planes(dev, ref, 1, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, 8, 8, 4, 0, LAM26, hadamard=8)
    ms = accel.MotionSearch(dev, p)
    with pytest.raises(accel.Declined):
        ms.run_plain(cur, pl, 1)
    p = accel.search_params(cw, ch, pad, 16, 16, 4, 0, LAM26, hadamard=6)
    ms = accel.MotionSearch(dev, p)
    with pytest.raises(RuntimeError):
        ms.run_plain(cur, pl, 1)
    for x in (cur, ref, pl):
        x.free()


def test_qpel_window_ties(dev):
    """Ramp ref = 4x, cur = 4x + 1, lambda 0: the five dx = +1 positions tie at cost 0
    and the raster-first (dy = -2) wins; a flat frame keeps the start."""
    cw, ch, pad = 64, 48, 16
    ramp = np.tile((4 * np.arange(cw) % 256).astype(np.uint8), (ch, 1))
    for ref, cur_f, want in ((ramp, ramp + 1, (1, -2)), (np.full((ch, cw), 77, np.uint8),) * 2 + ((0, 0),)):
        cur, refd = dev.upload(cur_f), dev.upload(ref)
        pl = accel.build_planes(dev, refd, 1, cw, ch, pad)
        p = accel.search_params(cw, ch, pad, 16, 16, 4, 4, 0.0)
        ms = accel.MotionSearch(dev, p)
        ds = dev.upload(np.zeros((ms.nblk, 2), np.int32))
        ms.run_refine(cur, pl, 1, ds)
        got = ms.results(1, with_evals=True)
        # the interior column of blocks (x = 16, 32) sees an unwrapped ramp
        for b in (1, 2, 5, 6):
            assert tuple(got["mv"][b]) == want and got["cost"][b] == 0, b
        for x in (cur, refd, pl, ds):
            x.free()


@pytest.mark.parametrize("fs_kernel", [0, 1, 2])
def test_full_search_kernel_variants(dev, fs_kernel):
    """Every integer-search kernel variant gives the oracle's result (CIF, +-16, and a
    window that leaves the padded planes)."""
    dev.tune("fs_kernel", fs_kernel)
    try:
        frames = synth.cif_clip(3, seed=40 + fs_kernel)
        _search_case(dev, frames, 352, 288, 32, 16, 16, 16, 0, LAM26)
        nb = 22 * 18 * 2
        centres = (4 * np.random.default_rng(fs_kernel).integers(-40, 41, (nb, 2))).astype(np.int32)
        _search_case(dev, frames, 352, 288, 32, 16, 16, 16, 2, LAM26, centres)
        for bw, bh in [(8, 8), (4, 4), (16, 8)]:
            _search_case(dev, frames[:2], 352, 288, 32, bw, bh, 9, 0, LAM26)
    finally:
        dev.tune("fs_kernel", 0)


def test_plain_search_1080p_frame_pair(dev):
    """Config C2 geometry (1920x1088, +-32, 4x4-Hadamard qpel refine) on one frame pair."""
    frames = synth.hd_clip(2, seed=3)
    _search_case(dev, frames, 1920, 1088, 32, 16, 16, 32, 0, LAM26)


def test_plain_search_batch_equals_single(dev):
    """One launch over several frames == one launch per frame."""
    cw, ch, pad = 352, 288, 32
    frames = synth.cif_clip(5, seed=9)
    p = accel.search_params(cw, ch, pad, 16, 16, 16, 0, LAM26)
    cur, refs = dev.upload(frames[1:]), dev.upload(frames[:-1])
    pl = accel.build_planes(dev, refs, 4, cw, ch, pad)
    ms = accel.MotionSearch(dev, p)
    ms.run_plain(cur, pl, 4)
    batch = ms.results(4)
    pb = accel.planes_bytes(cw, ch, pad)
    for f in range(4):
        ms1 = accel.MotionSearch(dev, p)
        ms1.run_plain(int(cur) + f * cw * ch, int(pl) + f * pb, 1)
        one = ms1.results(1)
        sl = slice(f * ms.nblk, (f + 1) * ms.nblk)
        assert np.array_equal(one["mv"], batch["mv"][sl]) and np.array_equal(one["cost"], batch["cost"][sl])


@pytest.mark.parametrize("npreds,steps,sadfp,snap,pat,had", [(1, (16, 8, 4), False, True, 0, 4),
                                                             (3, (16, 8, 4), True, True, 0, 4),
                                                             (0, (64, 32, 16, 8, 4), False, False, 2, 4),
                                                             (2, (8, 4), True, False, 1, 4),
                                                             (2, (16, 8, 4), False, True, 4, 4),
                                                             (1, (16, 8, 4), True, True, -1, 4),
                                                             (2, (16, 8, 4), False, True, 0, 8)])
def test_motion_search_reference_mode(dev, npreds, steps, sadfp, snap, pat, had):
    """The reference-mode motion search (seeds, integer snap, diamond rungs, the B2
    SAD phase, sub-pel stage, seed kept if cheaper) == the oracle's restatement for
    every block, with fractional and far predictors."""
    cw, ch, pad = 176, 144, 32
    frames = synth.cif_clip(2, seed=npreds + len(steps) + pat, w=cw, h=ch)
    cur, ref = dev.upload(frames[1]), dev.upload(frames[0])
    pl = accel.build_planes(dev, ref, 1, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, 16, 16, 16, pat, LAM26, hadamard=had)
    nb = (cw // 16) * (ch // 16)
    rng = np.random.default_rng(npreds * 7 + pat)
    preds = rng.integers(-90, 91, (nb, max(npreds, 1), 2)).astype(np.int32)
    dp = dev.upload(preds)
    mv, cost = accel.motion_search_ref(dev, p, 1, cur, pl, dp, npreds, steps=steps, sadfp=sadfp, snap=snap)
    cfg = O.cfg(cw, ch, pad, 16, 16, 16, pat, LAM26, hadamard=had)
    opl = O.planes(frames[0], cw, ch, pad)
    for b in range(nb):
        bx, by = (b % (cw // 16)) * 16, (b // (cw // 16)) * 16
        want = O.motion_search_ref(cfg, frames[1], opl, bx, by, preds[b, :npreds], steps=steps, sadfp=sadfp, snap=snap)
        assert (tuple(mv[b]), int(cost[b])) == want, (b, preds[b, :npreds])
    for x in (cur, ref, pl, dp):
        x.free()


@pytest.mark.parametrize("cw,ch,nf,nrefs,lam_ref,steps,sadfp,snap,pat,had", [
    (176, 144, 3, 1, 0.0, (16, 8, 4), False, True, 0, 4),
    (352, 288, 1, 1, 0.0, (16, 8, 4), True, True, 4, 4),
    (16, 96, 2, 1, 0.0, (64, 32, 16, 8, 4), False, False, 2, 4),
    (320, 16, 2, 1, 0.0, (8, 4), True, False, 1, 4),
    (176, 144, 2, 1, 0.0, (16, 8, 4), False, True, 0, 8),
    (176, 144, 2, 2, 6.4, (16, 8, 4), False, True, 0, 4),
    (176, 144, 2, 3, 3.2, (16, 8, 4), True, True, 4, 4),
    (64, 80, 3, 4, 0.0, (16, 8, 4), False, True, 0, 8)])
def test_motion_search_wavefront(dev, cw, ch, nf, nrefs, lam_ref, steps, sadfp, snap, pat, had):
    """Every macroblock decided in macroblock order with its best_part result over
    nrefs references, predict_mv of its decided neighbours per reference (one launch
    per anti-diagonal, all frames at once) == the oracle's raster-order frame
    decision, one-column and one-row frames included."""
    pad = 32
    clip = synth.cif_clip(nf + nrefs, seed=cw + nf + pat + nrefs, w=cw, h=ch)
    cur, refs = dev.upload(clip[nrefs:]), dev.upload(clip[:-1])
    pl = accel.build_planes(dev, refs, nf + nrefs - 1, cw, ch, pad)
    p = accel.search_params(cw, ch, pad, 16, 16, 16, pat, LAM26, hadamard=had)
    nb = (cw // 16) * (ch // 16)
    cfg = O.cfg(cw, ch, pad, 16, 16, 16, pat, LAM26, hadamard=had)
    opl = [O.planes(clip[k], cw, ch, pad) for k in range(nf + nrefs - 1)]
    want = [O.motion_search_frame(cfg, clip[f + nrefs], [opl[f + nrefs - 1 - r] for r in range(nrefs)],
                                  lambda_ref=lam_ref, steps=steps, sadfp=sadfp, snap=snap) for f in range(nf)]
    if nrefs > 1:
        assert len(set(want[0][0].tolist())) > 1  # the case exercises the reference choice
    try:
        for mode in (0, 1):  # one launch per anti-diagonal; fixed-point passes
            dev.tune("decide_mode", mode)
            print(f"wavefront {cw}x{ch} frames {nf} refs {nrefs} steps {steps} pat {pat} had {had} mode {mode}",
                  flush=True)
            out = accel.motion_search_wavefront(dev, p, nf, cur, pl, nrefs=nrefs, lambda_ref=lam_ref, steps=steps,
                                                sadfp=sadfp, snap=snap)
            ref = dev.download(out[0], (nf * nb,), np.int32)
            mv = dev.download(out[1], (nf * nb, 2), np.int32)
            cost = dev.download(out[2], (nf * nb,), np.int32)
            for x in out:
                x.free()
            for f in range(nf):
                w_ref, w_mv, w_c = want[f]
                sl = slice(f * nb, (f + 1) * nb)
                bad = np.nonzero((ref[sl] != w_ref) | (mv[sl] != w_mv).any(1) | (cost[sl] != w_c))[0]
                assert bad.size == 0, (mode, f, bad[:8], ref[sl][bad[:4]], w_ref[bad[:4]], mv[sl][bad[:4]],
                                       w_mv[bad[:4]])
    finally:
        dev.tune("decide_mode", 2)
    for x in (cur, refs, pl):
        x.free()


@pytest.mark.parametrize("source", ["decision", "random"])
def test_decision_predict(dev, source):
    """P_Skip vectors and the luma predictions of decided frames (the decision's own
    output, or arbitrary references and vectors up to 150 px outside the frame) ==
    the oracle's mc_luma on the unpadded reference frames."""
    cw, ch, pad, nf, nrefs = 176, 144, 32, 2, 3
    clip = synth.cif_clip(nf + nrefs, seed=77, w=cw, h=ch)
    cur, refs = dev.upload(clip[nrefs:]), dev.upload(clip[:-1])
    pl = accel.build_planes(dev, refs, nf + nrefs - 1, cw, ch, pad)
    nb = (cw // 16) * (ch // 16)
    if source == "decision":
        p = accel.search_params(cw, ch, pad, 16, 16, 16, 0, LAM26)
        ref_d, mv_d, cost_d = accel.motion_search_wavefront(dev, p, nf, cur, pl, nrefs=nrefs, lambda_ref=3.0)
        cost_d.free()
        ref = dev.download(ref_d, (nf * nb,), np.int32)
        mv = dev.download(mv_d, (nf * nb, 2), np.int32)
    else:
        rng = np.random.default_rng(5)
        ref = rng.integers(0, nrefs, nf * nb).astype(np.int32)
        mv = rng.integers(-600, 601, (nf * nb, 2)).astype(np.int32)
        zero = rng.random(nf * nb) < 0.2
        mv[zero], ref[zero] = 0, 0
        ref_d, mv_d = dev.upload(ref), dev.upload(mv)
    print(f"decision_predict {source} {cw}x{ch} frames {nf} refs {nrefs}", flush=True)
    smv, pred, sp = accel.decision_predict(dev, cw, ch, pad, nf, nrefs, pl, ref_d, mv_d)
    for f in range(nf):
        sl = slice(f * nb, (f + 1) * nb)
        w_smv, w_pred, w_sp = O.decision_predict(cw, ch, [clip[f + nrefs - 1 - r] for r in range(nrefs)], ref[sl],
                                                 mv[sl])
        assert np.array_equal(smv[sl], w_smv), f
        assert np.array_equal(pred[f], w_pred), (f, np.argwhere(pred[f] != w_pred)[:4])
        assert np.array_equal(sp[f], w_sp), (f, np.argwhere(sp[f] != w_sp)[:4])
    for x in (cur, refs, pl, ref_d, mv_d):
        x.free()


def test_decision_entry_points_decline_and_reject(dev):
    """Shapes and sizes outside the encoder-order decision's and the predictions' scope
    decline (the caller keeps its own path); malformed arguments are errors; nothing is
    launched in either case."""
    cw, ch, pad = 64, 48, 32
    clip = synth.cif_clip(2, seed=1, w=cw, h=ch)
    cur, refs = dev.upload(clip[1:]), dev.upload(clip[:1])
    pl = accel.build_planes(dev, refs, 1, cw, ch, pad)
    p16 = accel.search_params(cw, ch, pad, 16, 16, 16, 0, LAM26)
    p8 = accel.search_params(cw, ch, pad, 8, 8, 16, 0, LAM26)
    with pytest.raises(accel.Declined):
        accel.motion_search_wavefront(dev, p8, 1, cur, pl)
    with pytest.raises(accel.Declined):
        accel.motion_search_wavefront(dev, p16, 1, cur, pl, nrefs=5)
    with pytest.raises(RuntimeError):
        accel.motion_search_wavefront(dev, p16, 1, cur, pl, steps=(0,))
    with pytest.raises(RuntimeError):
        accel.motion_search_wavefront(dev, p16, 1, cur, pl, lambda_ref=-1.0)
    ref_d, mv_d, cost_d = accel.motion_search_wavefront(dev, p16, 1, cur, pl)
    with pytest.raises(accel.Declined):
        accel.decision_predict(dev, cw + 8, ch, pad, 1, 1, pl, ref_d, mv_d)
    with pytest.raises(accel.Declined):
        accel.decision_predict(dev, cw, ch, 3, 1, 1, pl, ref_d, mv_d)
    with pytest.raises(accel.Declined):
        accel.decision_predict(dev, cw, ch, pad, 1, 0, pl, ref_d, mv_d)
    with pytest.raises(RuntimeError):
        accel.decision_predict(dev, cw, ch, pad, 1, 1, pl, ref_d, mv_d, planes_stride=16)
    for x in (cur, refs, pl, ref_d, mv_d, cost_d):
        x.free()
