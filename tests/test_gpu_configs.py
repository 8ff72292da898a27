"""The workload configurations at their real geometry, against the oracle:
  C3  the batched multi-stream search with CIF, 1280x720 and 1920x1088 streams,
      1-4 references, five partition shapes, +-32 (the whole CIF stream, sampled
      blocks of the others), eagerly and after HIP-graph replays;
  C4  the learned-proposal search on a 1920x1088 frame pair with the trained
      network, +-32 fall-back, the +-2 quarter-pel window: every block re-costed;
  C5  the 3840x2160 learned-seeded 3-reference +-64 lookahead: frame 0 whole,
      sampled blocks of frames 1-3, the whole window's propagation."""
import os

import numpy as np
import pytest

from oracle import mvnet_ref as MR
from oracle.oracle import Oracle
from rusty_h264_amd import accel, mvnet, synth

pytestmark = pytest.mark.gpu
O = Oracle(fast=True)   # the -march=x86-64-v3 build of the same source (equal to the -O2 checker, CPU-tested)
NT = min(os.cpu_count() or 1, 16)
LAM = accel.lambda_me(26)
PAD = 32


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    print("device:", d.name(), flush=True)
    yield d
    d.close()


# ------------------------------------------------------------------ C3

SHAPES = [(16, 16), (16, 8), (8, 16), (8, 8), (4, 4)]
C3_STREAMS = [(352, 288, 4, "cif"), (1280, 720, 2, "hd"), (1920, 1088, 1, "hd"), (1280, 720, 3, "hd")]


def _best_part(results, b, nrefs):
    """best_part's reference choice (mb16.rs:4938-4957): cost + (int64)(lambda * ref_bits(r)), first minimum."""
    best = None
    for r, (mv, cost) in enumerate(results):
        c = int(cost[b]) + int(LAM * accel.ref_bits(r, nrefs))
        if best is None or c < best[0]:
            best = (c, r, int(mv[b][0]), int(mv[b][1]))
    return best[1], best[2], best[3], best[0]


def test_c3_batched_real_sizes(dev):
    R = 32
    streams, host = [], []
    for i, (cw, ch, nrefs, gen) in enumerate(C3_STREAMS):
        clip = (synth.cif_clip if gen == "cif" else synth.hd_clip)(nrefs + 1, seed=1000 + i, w=cw, h=ch)
        frames = dev.upload(clip[::-1].copy())          # frame 0 = current, 1.. = references, newest first
        pb = accel.planes_bytes(cw, ch, PAD)
        planes = dev.alloc(nrefs * pb)
        accel.build_planes(dev, int(frames) + cw * ch, nrefs, cw, ch, PAD, out=planes)
        streams.append({"cur": frames, "planes": planes, "cw": cw, "ch": ch, "nrefs": nrefs})
        host.append((clip[nrefs], [clip[nrefs - 1 - r] for r in range(nrefs)]))
    print("C3 streams", [(s["cw"], s["ch"], s["nrefs"]) for s in streams], flush=True)
    bt = accel.Batch(dev, streams, SHAPES, PAD, R, 0, LAM)
    bt.run()
    eager = bt.results()
    rng = np.random.default_rng(3)
    for i, (cw, ch, nrefs, _) in enumerate(C3_STREAMS):
        cur, refs = host[i]
        opl = [O.planes(r, cw, ch, PAD, fused=True) for r in refs]
        for bw, bh in SHAPES:
            cfg = O.cfg(cw, ch, PAD, bw, bh, R, 0, LAM)
            nbx, nb = cw // bw, (cw // bw) * (ch // bh)
            got = eager[i][(bw, bh)]
            if i == 0:      # the whole CIF stream
                per = [O.plain_search_frame(cfg, cur, pl, nthreads=NT) for pl in opl]
                per = [(p["mv"], p["cost"]) for p in per]
                blocks = range(nb)
            else:           # sampled blocks, corners included
                blocks = sorted(set(rng.integers(0, nb, 20).tolist()) | {0, nbx - 1, nb - nbx, nb - 1})
                per = []
                for pl in opl:
                    mv, cost = np.zeros((nb, 2), np.int64), np.zeros(nb, np.int64)
                    for b in blocks:
                        bx, by = (b % nbx) * bw, (b // nbx) * bh
                        imv, _ = O.full_search_block(cfg, cur, pl, bx, by)
                        m, c, _ = O.subpel_refine_block(cfg, cur, pl, bx, by, *imv)
                        mv[b], cost[b] = m, c
                    per.append((mv, cost))
            for b in blocks:
                assert tuple(int(v) for v in got[b]) == _best_part(per, b, nrefs), (i, (bw, bh), b)
    # graph replays, with eager work between them, reproduce the eager results
    bt.capture()
    for rep in range(2):
        junk = dev.upload(np.arange(4096, dtype=np.int32))
        bt.replay()
        dev.sync()
        again = bt.results()
        for i in range(len(C3_STREAMS)):
            for s in SHAPES:
                assert np.array_equal(again[i][s], eager[i][s]), (rep, i, s)
        junk.free()
    bt.close()
    for s in streams:
        s["cur"].free()
        s["planes"].free()


# ------------------------------------------------------------------ C4

def test_c4_learned_1080p(dev):
    """rh264_learned_search at 1920x1088 with the trained network, local range 2, the
    +-32 fall-back above the default threshold (both paths occur), the +-2 quarter-pel
    window: all 8160 blocks re-costed by the oracle around the GPU network's centres."""
    if not os.path.exists(mvnet.WEIGHTS):
        pytest.skip("no trained weights")
    cw, ch, Rf, Rl = 1920, 1088, 32, mvnet.LOCAL_RANGE
    frames = synth.hd_clip(2, seed=4242, w=cw, h=ch)
    clip = dev.upload(frames)
    planes = accel.build_planes(dev, clip, 1, cw, ch, PAD)
    p = accel.search_params(cw, ch, PAD, 16, 16, Rf, mvnet.REFINE_PAT, LAM)
    net = mvnet.MVNet(dev, 1, cw, ch, mvnet.load_weights())
    nbx, nblk = cw // 16, (cw // 16) * (ch // 16)
    t8 = dev.alloc(4 * nblk * 8)
    net.forward(int(clip) + cw * ch, int(clip), cw * ch, 1, cw, ch, None, t8)
    top = dev.download(t8, (nblk, 8), np.int32)
    fb_cost = mvnet.FALLBACK_COST
    mv, cost, fb = mvnet.learned_search(dev, net, p, 1, int(clip) + cw * ch, int(clip), cw * ch, planes,
                                        accel.planes_bytes(cw, ch, PAD), Rl, fb_cost)
    gmv = dev.download(mv, (nblk, 2), np.int32)
    gcost = dev.download(cost, (nblk,), np.int32)
    gfb = dev.download(fb, (nblk,), np.int32)
    net.close()
    loc = O.cfg(cw, ch, PAD, 16, 16, Rl, mvnet.REFINE_PAT, LAM)
    full = O.cfg(cw, ch, PAD, 16, 16, Rf, mvnet.REFINE_PAT, LAM)
    opl = O.planes(frames[0], cw, ch, PAD, fused=True)
    nfb = 0
    for b in range(nblk):
        bx, by = (b % nbx) * 16, (b // nbx) * 16
        best = None
        for c in MR.bin_centre(top[b]):
            m, cst = O.full_search_block(loc, frames[1], opl, bx, by, int(c[0]), int(c[1]))
            if best is None or cst < best[1]:
                best = (m, cst)
        isfb = best[1] > fb_cost
        if isfb:
            best = O.full_search_block(full, frames[1], opl, bx, by)
        nfb += isfb
        m, cst, _ = O.subpel_refine_block(full, frames[1], opl, bx, by, *best[0])
        assert (int(gfb[b]), tuple(gmv[b]), int(gcost[b])) == (int(isfb), m, cst), b
    print(f"C4: {nfb} of {nblk} blocks fell back", flush=True)
    assert 0 < nfb < nblk


# ------------------------------------------------------------------ C5

def test_c5_lookahead_4k(dev):
    """The 3840x2160 learned-seeded lookahead (3 references, +-64, local range 2,
    fall-back above an 8x8 SAD of 1024, proposals scaled by the distance) over a
    4-frame window: frame 0's intra costs whole, 96 sampled blocks of each of frames
    1-3 against the oracle fed the GPU network's own proposals, the propagation of
    the whole window."""
    if not os.path.exists(mvnet.WEIGHTS):
        pytest.skip("no trained weights")
    cw, ch, n, nrefs, R, Rl, fbs = 3840, 2160, 4, 3, 64, 2, 1024
    frames = synth.hd_clip(n, seed=2031, w=cw, h=ch)
    mb_w, mb_h = cw // 16, ch // 16
    nm = mb_w * mb_h
    fbuf = dev.upload(frames)
    net = mvnet.MVNet(dev, n - 1, cw, ch, mvnet.load_weights())
    mt = accel.MbTree(dev, cw, ch, n)
    mt.run_multi_learned(fbuf, net, nrefs=nrefs, search_range=R, local_range=Rl, fallback_sad=fbs)
    got = mt.results()
    # the scaled proposals (the library's default): the network over (j + 1, j), times k
    cmax = ((max(R, 16 + Rl) + 8 + 3) & ~3) - 8 - Rl
    t8 = dev.alloc(4 * (n - 1) * nm * 8)
    net.forward(int(fbuf) + cw * ch, int(fbuf), cw * ch, n - 1, cw, ch, None, t8)
    top = dev.download(t8, (n - 1, nm, 8), np.int32)
    net.close()
    cen = np.zeros((n, nrefs, nm, 8, 2), np.int32)
    for k in range(1, nrefs + 1):
        cen[k:, k - 1, :, :, 0] = np.clip((top[k - 1:] % 17 - 8) * 2 * k, -cmax, cmax)
        cen[k:, k - 1, :, :, 1] = np.clip((top[k - 1:] // 17 - 8) * 2 * k, -cmax, cmax)
    halves = got["half"]
    for f in range(n):
        assert np.array_equal(halves[f], O.downsample2x(frames[f], cw, ch)), f
    intra0, _, _, _ = O.mbtree_costs_multi(halves[:1], mb_w, mb_h, nrefs, R, nthreads=NT)
    assert np.array_equal(got["intra"][0], intra0[0]) and np.array_equal(got["inter"][0], intra0[0])
    rng = np.random.default_rng(5)
    fell = 0
    for f in range(1, n):
        sample = sorted(set(rng.integers(0, nm, 92).tolist()) | {0, mb_w - 1, nm - mb_w, nm - 1})
        for m in sample:
            c, mv, k, mask = O.la_block_learned(halves, cw // 2, ch // 2, f, nrefs, R, Rl, fbs, cen[f, :, m],
                                                (m % mb_w) * 8, (m // mb_w) * 8)
            want = (min(c, int(got["intra"][f, m])), mv, k, mask)
            have = (int(got["inter"][f, m]), tuple(got["mv"][f, m]), int(got["ref"][f, m]), int(got["fb"][f, m]))
            assert have == want, (f, m, have, want)
            fell += mask != 0
    print(f"C5: {fell} sampled blocks used a fall-back search", flush=True)
    assert np.array_equal(got["prop"], O.propagate_multi_i32(got["intra"], got["inter"], got["mv"], got["ref"],
                                                             mb_w, mb_h))
    fbuf.free()
