"""Batched multi-stream search: uneven streams, 1..4 references, five partition
shapes; per-stream results equal the oracle's plain search per (shape, reference)
followed by best_part's reference choice.  Also: autotuned variants and HIP-graph
replays return the same results."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from rusty_h264_amd import accel, synth

pytestmark = pytest.mark.gpu
O = Oracle()
NT = min(os.cpu_count() or 1, 16)
LAM = accel.lambda_me(26)
SHAPES = [(16, 16), (16, 8), (8, 16), (8, 8), (4, 4)]
PAD, R = 32, 6


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    yield d
    d.close()


def _streams(dev, specs):
    streams, host = [], []
    for i, (cw, ch, nrefs) in enumerate(specs):
        clip = synth.cif_clip(nrefs + 1, seed=100 + i, w=cw, h=ch)
        cur = clip[nrefs]
        refs = [clip[nrefs - 1 - r] for r in range(nrefs)]  # ref 0 = previous frame
        pl = np.concatenate([O.planes(r, cw, ch, PAD).reshape(-1) for r in refs])
        pb = accel.planes_bytes(cw, ch, PAD)
        pls = np.zeros(nrefs * pb, np.uint8)
        for r in range(nrefs):
            pls[r * pb:r * pb + 4 * (cw + 2 * PAD) * (ch + 2 * PAD)] = pl[r * 4 * (cw + 2 * PAD) * (ch + 2 * PAD):
                                                                         (r + 1) * 4 * (cw + 2 * PAD) * (ch + 2 * PAD)]
        streams.append({"cur": dev.upload(cur), "planes": dev.upload(pls), "cw": cw, "ch": ch, "nrefs": nrefs})
        host.append((cur, refs))
    return streams, host


def _expected(host, cw, ch, shape):
    cur, refs = host
    per = []
    for r in refs:
        cfg = O.cfg(cw, ch, PAD, shape[0], shape[1], R, 0, LAM)
        per.append(O.plain_search_frame(cfg, cur, O.planes(r, cw, ch, PAD), nthreads=NT))
    nb = len(per[0]["cost"])
    out = np.zeros((nb, 4), np.int64)
    for b in range(nb):
        best = None
        for ri, p in enumerate(per):
            c = int(p["cost"][b]) + int(LAM * accel.ref_bits(ri, len(refs)))
            if best is None or c < best[0]:
                best = (c, ri, int(p["mv"][b][0]), int(p["mv"][b][1]))
        out[b] = (best[1], best[2], best[3], best[0])
    return out


SPECS = [(96, 64, 1), (64, 48, 2), (128, 96, 3), (48, 32, 4), (112, 80, 2)]


def test_batch_matches_per_stream_oracle(dev):
    streams, host = _streams(dev, SPECS)
    bt = accel.Batch(dev, streams, SHAPES, PAD, R, 0, LAM)
    bt.run()
    got = bt.results()
    for i, (cw, ch, _) in enumerate(SPECS):
        for shape in SHAPES:
            want = _expected(host[i], cw, ch, shape)
            bad = np.nonzero((got[i][shape] != want).any(1))[0]
            assert bad.size == 0, f"stream {i} shape {shape}: {bad.size} blocks differ, first {bad[:5]}"
    bt.close()


def test_batch_autotune_and_graph_replay(dev):
    streams, host = _streams(dev, SPECS[:3])
    bt = accel.Batch(dev, streams, SHAPES, PAD, R, 0, LAM)
    bt.run()
    ref = bt.results()
    choices = bt.autotune(reps=1)
    assert len(choices) == len(SHAPES)
    bt.run()
    tuned = bt.results()
    for i in range(3):
        for s in SHAPES:
            assert np.array_equal(tuned[i][s], ref[i][s])
    bt.capture()
    for rep in range(3):
        # eager work between replays, then the replay must reproduce the eager result
        scratch = dev.upload(np.arange(1000, dtype=np.int32))
        bt.replay()
        dev.sync()
        got = bt.results()
        for i in range(3):
            for s in SHAPES:
                assert np.array_equal(got[i][s], ref[i][s]), (rep, i, s)
        scratch.free()
    bt.close()


def _mvps(rng, cw, ch, shapes, spread):
    """Per-shape block predictors of one stream, shapes in order (quarter-pel)."""
    per = [rng.integers(-spread, spread + 1, ((cw // bw) * (ch // bh), 2)).astype(np.int32) for bw, bh in shapes]
    return per, np.concatenate([p.reshape(-1) for p in per])


@pytest.mark.parametrize("pad,rng_r", [(32, 6), (8, 12)])
def test_batch_fused_shapes_with_predictors(dev, pad, rng_r):
    """The five macroblock shapes of one batch search in one fused pass (4x4 SADs
    summed into the larger partitions).  With per-shape block predictors, every
    shape's records equal the oracle's plain search per reference plus best_part's
    choice, and equal what single-shape batches (each searched on its own) return.
    pad 8 with range 12 puts the edge macroblocks' windows outside the padded plane
    (clamped reads)."""
    rng = np.random.default_rng(11 + pad)
    specs = [(96, 64, 2), (64, 48, 1), (48, 32, 3)]
    streams, host, preds = [], [], []
    for i, (cw, ch, nrefs) in enumerate(specs):
        clip = synth.cif_clip(nrefs + 1, seed=300 + i, w=cw, h=ch)
        cur = clip[nrefs]
        refs = [clip[nrefs - 1 - r] for r in range(nrefs)]
        pb = accel.planes_bytes(cw, ch, pad)
        pls = np.zeros(nrefs * pb, np.uint8)
        for r in range(nrefs):
            pl = O.planes(refs[r], cw, ch, pad).reshape(-1)
            pls[r * pb:r * pb + pl.size] = pl
        per, flat = _mvps(rng, cw, ch, SHAPES, 90)
        streams.append({"cur": dev.upload(cur), "planes": dev.upload(pls), "cw": cw, "ch": ch, "nrefs": nrefs,
                        "mvps": dev.upload(flat)})
        host.append((cur, refs))
        preds.append(per)
    print("fused batch", pad, rng_r, flush=True)
    bt = accel.Batch(dev, streams, SHAPES, pad, rng_r, 0, LAM)
    assert all(v == 3 for v, _ in bt.autotune(reps=1)), "the five shapes should search fused"
    bt.run()
    fused = bt.results()
    bt.close()
    for si, shape in enumerate(SHAPES):
        single = []
        for i, st in enumerate(streams):
            s1 = dict(st)
            s1["mvps"] = dev.upload(preds[i][si].reshape(-1))
            single.append(s1)
        b1 = accel.Batch(dev, single, [shape], pad, rng_r, 0, LAM)
        b1.run()
        alone = b1.results()
        b1.close()
        for i, (cw, ch, nrefs) in enumerate(specs):
            assert np.array_equal(fused[i][shape], alone[i][shape]), (i, shape)
            cur, refs = host[i]
            cfg = O.cfg(cw, ch, pad, shape[0], shape[1], rng_r, 0, LAM)
            per = []
            for r in refs:
                p = O.plain_search_frame(cfg, cur, O.planes(r, cw, ch, pad), mvps=preds[i][si], nthreads=NT)
                per.append(p)
            want = np.zeros_like(fused[i][shape])
            for b in range(want.shape[0]):
                best = None
                for ri, p in enumerate(per):
                    c = int(p["cost"][b]) + int(LAM * accel.ref_bits(ri, nrefs))
                    if best is None or c < best[0]:
                        best = (c, ri, int(p["mv"][b][0]), int(p["mv"][b][1]))
                want[b] = (best[1], best[2], best[3], best[0])
            bad = np.nonzero((fused[i][shape] != want).any(1))[0]
            assert bad.size == 0, f"pad {pad} stream {i} shape {shape}: {bad.size} blocks differ, first {bad[:5]}"
        for s1 in single:
            s1["mvps"].free()
    for st in streams:
        st["cur"].free()
        st["planes"].free()
        st["mvps"].free()
