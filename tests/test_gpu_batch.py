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
