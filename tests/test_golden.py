"""The oracle reproduces the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py) exactly.  CPU only."""
import os

import numpy as np

from oracle.oracle import Oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
O = Oracle()
SHAPES = [(16, 16), (16, 8), (8, 16), (8, 8), (8, 4), (4, 8), (4, 4)]
LAM26 = (0.85 * 2.0 ** ((26 - 12) / 3.0)) ** 0.5


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_block_metric_fixtures():
    d = _load("block_metrics.npz")
    corpus, lcg = d["corpus"], d["lcg"]
    for w, h in SHAPES:
        got = np.array([[O.sad(a, 16, b, 16, w, h), O.satd_wels(a, 16, b, 16, w, h), O.satd_px(a, 16, b, 16, w, h)]
                        for a, b in corpus])
        assert np.array_equal(got, d[f"corpus_{w}x{h}"])
        off, lv = 0, []
        for _ in range(32):
            a = lcg[off:off + 480]; off += 480
            b = lcg[off:off + 400]; off += 400
            lv.append([O.sad(a, 24, b, 20, w, h), O.satd_wels(a, 24, b, 20, w, h), O.satd_px(a, 24, b, 20, w, h)])
        assert np.array_equal(np.array(lv), d[f"lcg_{w}x{h}"])


def test_plane_fixture():
    d = _load("planes_48x32_pad8.npz")
    assert np.array_equal(O.planes(d["frame"], 48, 32, 8), d["planes"])
    assert np.array_equal(O.planes(d["frame"], 48, 32, 8, fused=True), d["planes"])


def test_search_fixture():
    d = _load("search_96x64.npz")
    frames = d["frames"]
    pl = O.planes(frames[0], 96, 64, 32)
    for bw, bh in SHAPES:
        for pat in (0, 2):
            r = O.plain_search_frame(O.cfg(96, 64, 32, bw, bh, 7, pat, LAM26), frames[1], pl)
            for k, v in r.items():
                assert np.array_equal(v, d[f"{bw}x{bh}_p{pat}_{k}"]), (bw, bh, pat, k)


def test_mbtree_fixture():
    d = _load("mbtree_64x48.npz")
    halves = [O.downsample2x(f, 64, 48) for f in d["frames"]]
    for i, h in enumerate(halves):
        a, b, m = O.frame_costs_half(h, halves[i - 1] if i else None, 4, 3)
        assert np.array_equal(a, d["intra"][i]) and np.array_equal(b, d["inter"][i]) and np.array_equal(m, d["mv"][i])
    assert np.array_equal(O.propagate_f64(d["intra"], d["inter"], d["mv"], 4, 3), d["prop_f64"])
    assert np.array_equal(O.propagate_i32(d["intra"], d["inter"], d["mv"], 4, 3), d["prop_i32"])
