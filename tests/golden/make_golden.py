"""Writes the golden fixtures in tests/golden/ from the CPU oracle.

The reference's tests hold no absolute expected values for this path (they are
all differential: SIMD vs its scalar twin), so these fixtures freeze the oracle's
outputs -- after the oracle has been pinned by tests/test_oracle.py -- on the
reference's own test-input generators (the closed-form corpus of
crates/rusty_h264-accel/src/satd_sad.rs:390-399, the LCG of satd_sad.rs:449-460)
and on small seeded frames.  They guard the oracle against regressions and give
the GPU tests a second, file-held expectation.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402
from rusty_h264_amd import synth  # noqa: E402

SHAPES = [(16, 16), (16, 8), (8, 16), (8, 8), (8, 4), (4, 8), (4, 4)]
LAM26 = (0.85 * 2.0 ** ((26 - 12) / 3.0)) ** 0.5


def block_metrics(o):
    out = {}
    corpus = np.stack([np.stack(synth.corpus_pair(s)) for s in range(96)])  # (96, 2, 16, 16)
    out["corpus"] = corpus
    lcg = synth.lcg_u8(0x777, 32 * (24 * 20 + 20 * 20))
    out["lcg"] = lcg
    for w, h in SHAPES:
        vals = np.array([[o.sad(a, 16, b, 16, w, h), o.satd_wels(a, 16, b, 16, w, h), o.satd_px(a, 16, b, 16, w, h)]
                         for a, b in corpus], np.int64)
        out[f"corpus_{w}x{h}"] = vals
        lv = []
        off = 0
        for _ in range(32):
            a = lcg[off:off + 24 * 20]; off += 24 * 20
            b = lcg[off:off + 20 * 20]; off += 20 * 20
            lv.append([o.sad(a, 24, b, 20, w, h), o.satd_wels(a, 24, b, 20, w, h), o.satd_px(a, 24, b, 20, w, h)])
        out[f"lcg_{w}x{h}"] = np.array(lv, np.int64)
    np.savez_compressed(os.path.join(HERE, "block_metrics.npz"), **out)


def planes(o):
    y = np.random.default_rng(0x5EED).integers(0, 256, (32, 48), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "planes_48x32_pad8.npz"), frame=y, planes=o.planes(y, 48, 32, 8))


def search(o):
    frames = synth.cif_clip(2, seed=7, w=96, h=64)
    out = {"frames": frames}
    pl = o.planes(frames[0], 96, 64, 32)
    for bw, bh in SHAPES:
        for pat in (0, 2):
            r = o.plain_search_frame(o.cfg(96, 64, 32, bw, bh, 7, pat, LAM26), frames[1], pl)
            for k, v in r.items():
                out[f"{bw}x{bh}_p{pat}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "search_96x64.npz"), **out)


def mbtree(o):
    frames = synth.cif_clip(5, seed=11, w=64, h=48)
    halves = [o.downsample2x(f, 64, 48) for f in frames]
    intra, inter, mv = [], [], []
    for i, h in enumerate(halves):
        a, b, m = o.frame_costs_half(h, halves[i - 1] if i else None, 4, 3)
        intra.append(a); inter.append(b); mv.append(m)
    intra, inter, mv = np.stack(intra), np.stack(inter), np.stack(mv)
    np.savez_compressed(os.path.join(HERE, "mbtree_64x48.npz"), frames=frames, intra=intra, inter=inter, mv=mv,
                        prop_f64=o.propagate_f64(intra, inter, mv, 4, 3),
                        prop_i32=o.propagate_i32(intra, inter, mv, 4, 3))


if __name__ == "__main__":
    o = Oracle()
    block_metrics(o)
    planes(o)
    search(o)
    mbtree(o)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
