"""The walker comparisons catch plausible mistakes in the oracle.

Each case plants one mistake in a copy of oracle/me_oracle.c (compiled with gcc
into a temporary directory, on the CPU), then runs the trace comparisons of
test_oracle_walkers.py against the mutated build: at least one must disagree.
The planted mistakes are the ones a restatement of the two search drivers could
plausibly make (ring order, a dropped repeat loop, a truncating snap, a tie rule,
the fixed-centre pricing, the rescue walk's centre, the rescue learning window's
count and latch).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle
from tests import refimpl
from tests.test_oracle_walkers import CASES, _clip, decision_walk, latch_clip, tie_corpus

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

RING = "const int ring[4][2] = {{step, 0}, {-step, 0}, {0, step}, {0, -step}};"
RING_SWAPPED = "const int ring[4][2] = {{-step, 0}, {step, 0}, {0, step}, {0, -step}};"

# (name, text, replacement, which occurrence (0-based), driver)
MUTATIONS = [
    ("diamond ring order", RING, RING_SWAPPED, 0, "ms"),
    ("diamond rung walked once", "            if (!improved) break;\n        }\n    }\n    /* 4.",
     "            break;\n        }\n    }\n    /* 4.", 0, "ms"),
    ("snap truncates", "bx = div_euclid4(bx + 2) * 4;", "bx = ((bx + 2) / 4) * 4;", 0, "ms"),
    ("seed ties move", "if (pc < bc) { bc = pc; bx = preds[2 * i];", "if (pc <= bc) { bc = pc; bx = preds[2 * i];", 0,
     "ms"),
    ("fixed-centre ties move", "if (cc < bcc) { bcc = cc; bi = i; }", "if (cc <= bcc) { bcc = cc; bi = i; }", 0, "ms"),
    ("fixed-centre batch priced as the cost closure", "const double lam_fc = dc->sadfp ? dc->lambda_fp : dc->lambda;",
     "const double lam_fc = dc->sadfp ? dc->lambda_fp : (dc->mv_tab ? dc->lambda * 0.25 : dc->lambda);", 0, "ms"),
    ("fixed-centre pass cascades", "if (fc && (bx & 3) == 0 && (by & 3) == 0) {", "if (0) {", 0, "ms"),
    ("rescue walk keeps its centre", "if (cc < bc) { bc = cc; bx = x; by = y; }\n                }\n            if (events)",
     "if (cc < bc) { bc = cc; gbx = x; gby = y; }\n                }\n            bx = gbx; by = gby;\n"
     "            if (events)", 0, "ms"),
    ("window counts searches that did not rescue", "if (ev[0] && resc_n < dc->rescue_learn) {",
     "if (resc_n < dc->rescue_learn) {", 0, "fr"),
    ("window closes on payoff <=", "resc_big * 100 < dc->rescue_learn * dc->rescue_payoff",
     "resc_big * 100 <= dc->rescue_learn * dc->rescue_payoff", 0, "fr"),
    ("window tallies the closing search's payoff after its check",
     "                    if (ev[1]) resc_big++;\n                    if (resc_n == dc->rescue_learn && resc_big * 100 < "
     "dc->rescue_learn * dc->rescue_payoff)\n                        resc_off = 1;",
     "                    if (resc_n == dc->rescue_learn && resc_big * 100 < dc->rescue_learn * dc->rescue_payoff)\n"
     "                        resc_off = 1;\n                    if (ev[1]) resc_big++;", 0, "fr"),
    ("inter_cost ring order", RING, RING_SWAPPED, 1, "ic"),
    ("inter_cost rung walked once", "                if (!moved) break;", "                break;", 0, "ic"),
    ("inter_cost without the early out", "    if (best != 0) {\n        if (seedx", "    if (1) {\n        if (seedx", 0,
     "ic"),
]


def _mutated(tmp_path, text, repl, which):
    d = tmp_path / "orc"
    d.mkdir()
    for f in ("me_oracle.c", "me_oracle.h"):
        shutil.copy(os.path.join(SRC, f), d / f)
    src = (d / "me_oracle.c").read_text()
    at = -1
    for _ in range(which + 1):
        at = src.index(text, at + 1)
    (d / "me_oracle.c").write_text(src[:at] + repl + src[at + len(text):])
    so = d / "liboracle_mut.so"
    subprocess.run(["gcc", "-std=c11", "-O1", "-fPIC", "-shared", "-fopenmp", "-o", str(so), str(d / "me_oracle.c"),
                    "-lm"], check=True)
    return Oracle(path=str(so))


def _ms_agrees(O):
    cw, ch, pad, lam = 64, 48, 32, 4.65
    for name, (bw, bh), kw, pat in CASES:
        cur, ref = _clip(cw, ch, 7 + len(name), (3, -5))
        pl = O.planes(ref, cw, ch, pad)
        cfg = O.cfg(cw, ch, pad, bw, bh, 16, pat, lam)
        rng = np.random.default_rng(len(name))
        blocks = [(x, y) for y in range(0, ch, bh) for x in range(0, cw, bw)]
        for k, i in enumerate(rng.choice(len(blocks), min(len(blocks), 6), replace=False)):
            lx, ly = blocks[i]
            npreds = k % 4
            preds = [tuple(int(v) for v in rng.integers(-40, 41, 2)) for _ in range(npreds)]
            if npreds and k % 2:
                preds[0] = (4 * int(rng.integers(-6, 7)), 4 * int(rng.integers(-6, 7)))
            got = O.motion_search_trace(cfg, cur, pl, lx, ly, preds, **kw)
            want = refimpl.motion_search_walk(cur, ref, lx, ly, bw, bh, preds, lam, lam * 0.5, subpel_pat=pat, **kw)
            if got != want:
                return False
    for cur, ref, lx, ly, preds, kw, pat in tie_corpus():
        cw, ch = cur.shape[1], cur.shape[0]
        cfg = O.cfg(cw, ch, pad, 16, 16, 16, pat, 0.0, lam_fp=0.0)
        got = O.motion_search_trace(cfg, cur, O.planes(ref, cw, ch, pad), lx, ly, preds, **kw)
        if got != refimpl.motion_search_walk(cur, ref, lx, ly, 16, 16, preds, 0.0, 0.0, subpel_pat=pat, **kw):
            return False
    return True


def _ic_agrees(O):
    cw, ch = 48, 40
    for seed, shift in ((1, (2, 1)), (2, (-5, 7)), (3, (11, -9))):
        cur, ref = _clip(cw, ch, 50 + seed, shift)
        rng = np.random.default_rng(seed)
        for by in range(0, ch, 8):
            for bx in range(0, cw, 8):
                s = tuple(4 * int(v) for v in rng.integers(-5, 6, 2)) if (bx + by) % 16 else (0, 0)
                if O.inter_cost_trace(cur, cw, ch, ref, bx, by, 8, s) != refimpl.inter_cost_walk(cur, ref, bx, by, 8, s):
                    return False
    if O.inter_cost_trace(ref, cw, ch, ref, 8, 8, 8, (12, -4))[2] != refimpl.inter_cost_walk(ref, ref, 8, 8, 8,
                                                                                              (12, -4))[2]:
        return False
    return True


def _fr_agrees(O):
    """The frame decision's learning window against decision_walk (latch_clip)."""
    cw, ch, lam = 96, 64, 4.0
    cur, ref = latch_clip(cw, ch, 1, (10, -13))
    kw = dict(fc=True, rescue=True, rescue_range=12, rescue_thresh=0)
    cfg = O.cfg(cw, ch, 32, 16, 16, 32, 0, lam)
    pl = [O.planes(ref, cw, ch, 32)]
    # (10, 30): 3 of 10 paid, exactly the threshold (stays open); (9, 20): the 9th
    # search's payoff keeps it open
    for learn, payoff in ((4, 50), (14, 50), (13, 10), (12, 0), (16, 7), (10, 30), (9, 20)):
        dec, _ = decision_walk(cur, [ref], lam, 0.0, learn, payoff, kw)
        _, mv, cost = O.motion_search_frame(cfg, cur, pl, rescue_learn=learn, rescue_payoff=payoff, **kw)
        for (x, y), (_, m, c) in dec.items():
            i = y * (cw // 16) + x
            if (int(mv[i, 0]), int(mv[i, 1])) != tuple(m) or int(cost[i]) != c:
                return False
    return True


def test_unmutated_copy_agrees(tmp_path):
    """The copy-and-compile route itself changes nothing."""
    O = _mutated(tmp_path, "/* 1. seeds */", "/* 1. seeds */", 0)
    assert _ms_agrees(O) and _ic_agrees(O) and _fr_agrees(O)


@pytest.mark.parametrize("name,text,repl,which,driver", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_is_caught(tmp_path, name, text, repl, which, driver):
    O = _mutated(tmp_path, text, repl, which)
    agrees = {"ms": _ms_agrees, "ic": _ic_agrees, "fr": _fr_agrees}[driver](O)
    assert not agrees, f"the walkers did not notice: {name}"
