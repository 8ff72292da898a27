"""The mb-tree QP-offset policy (gop_qp_offsets, crates/rusty_h264-encoder/src/
mbtree.rs:369-593): the oracle pinned by closed forms, then the library's host
function (rh264_mbtree_qp_offsets) against the oracle, bit-exact in f64."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rusty_h264_amd import _lib, accel

O = Oracle()


def _half_away(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def _two_frame_gop(rng, mb_w=3, mb_h=2):
    """Frame 1 predicts every block perfectly from the co-located block of frame 0
    (inter 0, vector 0): each block of frame 0 receives exactly frame 1's intra."""
    intra = rng.integers(50, 5000, (2, mb_h, mb_w)).astype(np.int32)
    inter = intra.copy()
    inter[1] = 0
    mv = np.zeros((2, mb_h, mb_w, 2), np.int32)
    return intra, inter, mv


def test_policy_closed_form_two_frames():
    rng = np.random.default_rng(3)
    intra, inter, mv = _two_frame_gop(rng)
    s = 0.9
    raw = np.concatenate([-s * np.log2((intra[0] + intra[1]) / intra[0]).ravel(), np.zeros(intra[1].size)])
    cen = raw - raw.mean()
    off, st = O.mbtree_qp_offsets(intra, inter, mv, strength=s, res_min=0.0, sd_min=0.0)
    assert np.array_equal(off.ravel(), np.clip(_half_away(cen), -6, 6).astype(np.int32))
    assert st["residual_frac"] == 0.0 and st["eff_strength"] == s
    assert st["spread"] == pytest.approx(np.sqrt((cen ** 2).mean()) / s, rel=1e-12)
    # the spread is strength-invariant
    for s2 in (0.5, 2.0):
        assert O.mbtree_qp_offsets(intra, inter, mv, strength=s2, res_min=0.0)[1]["spread"] == \
            pytest.approx(st["spread"], rel=1e-12)


def test_policy_latches_and_no_ops():
    rng = np.random.default_rng(4)
    intra, inter, mv = _two_frame_gop(rng)
    # residual fraction 0 < res_min: latched off, all zero
    off, st = O.mbtree_qp_offsets(intra, inter, mv, res_min=0.03)
    assert not off.any() and st["eff_strength"] == 0.0
    # spread below sd_min: all zero
    spread = O.mbtree_qp_offsets(intra, inter, mv, res_min=0.0)[1]["spread"]
    assert not O.mbtree_qp_offsets(intra, inter, mv, res_min=0.0, sd_min=spread * 1.01)[0].any()
    assert O.mbtree_qp_offsets(intra, inter, mv, res_min=0.0, sd_min=spread * 0.99)[0].any()
    # strength <= 0 and nothing propagated (inter == intra) are no-ops
    assert not O.mbtree_qp_offsets(intra, inter, mv, strength=0.0, res_min=0.0)[0].any()
    assert not O.mbtree_qp_offsets(intra, intra, mv, res_min=0.0)[0].any()
    # large propagation clamps at +-6
    big = intra.copy()
    big[0] = 1
    big[1] = 10 ** 6
    assert np.abs(O.mbtree_qp_offsets(big, np.where(np.arange(2)[:, None, None] == 1, 0, big), mv,
                                      res_min=0.0)[0]).max() == 6


def _random_gop(rng, n, mb_w, mb_h):
    intra = rng.integers(1, 20000, (n, mb_h, mb_w)).astype(np.int32)
    inter = (intra * rng.random((n, mb_h, mb_w))).astype(np.int32)
    mv = rng.integers(-300, 301, (n, mb_h, mb_w, 2)).astype(np.int32)
    return intra, inter, mv


@pytest.mark.parametrize("seed,n,mb_w,mb_h,strength,res_min,sd_min", [
    (0, 8, 11, 9, 0.9, 0.03, 0.0), (1, 5, 22, 18, 0.5, 0.0, 0.0), (2, 12, 7, 5, 2.0, 0.03, 0.4),
    (3, 1, 4, 4, 0.9, 0.03, 0.0), (4, 3, 1, 1, 0.9, 0.0, 0.0), (5, 6, 30, 2, 0.9, 0.6, 0.0)])
def test_library_policy_matches_oracle(seed, n, mb_w, mb_h, strength, res_min, sd_min):
    """Host logic of the library: rh264_mbtree_qp_offsets == the oracle, offsets and
    f64 statistics identical, across sizes, strengths and both latches."""
    rng = np.random.default_rng(seed)
    intra, inter, mv = _random_gop(rng, n, mb_w, mb_h)
    want, wst = O.mbtree_qp_offsets(intra, inter, mv, strength, res_min, sd_min)
    got, gst = accel.mbtree_qp_offsets(_lib.load(), intra, inter, mv, strength, res_min, sd_min)
    assert np.array_equal(got, want)
    assert gst == wst


def test_library_policy_rejects_invalid_costs():
    lib = _lib.load()
    intra, inter, mv = _random_gop(np.random.default_rng(9), 2, 3, 3)
    inter[1, 0, 0] = intra[1, 0, 0] + 1
    with pytest.raises(RuntimeError):
        accel.mbtree_qp_offsets(lib, intra, inter, mv)


# ------------------------------------------------------------ the grain latch

def _grain_textbook(cur, ref):
    """grain_signature (signals.rs:528-533) by numpy: the median of the per-MB variances
    of the even-row/column samples; the lower quartile of the zero-motion SAD / 256 of
    the interior walk's ~48 blocks; the mean over ~24 of those blocks of (SAD0 - best of
    the +-8 step-4 grid inside the frame) / (SAD0 + 1), summed in walk order."""
    ch, cw = cur.shape
    mbw, mbh = cw // 16, ch // 16
    c = cur.astype(np.int64)
    r = ref.astype(np.int64)
    sub = c[0::2, 0::2].reshape(mbh, 8, mbw, 8).transpose(0, 2, 1, 3).reshape(mbh * mbw, 64)
    s, q = sub.sum(1), (sub * sub).sum(1)
    var = np.sort((q - s * s // 64) // 64)
    med = int(var[len(var) // 2])
    floor_v = mg = 0.0
    if mbw >= 6 and mbh >= 6:
        inner = (mbw - 4) * (mbh - 4)

        def walk(target):
            return [((2 + i % (mbw - 4)) * 16, (2 + i // (mbw - 4)) * 16) for i in range(0, inner, max(inner // target, 1))]

        def sad(bx, by, rx, ry):
            return int(np.abs(c[by:by + 16, bx:bx + 16] - r[ry:ry + 16, rx:rx + 16]).sum())

        fl = sorted(sad(bx, by, bx, by) / 256.0 for bx, by in walk(48))
        floor_v = fl[len(fl) // 4]
        acc, n = 0.0, 0
        for bx, by in walk(24):
            s0 = sad(bx, by, bx, by)
            best = min([s0] + [sad(bx, by, bx + dx, by + dy) for dy in range(-8, 9, 4) for dx in range(-8, 9, 4)
                               if 0 <= bx + dx and bx + dx + 16 <= cw and 0 <= by + dy and by + dy + 16 <= ch])
            acc += (s0 - best) / (s0 + 1)
            n += 1
        mg = acc / n
    return med < 200 and floor_v > 5.0 and mg < 0.1, {"median_var": float(med), "floor": floor_v, "mgain": mg}


def _grain_pairs():
    rng = np.random.default_rng(17)
    yy, xx = np.mgrid[0:144, 0:176]
    flat = np.full((144, 176), 120.0)
    tex = 128 + 60 * np.sin(xx / 3.0) * np.cos(yy / 4.0)
    noisy = lambda b, s: np.clip(b + rng.normal(0, s, b.shape), 0, 255).astype(np.uint8)
    return {
        "grain": (noisy(flat, 7), noisy(flat, 7)),                      # noise only: grain
        "static": (noisy(flat, 7),) * 2,                                # identical frames: no residual
        "motion": (np.roll(noisy(tex, 2), 4, 1), noisy(tex, 2)),        # texture that moves: not grain
        "texture": (noisy(tex + 40 * rng.random(tex.shape), 7), noisy(tex, 7)),  # var >= 200
        "small": (noisy(flat[:80, :80], 7), noisy(flat[:80, :80], 7)),  # below 6x6 MBs: fails open
        "random": tuple(rng.integers(0, 256, (96, 128), dtype=np.uint8) for _ in range(2)),
    }


def test_grain_signature_textbook_and_closed_forms():
    want_flags = {"grain": True, "static": False, "motion": False, "texture": False, "small": False,
                  "random": False}
    for name, (cur, ref) in _grain_pairs().items():
        g, st = O.grain_signature(cur, ref)
        tg, tst = _grain_textbook(cur, ref)
        assert (g, st) == (tg, tst), name
        assert g == want_flags[name], (name, st)
    g, st = O.grain_signature(*_grain_pairs()["static"])
    assert st["floor"] == 0.0 and st["mgain"] == 0.0


def test_grain_latch_zeroes_the_gop():
    rng = np.random.default_rng(8)
    intra, inter, mv = _two_frame_gop(rng)
    off, st = O.mbtree_qp_offsets(intra, inter, mv, res_min=0.0, grain=True)
    assert not off.any() and st == {"residual_frac": 0.0, "spread": 0.0, "eff_strength": 0.0}
    assert O.mbtree_qp_offsets(intra, inter, mv, res_min=0.0)[0].any()
    # a single-frame GOP has no pair: the latch does not apply
    one = O.mbtree_qp_offsets(intra[:1], inter[:1], mv[:1], res_min=0.0, grain=True)
    assert one[1]["eff_strength"] == 0.9


def test_library_grain_equals_oracle():
    lib = _lib.load()
    for name, (cur, ref) in _grain_pairs().items():
        assert accel.grain_signature(lib, cur, ref) == O.grain_signature(cur, ref), name
    rng = np.random.default_rng(9)
    intra, inter, mv = _two_frame_gop(rng)
    for grain in (False, True):
        a = accel.mbtree_qp_offsets(lib, intra, inter, mv, res_min=0.0, grain=grain)
        b = O.mbtree_qp_offsets(intra, inter, mv, res_min=0.0, grain=grain)
        assert np.array_equal(a[0], b[0]) and a[1] == b[1]
