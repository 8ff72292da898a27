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
