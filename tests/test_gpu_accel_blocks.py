"""GPU parity of the per-block family (the accel crate's call shape, host pointers):
every function against the oracle's restatement of the reference function it
replaces, its declines where the reference returns None / false, and a C caller of
the library (tests/cabi_smoke.c, built by __graft_entry__.build())."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle
from rusty_h264_amd import _lib, synth

pytestmark = pytest.mark.gpu
O = Oracle()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
L = None


@pytest.fixture(scope="module", autouse=True)
def lib():
    global L
    L = _lib.load()
    yield L
    L.rh264_accel_release()


def _ptr(a):
    return a.ctypes.data


def _rand_plane(rng, h, w):
    return rng.integers(0, 256, (h, w), dtype=np.uint8)


@pytest.mark.parametrize("name,w,h,fn", [("satd_16x16", 16, 16, "wels"), ("satd_16x8", 16, 8, "wels"),
                                         ("satd_8x16", 8, 16, "wels"), ("satd_8x8", 8, 8, "wels"),
                                         ("satd_4x4", 4, 4, "wels"), ("sad_16x16", 16, 16, "sad"),
                                         ("sad_16x8", 16, 8, "sad"), ("sad_8x16", 8, 16, "sad")])
def test_pair_functions(name, w, h, fn):
    """satd_* (Wels domain) and sad_* with arbitrary strides == the oracle, including
    the corpus pairs and 0/255 extremes."""
    rng = np.random.default_rng(len(name))
    a, b = _rand_plane(rng, 40, 37), _rand_plane(rng, 33, 29)
    ext = np.where(rng.integers(0, 2, (h, w)) == 1, 255, 0).astype(np.uint8)
    cases = [(a, 37, 3 * 37 + 5, b, 29, 2 * 29 + 1), (a, 37, 0, b, 29, 0),
             (ext, w, 0, np.ascontiguousarray(255 - ext), w, 0)]
    for seed in (0, 17, 95):
        ca, cb = synth.corpus_pair(seed)
        cases.append((ca, 16, 0, cb, 16, 0))
    ref = O.satd_wels if fn == "wels" else O.sad
    out = C.c_int32()
    for pa, sa, oa, pb, sb, ob in cases:
        pa = np.ascontiguousarray(pa).reshape(-1)
        pb = np.ascontiguousarray(pb).reshape(-1)
        rc = getattr(L, "rh264_" + name)(_ptr(pa) + oa, sa, _ptr(pb) + ob, sb, C.byref(out))
        assert rc == 0, L.rh264_last_error()
        assert out.value == ref(pa[oa:], sa, pb[ob:], sb, w, h)


@pytest.mark.parametrize("w,h", [(16, 16), (16, 8), (8, 16), (8, 8)])
def test_x4_family(w, h):
    """sad_x4 / satd_x4 (one base, four offsets), satd_x4p (four operands) and
    satd_avg_x4 / satd_avg (fused average) == the oracle's SAD, sum|H d| and
    satd_avg; satd_avg(a, a) == satd_x4 (satd_avg_compare.rs:165-167)."""
    rng = np.random.default_rng(w * 100 + h)
    rs, ss = 48, 24
    src = _rand_plane(rng, 20, ss).reshape(-1)
    base = _rand_plane(rng, 40, rs).reshape(-1)
    other = [_rand_plane(rng, 40, rs).reshape(-1) for _ in range(4)]
    o = (C.c_size_t * 4)(0, 7, 3 * rs + 1, 9 * rs + 20)
    out = (C.c_uint32 * 4)()
    for kind, fn in (("sad_x4", O.sad), ("satd_x4", O.satd_px)):
        assert getattr(L, "rh264_" + kind)(_ptr(src), ss, _ptr(base), o, rs, w, h, out) == 0
        assert list(out) == [fn(src, ss, base[o[i]:], rs, w, h) for i in range(4)], kind
    ptrs = (C.c_void_p * 4)(*[_ptr(x) for x in other])
    assert L.rh264_satd_x4p(_ptr(src), ss, ptrs, o, rs, w, h, out) == 0
    assert list(out) == [O.satd_px(src, ss, other[i][o[i]:], rs, w, h) for i in range(4)]
    pa = (C.c_void_p * 4)(*[_ptr(x) for x in other])
    pb = (C.c_void_p * 4)(_ptr(base), _ptr(other[2]), _ptr(base), _ptr(other[0]))
    ob = (C.c_size_t * 4)(5, 2 * rs, rs + 3, 0)
    bases = [base, other[2], base, other[0]]
    assert L.rh264_satd_avg_x4(_ptr(src), ss, pa, o, pb, ob, rs, w, h, out) == 0
    for i in range(4):
        assert out[i] == O.satd_avg(src, ss, other[i][o[i]:], bases[i][ob[i]:], rs, w, h), i
    one = C.c_uint32()
    assert L.rh264_satd_avg(_ptr(src), ss, _ptr(other[1]) + 11, _ptr(base) + rs, rs, w, h, C.byref(one)) == 0
    assert one.value == O.satd_avg(src, ss, other[1][11:], base[rs:], rs, w, h)
    assert L.rh264_satd_avg(_ptr(src), ss, _ptr(base), _ptr(base), rs, w, h, C.byref(one)) == 0
    assert one.value == O.satd_px(src, ss, base, rs, w, h)


def test_x4_declines_uncovered_shapes():
    """x4_shape (satd_avg.rs:342-344) is 16x16/16x8/8x16/8x8: others return None."""
    src = np.zeros(256, np.uint8)
    o = (C.c_size_t * 4)(0, 0, 0, 0)
    out = (C.c_uint32 * 4)()
    one = C.c_uint32()
    for w, h in ((4, 4), (8, 4), (4, 8), (32, 16)):
        assert L.rh264_sad_x4(_ptr(src), 16, _ptr(src), o, 16, w, h, out) == 1
        assert L.rh264_satd_x4(_ptr(src), 16, _ptr(src), o, 16, w, h, out) == 1
        assert L.rh264_satd_avg(_ptr(src), 16, _ptr(src), _ptr(src), 16, w, h, C.byref(one)) == 1


@pytest.mark.parametrize("cw,ch", [(64, 48), (176, 144), (352, 288)])
def test_hpel_fused(cw, ch):
    """hpel_fused over the padded plane (pad 32) == the oracle's fused builder
    (inter.rs:908-987, itself pinned against the tile walk and the separable filter),
    including 0/255 checkerboards."""
    rng = np.random.default_rng(cw)
    for y in (rng.integers(0, 256, (ch, cw), dtype=np.uint8),
              np.where((np.add.outer(np.arange(ch), np.arange(cw)) % 2) == 0, 255, 0).astype(np.uint8)):
        pad = 32
        pw, ph = cw + 2 * pad, ch + 2 * pad
        f = np.ascontiguousarray(np.pad(y, pad, mode="edge"))
        h, v, c = (np.zeros((ph, pw), np.uint8) for _ in range(3))
        assert L.rh264_hpel_fused(_ptr(f), pw, ph, _ptr(h), _ptr(v), _ptr(c)) == 0, L.rh264_last_error()
        want = O.planes(y, cw, ch, pad, fused=True)
        assert np.array_equal(h, want[1]) and np.array_equal(v, want[2]) and np.array_equal(c, want[3])
    f = np.zeros(20 * 20, np.uint8)
    assert L.rh264_hpel_fused(_ptr(f), 20, 20, _ptr(f), _ptr(f), _ptr(f)) == 1   # pw < 24: false
    assert L.rh264_hpel_fused(_ptr(f), 24, 12, _ptr(f), _ptr(f), _ptr(f)) == 1   # outside the kernel's geometry


@pytest.mark.parametrize("w,h", [(16, 16), (16, 8), (8, 16), (8, 8)])
def test_mectx(w, h):
    """MeCtx::new + eval (mectx.rs:50-166): every quarter phase, inside and at the
    validated window's bounds, == the oracle's SATD of the clamped prediction; None
    outside the window; new declines what the reference's does."""
    cw, ch, pad = 96, 64, 32
    rng = np.random.default_rng(w + 3 * h)
    y = rng.integers(0, 256, (ch, cw), dtype=np.uint8)
    cur = rng.integers(0, 256, (ch, cw), dtype=np.uint8)
    planes = O.planes(y, cw, ch, pad)
    pw, ph = cw + 2 * pad, ch + 2 * pad
    F, H, V, Cp = (np.ascontiguousarray(planes[k]).reshape(-1) for k in range(4))
    lx, ly = 32, 16
    srcp = cur.reshape(-1)[ly * cw + lx:]
    m = C.c_void_p()
    rc = L.rh264_mectx_new(_ptr(srcp), srcp.size, cw, _ptr(F), _ptr(H), _ptr(V), _ptr(Cp), pw * ph, pw, pad, pw, ph,
                           lx, ly, w, h, C.byref(m))
    assert rc == 0, L.rh264_last_error()
    cfg = O.cfg(cw, ch, pad, w, h, 8, 0, 0.0)
    mvs = [(fx + 4 * dx, fy + 4 * dy) for fx in range(4) for fy in range(4) for dx in (-7, 0, 5) for dy in (-3, 2)]
    # the window's bounds: px = lx + (mvx >> 2) + pad in 0..pw - w - 1
    mvs += [(4 * (-lx - pad), 0), (4 * (pw - w - 1 - lx - pad) + 3, 1), (4 * (-lx - pad) - 1, 0),
            (4 * (pw - w - lx - pad), 0), (0, 4 * (ph - h - ly - pad)), (0, 4 * (-ly - pad)) ]
    arr = np.array(mvs, np.int32).reshape(-1)
    out = np.zeros(len(mvs), np.uint32)
    ok = np.zeros(len(mvs), np.uint8)
    assert L.rh264_mectx_eval_n(m, len(mvs), _ptr(arr), _ptr(out), _ptr(ok)) == 0
    one = C.c_uint32()
    for i, (mx, my) in enumerate(mvs):
        px, py = lx + (mx >> 2) + pad, ly + (my >> 2) + pad
        inside = 0 <= px <= pw - w - 1 and 0 <= py <= ph - h - 1
        assert bool(ok[i]) == inside, (mx, my)
        rc = L.rh264_mectx_eval(m, mx, my, C.byref(one))
        assert rc == (0 if inside else 1)
        if inside:
            want = O.satd_cost(cfg, cur, planes, lx, ly, mx, my)
            assert out[i] == want and one.value == want, (mx, my)
    assert L.rh264_mectx_free(m) == 0
    # declines: stride != pw, an uncovered shape, a short source slice
    for args in ((pw + 1, w, h, srcp.size), (pw, 4, 4, srcp.size), (pw, w, h, 3)):
        st, ww, hh, sl = args
        assert L.rh264_mectx_new(_ptr(srcp), sl, cw, _ptr(F), _ptr(H), _ptr(V), _ptr(Cp), pw * ph, st, pad, pw, ph,
                                 lx, ly, ww, hh, C.byref(m)) == 1


def test_c_caller():
    """A C program linked against the library (no Python in between) gets the same
    values: one per-block entry, a decline, and one batched search."""
    exe = os.path.join(ROOT, "tests", "cabi_smoke")
    if not os.path.exists(exe):
        pytest.skip("tests/cabi_smoke is built by __graft_entry__.build()")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cabi_smoke ok" in r.stdout
