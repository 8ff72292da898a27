"""rh264_plain_search_noalloc (the entry point that allocates nothing) == rh264_plain_search
through the C ABI: two calls in a row, the second smaller, so the static scratch's
leftover keys from the first call must not leak into the second; a 1080p batch; the
size limit is an error, not a fault."""
import ctypes as C

import numpy as np
import pytest

from rusty_h264_amd import _lib, accel, synth

pytestmark = pytest.mark.gpu
LAM26 = accel.lambda_me(26)


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    print("device:", d.name(), flush=True)
    yield d
    d.close()


def _run(dev, fn, p, nframes, cur, planes, pb):
    nb = (p.cw // p.bw) * (p.ch // p.bh) * nframes
    outs = [dev.alloc(8 * nb), dev.alloc(4 * nb), dev.alloc(8 * nb), dev.alloc(4 * nb)]
    rc = getattr(dev.lib, fn)(dev.ctx, C.byref(p), nframes, int(cur), p.cw * p.ch, int(planes), pb, None, None,
                              *[o.ptr for o in outs])
    assert rc == 0, dev.lib.rh264_last_error()
    res = [dev.download(outs[0], (nb, 2), np.int32), dev.download(outs[1], (nb,), np.int32),
           dev.download(outs[2], (nb, 2), np.int32), dev.download(outs[3], (nb,), np.int32)]
    for o in outs:
        o.free()
    return res


@pytest.mark.parametrize("case", ["cif", "1080p"])
def test_noalloc_equals_plain_search(dev, case):
    if case == "cif":
        calls = [(352, 288, 4, 16, 16, 16, 0), (176, 144, 1, 8, 8, 8, 2)]   # second call smaller
    else:
        calls = [(1920, 1088, 15, 16, 16, 32, 0), (352, 288, 2, 16, 16, 16, 4)]
    for cw, ch, nf, bw, bh, rng, pat in calls:
        print("call", cw, ch, nf, bw, bh, flush=True)
        clip = synth.hd_clip(nf + 1, seed=cw + nf, w=cw, h=ch) if cw > 352 else \
            synth.cif_clip(nf + 1, seed=cw + nf, w=cw, h=ch)
        frames = dev.upload(clip)
        pb = accel.planes_bytes(cw, ch, 32)
        planes = accel.build_planes(dev, int(frames), nf, cw, ch, 32)
        p = accel.search_params(cw, ch, 32, bw, bh, rng, pat, LAM26)
        cur = int(frames) + cw * ch
        want = _run(dev, "rh264_plain_search", p, nf, cur, planes, pb)
        got = _run(dev, "rh264_plain_search_noalloc", p, nf, cur, planes, pb)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
        frames.free()
        planes.free()


def test_noalloc_size_limit(dev):
    """More blocks than the static scratch holds is an error returned before any launch."""
    cw, ch = 1920, 1088
    p = accel.search_params(cw, ch, 32, 4, 4, 8, 0, LAM26)  # 130,560 blocks a frame
    assert (cw // 4) * (ch // 4) * 3 > (1 << 18)
    rc = dev.lib.rh264_plain_search_noalloc(dev.ctx, C.byref(p), 3, None, cw * ch, None, accel.planes_bytes(cw, ch, 32),
                                            None, None, None, None, C.c_void_p(8), C.c_void_p(8))
    assert rc < 0 and b"exceed" in dev.lib.rh264_last_error()
