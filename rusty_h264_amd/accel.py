"""Host-side mirror of the rusty_h264-accel motion-search interface over the
HIP provider (include/rh264_me.h).

The reference's accel crate exposes per-block kernels (crates/rusty_h264-accel/
src/lib.rs:30-58) that the encoder calls one block at a time.  Here the same
operations are served in batches on the GPU: `block_metric` covers the
sad/satd/satd_avg family, `build_planes` the plane cache (`build_hpel_planes`,
crates/rusty_h264-common/src/inter.rs:816-843), and `MotionSearch` the search
itself (full search, sub-pel ring refinement).  Return values use the same
domains as the reference: SAD, Wels SATD (per-4x4 rounded) or the encoder's
unhalved sum |H d|.

Declined calls (return code 1, an unsupported shape or geometry) raise
`Declined`: the caller then runs its own path, as the reference does when an
accel kernel returns None.  Errors raise RuntimeError.  Nothing here computes on
the CPU.
"""
import ctypes as C
import math

import numpy as np

from . import _lib

LR_N = _lib.LR_N
METRIC_SAD, METRIC_SATD_WELS, METRIC_SATD, METRIC_SATD_AVG, METRIC_SATD8 = 0, 1, 2, 3, 4
SHAPES = ((16, 16), (16, 8), (8, 16), (8, 8), (8, 4), (4, 8), (4, 4))
HPEL_PAD_DEFAULT = 32  # inter.rs:761


class Declined(Exception):
    """The provider does not cover this shape/geometry (accel `None`)."""


def _check(rc, what):
    if rc == 0:
        return
    if rc == 1:
        raise Declined(what)
    raise RuntimeError(f"{what}: {_lib.load().rh264_last_error().decode()}")


def lambda_me(qp, scale=1.0):
    """Motion-search lambda: sqrt(0.85 * 2^((qp-12)/3)) * scale (mb16.rs:8642, 8838)."""
    return math.sqrt(0.85 * 2.0 ** ((qp - 12) / 3.0)) * scale


def lambda_table(lam):
    """lr[b] = (int64)(lam * b): `(lam * rate as f64) as i64`, mb16.rs:2563."""
    out = (C.c_int32 * LR_N)()
    _lib.load().rh264_lambda_table(float(lam), out)
    return np.frombuffer(out, dtype=np.int32).copy()


class DevBuf:
    """A device allocation owned by a Device."""

    def __init__(self, dev, nbytes):
        self.dev = dev
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(dev.lib.rh264_alloc(dev.ctx, self.nbytes, C.byref(p)), "rh264_alloc")
        self.ptr = p.value

    def free(self):
        if self.ptr:
            _check(self.dev.lib.rh264_free(self.dev.ctx, self.ptr), "rh264_free")
            self.ptr = None

    def __int__(self):
        return self.ptr


class Device:
    """One GPU context (rh264_ctx): a stream, timing events and scratch."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        ctx = C.c_void_p()
        _check(self.lib.rh264_init(device, C.byref(ctx)), "rh264_init")
        self.ctx = ctx.value
        self._bufs = []

    def name(self):
        buf = C.create_string_buffer(256)
        _check(self.lib.rh264_device_name(self.ctx, buf, 256), "rh264_device_name")
        return buf.value.decode()

    def alloc(self, nbytes):
        b = DevBuf(self, nbytes)
        self._bufs.append(b)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        b = self.alloc(arr.nbytes)
        _check(self.lib.rh264_h2d(self.ctx, b.ptr, arr.ctypes.data, arr.nbytes), "rh264_h2d")
        return b

    def download(self, buf, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        _check(self.lib.rh264_d2h(self.ctx, out.ctypes.data, int(buf), out.nbytes), "rh264_d2h")
        return out

    def sync(self):
        _check(self.lib.rh264_sync(self.ctx), "rh264_sync")

    def tune(self, knob, value):
        """Select a kernel variant (rh264_tune); results are identical for every variant."""
        _check(self.lib.rh264_tune(self.ctx, knob.encode(), int(value)), "rh264_tune")

    def record(self, slot):
        _check(self.lib.rh264_event_record(self.ctx, slot), "rh264_event_record")

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        _check(self.lib.rh264_event_elapsed_ms(self.ctx, a, b, C.byref(ms)), "rh264_event_elapsed_ms")
        return ms.value

    def close(self):
        for b in self._bufs:
            b.free()
        self._bufs = []
        if self.ctx:
            self.lib.rh264_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _ptr(x):
    if x is None:
        return None
    return int(x)


def _need(buf, nbytes, what):
    """Refuse a device buffer smaller than the call will read or write."""
    if isinstance(buf, DevBuf) and buf.nbytes < nbytes:
        raise ValueError(f"{what}: buffer of {buf.nbytes} bytes, call needs {nbytes}")


def block_metric(dev, kind, w, h, src, ss, src_off, ref, rs, ref_off, ref_off2=None):
    """Batched accel block metrics on host arrays.

    src/ref are flat uint8 host arrays; src_off/ref_off(/ref_off2) int64 offsets
    of each block's top-left sample.  Returns int32 per block:
    SAD (satd_sad.rs:74-82), Wels SATD (satd_sad.rs:39-72), sum|H d| (satd_x4,
    satd_avg.rs:388-414) or the fused average form (satd_avg, satd_avg.rs:579-602).
    """
    n = len(src_off)
    dsrc, dref = dev.upload(np.asarray(src, np.uint8)), dev.upload(np.asarray(ref, np.uint8))
    dso = dev.upload(np.asarray(src_off, np.int64))
    dro = dev.upload(np.asarray(ref_off, np.int64))
    dro2 = dev.upload(np.asarray(ref_off2, np.int64)) if ref_off2 is not None else None
    out = dev.alloc(4 * max(n, 1))
    try:
        _check(dev.lib.rh264_block_metric(dev.ctx, kind, w, h, n, dsrc.ptr, ss, dso.ptr, dref.ptr, rs,
                                          dro.ptr, _ptr(dro2), out.ptr), "rh264_block_metric")
        return dev.download(out, (n,), np.int32)
    finally:
        for b in (dsrc, dref, dso, dro, dro2, out):
            if b is not None:
                b.free()


def sad_x4(dev, src, ss, base, offs, rs, w, h):
    """sad_x4 (satd_avg.rs:462-492): one source block vs four offsets into `base`."""
    return block_metric(dev, METRIC_SAD, w, h, src, ss, [0] * 4, base, rs, list(offs))


def satd_x4(dev, src, ss, base, offs, rs, w, h):
    """satd_x4 (satd_avg.rs:388-414): unhalved sum|H d| vs four offsets."""
    return block_metric(dev, METRIC_SATD, w, h, src, ss, [0] * 4, base, rs, list(offs))


def satd_avg_x4(dev, src, ss, plane, offs_a, offs_b, rs, w, h):
    """satd_avg_x4 (satd_avg.rs:536-571) with all operands inside one plane buffer."""
    return block_metric(dev, METRIC_SATD_AVG, w, h, src, ss, [0] * len(offs_a), plane, rs,
                        list(offs_a), list(offs_b))


def planes_bytes(cw, ch, pad=HPEL_PAD_DEFAULT):
    return int(_lib.load().rh264_planes_bytes(cw, ch, pad))


def build_planes(dev, frames_buf, nframes, cw, ch, pad=HPEL_PAD_DEFAULT, out=None):
    """Plane sets (F,H,V,C) for `nframes` device frames of cw x ch -> DevBuf."""
    pb = planes_bytes(cw, ch, pad)
    _need(frames_buf, nframes * cw * ch, "frames")
    if out is None:
        out = dev.alloc(pb * max(nframes, 1))
    _need(out, pb * nframes, "plane sets")
    _check(dev.lib.rh264_build_planes(dev.ctx, int(frames_buf), cw * ch, nframes, cw, ch, pad, int(out)),
           "rh264_build_planes")
    return out


def search_params(cw, ch, pad=HPEL_PAD_DEFAULT, bw=16, bh=16, search_range=32, subpel_pat=0,
                  lam=0.0, lam_fp=None, hadamard=4):
    """rh264_search_params; lam prices the SATD phase, lam_fp (default lam*0.5, the
    reference's RFF_ME_SADL default, mb16.rs:756-763) the SAD full-pel phase;
    hadamard 8 selects the 8x8 transform for the sub-pel stage (16x16 blocks)."""
    p = _lib.SearchParams()
    p.cw, p.ch, p.pad, p.bw, p.bh, p.range, p.subpel_pat = cw, ch, pad, bw, bh, search_range, subpel_pat
    p.hadamard = hadamard
    lr = lambda_table(lam)
    lr_fp = lambda_table(lam * 0.5 if lam_fp is None else lam_fp)
    for i in range(LR_N):
        p.lr[i] = int(lr[i])
        p.lr_fp[i] = int(lr_fp[i])
    return p


class MotionSearch:
    """The plain motion search over device-resident frames.

    Results are per block, frame-major: int_mv / mv (N, 2) int32 quarter-pel,
    int_cost / cost (N,) int32.
    """

    def __init__(self, dev, params):
        self.dev = dev
        self.p = params
        self.nblk = (params.cw // params.bw) * (params.ch // params.bh)
        self.pbytes = planes_bytes(params.cw, params.ch, params.pad)
        self._out = {}

    def _outputs(self, nframes):
        n = self.nblk * nframes
        if self._out.get("n", 0) < n:
            for b in self._out.get("bufs", []):
                b.free()
            bufs = [self.dev.alloc(8 * n), self.dev.alloc(4 * n), self.dev.alloc(8 * n),
                    self.dev.alloc(4 * n), self.dev.alloc(4 * n)]
            self._out = {"n": n, "bufs": bufs}
        return self._out["bufs"]

    def _check_inputs(self, cur_buf, planes_buf, nframes, centres, mvps, cur_stride=None, planes_stride=None):
        fb = self.p.cw * self.p.ch
        _need(cur_buf, (nframes - 1) * (cur_stride or fb) + fb, "current frames")
        _need(planes_buf, (nframes - 1) * (planes_stride or self.pbytes) + self.pbytes, "plane sets")
        _need(centres, 8 * self.nblk * nframes, "centres")
        _need(mvps, 8 * self.nblk * nframes, "mvps")

    def run_plain(self, cur_buf, planes_buf, nframes, centres=None, mvps=None, cur_stride=None,
                  planes_stride=None):
        """Full search + sub-pel refine (rh264_plain_search); asynchronous."""
        self._check_inputs(cur_buf, planes_buf, nframes, centres, mvps, cur_stride, planes_stride)
        imv, icost, mv, cost, _ = self._outputs(nframes)
        _check(self.dev.lib.rh264_plain_search(
            self.dev.ctx, C.byref(self.p), nframes, int(cur_buf),
            cur_stride or self.p.cw * self.p.ch, int(planes_buf), planes_stride or self.pbytes,
            _ptr(centres), _ptr(mvps), imv.ptr, icost.ptr, mv.ptr, cost.ptr), "rh264_plain_search")

    def run_full(self, cur_buf, planes_buf, nframes, centres=None, mvps=None):
        self._check_inputs(cur_buf, planes_buf, nframes, centres, mvps)
        imv, icost, _, _, _ = self._outputs(nframes)
        _check(self.dev.lib.rh264_full_search(
            self.dev.ctx, C.byref(self.p), nframes, int(cur_buf), self.p.cw * self.p.ch,
            int(planes_buf), self.pbytes, _ptr(centres), _ptr(mvps), imv.ptr, icost.ptr),
            "rh264_full_search")

    def run_refine(self, cur_buf, planes_buf, nframes, start_mv, mvps=None):
        self._check_inputs(cur_buf, planes_buf, nframes, start_mv, mvps)
        _, _, mv, cost, evals = self._outputs(nframes)
        _check(self.dev.lib.rh264_subpel_refine(
            self.dev.ctx, C.byref(self.p), nframes, int(cur_buf), self.p.cw * self.p.ch,
            int(planes_buf), self.pbytes, int(start_mv), _ptr(mvps), mv.ptr, cost.ptr, evals.ptr),
            "rh264_subpel_refine")

    def results(self, nframes, with_evals=False):
        n = self.nblk * nframes
        imv, icost, mv, cost, evals = self._outputs(nframes)
        d = self.dev
        r = {"int_mv": d.download(imv, (n, 2), np.int32), "int_cost": d.download(icost, (n,), np.int32),
             "mv": d.download(mv, (n, 2), np.int32), "cost": d.download(cost, (n,), np.int32)}
        if with_evals:
            r["evals"] = d.download(evals, (n,), np.int32)
        return r


def predict_blocks(dev, planes_buf, cw, ch, pad, bw, bh, xy, mv):
    """Quarter-pel prediction (mc_luma values, inter.rs:1193-1262) for n blocks."""
    xy = np.ascontiguousarray(xy, np.int32)
    mv = np.ascontiguousarray(mv, np.int32)
    n = len(xy)
    dxy, dmv = dev.upload(xy), dev.upload(mv)
    out = dev.alloc(max(n * bw * bh, 1))
    try:
        _check(dev.lib.rh264_predict_blocks(dev.ctx, int(planes_buf), cw, ch, pad, bw, bh, n, dxy.ptr,
                                            dmv.ptr, out.ptr), "rh264_predict_blocks")
        return dev.download(out, (n, bh, bw), np.uint8)
    finally:
        for b in (dxy, dmv, out):
            b.free()


class MbTree:
    """mb-tree lookahead costs and propagation over device-resident GOP frames
    (crates/rusty_h264-encoder/src/mbtree.rs, HalfRes mode)."""

    def __init__(self, dev, cw, ch, nframes):
        self.dev, self.cw, self.ch, self.n = dev, cw, ch, nframes
        self.mb_w, self.mb_h = cw // 16, ch // 16
        nm = self.mb_w * self.mb_h
        self.half = dev.alloc(nframes * (cw // 2) * (ch // 2))
        self.intra = dev.alloc(4 * nm * nframes)
        self.inter = dev.alloc(4 * nm * nframes)
        self.mv = dev.alloc(8 * nm * nframes)
        self.prop = dev.alloc(4 * nm * nframes)

    def run(self, frames_buf, frame_stride=None, propagate=True):
        """frames_buf must hold this object's nframes frames of cw x ch luma."""
        d = self.dev
        if isinstance(frames_buf, DevBuf) and frames_buf.nbytes < (self.n - 1) * (frame_stride or self.cw * self.ch) \
                + self.cw * self.ch:
            raise ValueError("frames_buf holds fewer frames than this MbTree was built for")
        _check(d.lib.rh264_mbtree_downsample(d.ctx, int(frames_buf), frame_stride or self.cw * self.ch, self.n,
                                             self.cw, self.ch, self.half.ptr), "rh264_mbtree_downsample")
        _check(d.lib.rh264_mbtree_costs(d.ctx, self.half.ptr, self.n, self.mb_w, self.mb_h, self.intra.ptr,
                                        self.inter.ptr, self.mv.ptr), "rh264_mbtree_costs")
        if propagate:
            _check(d.lib.rh264_mbtree_propagate_i32(d.ctx, self.n, self.mb_w, self.mb_h, self.intra.ptr,
                                                    self.inter.ptr, self.mv.ptr, self.prop.ptr),
                   "rh264_mbtree_propagate_i32")

    def run_multi(self, frames_buf, nrefs=3, search_range=64, frame_stride=None, propagate=True):
        """The build-defined multi-reference form (rh264_mbtree_costs_multi +
        rh264_mbtree_propagate_multi_i32): exhaustive +-search_range per reference,
        8x8 Hadamard costs, the cheapest of nrefs previous frames."""
        d = self.dev
        nm = self.mb_w * self.mb_h
        if not hasattr(self, "ref"):
            self.ref = d.alloc(4 * nm * self.n)
        _check(d.lib.rh264_mbtree_downsample(d.ctx, int(frames_buf), frame_stride or self.cw * self.ch, self.n,
                                             self.cw, self.ch, self.half.ptr), "rh264_mbtree_downsample")
        _check(d.lib.rh264_mbtree_costs_multi(d.ctx, self.half.ptr, self.n, self.mb_w, self.mb_h, nrefs, search_range,
                                              self.intra.ptr, self.inter.ptr, self.mv.ptr, self.ref.ptr),
               "rh264_mbtree_costs_multi")
        if propagate:
            _check(d.lib.rh264_mbtree_propagate_multi_i32(d.ctx, self.n, self.mb_w, self.mb_h, self.intra.ptr,
                                                          self.inter.ptr, self.mv.ptr, self.ref.ptr, self.prop.ptr),
                   "rh264_mbtree_propagate_multi_i32")

    def run_multi_learned(self, frames_buf, net, nrefs=3, search_range=64, local_range=2, fallback_sad=512,
                          frame_stride=None, propagate=True):
        """The learned-seeded multi-reference form (rh264_mbtree_costs_multi_learned):
        net is an mvnet.MVNet created for at least nframes - 1 pairs of cw x ch."""
        d = self.dev
        nm = self.mb_w * self.mb_h
        if not hasattr(self, "ref"):
            self.ref = d.alloc(4 * nm * self.n)
        if not hasattr(self, "fb"):
            self.fb = d.alloc(4 * nm * self.n)
        stride = frame_stride or self.cw * self.ch
        _check(d.lib.rh264_mbtree_downsample(d.ctx, int(frames_buf), stride, self.n, self.cw, self.ch, self.half.ptr),
               "rh264_mbtree_downsample")
        _check(d.lib.rh264_mbtree_costs_multi_learned(d.ctx, net.h, int(frames_buf), stride, self.half.ptr, self.n,
                                                      self.cw, self.ch, nrefs, search_range, local_range,
                                                      int(fallback_sad), self.intra.ptr, self.inter.ptr, self.mv.ptr,
                                                      self.ref.ptr, self.fb.ptr), "rh264_mbtree_costs_multi_learned")
        if propagate:
            _check(d.lib.rh264_mbtree_propagate_multi_i32(d.ctx, self.n, self.mb_w, self.mb_h, self.intra.ptr,
                                                          self.inter.ptr, self.mv.ptr, self.ref.ptr, self.prop.ptr),
                   "rh264_mbtree_propagate_multi_i32")

    def results(self):
        d, n, nm = self.dev, self.n, self.mb_w * self.mb_h
        r = {"half": d.download(self.half, (n, self.ch // 2, self.cw // 2), np.uint8),
             "intra": d.download(self.intra, (n, nm), np.int32),
             "inter": d.download(self.inter, (n, nm), np.int32),
             "mv": d.download(self.mv, (n, nm, 2), np.int32),
             "prop": d.download(self.prop, (n, nm), np.int32)}
        if hasattr(self, "ref"):
            r["ref"] = d.download(self.ref, (n, nm), np.int32)
        if hasattr(self, "fb"):
            r["fb"] = d.download(self.fb, (n, nm), np.int32)
        return r


def mv_cost_smooth():
    """The smooth mvd rate table (mb16.rs:507-516), 4096 uint16 entries (host)."""
    t = np.empty(4096, np.uint16)
    _lib.load().rh264_mv_cost_smooth(t.ctypes.data_as(C.POINTER(C.c_uint16)))
    return t


class DiamondCfg:
    """rh264_diamond with its rate table in device memory (mv_tab: host uint16 table
    of 4096 entries or None for Exp-Golomb lengths; lam / lam_fp: the lambdas that
    price a table's rates).  Use as a context manager or call close()."""

    def __init__(self, dev, steps=(16, 8, 4), sadfp=False, snap=True, fc=False, sad_only=False, mv_tab=None,
                 lam=0.0, lam_fp=0.0, rescue=False, rescue_range=24, rescue_thresh=3, rescue_fast=True,
                 rescue_var=800, rescue_learn=40, rescue_payoff=15):
        dc = _lib.Diamond()
        for i, st in enumerate(steps):
            dc.steps[i] = int(st)
        dc.nsteps, dc.sadfp, dc.snap, dc.fc, dc.sad_only = len(steps), int(sadfp), int(snap), int(fc), int(sad_only)
        dc.rescue, dc.rescue_range, dc.rescue_thresh = int(rescue), int(rescue_range), int(rescue_thresh)
        dc.rescue_fast, dc.rescue_var = int(rescue_fast), int(rescue_var)
        dc.rescue_learn, dc.rescue_payoff = int(rescue_learn), int(rescue_payoff)
        dc.lam, dc.lam_fp = float(lam), float(lam_fp)
        self.tab = None
        if mv_tab is not None:
            self.tab = dev.upload(np.ascontiguousarray(mv_tab, np.uint16))
            dc.mv_cost, dc.mv_tab = 1, self.tab.ptr
        self.dc = dc

    def close(self):
        if self.tab is not None:
            self.tab.free()
            self.tab = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def motion_search_ref(dev, params, nframes, cur, planes, preds=None, npreds=0, steps=(16, 8, 4), sadfp=False,
                      snap=True, cur_stride=None, planes_stride=None, events=False, **kw):
    """The reference-mode motion search (rh264_motion_search, mb16.rs:2433-3302) of every
    params.bw x params.bh block; preds: device buffer [nframes*nblk][npreds][2] (quarter
    pel), the first predictor pricing the rate; kw: the other DiamondCfg knobs.  Returns
    (mv [n][2], cost [n]) as host arrays, and the rescue events [n] with events=True."""
    n = (params.cw // params.bw) * (params.ch // params.bh) * nframes
    mv, cost, ev = dev.alloc(8 * n), dev.alloc(4 * n), dev.alloc(4 * n)
    try:
        with DiamondCfg(dev, steps=steps, sadfp=sadfp, snap=snap, **kw) as d:
            _check(dev.lib.rh264_motion_search(dev.ctx, C.byref(params), C.byref(d.dc), nframes, int(cur),
                                               cur_stride or params.cw * params.ch, int(planes),
                                               planes_stride or planes_bytes(params.cw, params.ch, params.pad),
                                               _ptr(preds), npreds, mv.ptr, cost.ptr, ev.ptr), "rh264_motion_search")
            out = dev.download(mv, (n, 2), np.int32), dev.download(cost, (n,), np.int32)
            return out + (dev.download(ev, (n,), np.int32),) if events else out
    finally:
        mv.free()
        cost.free()
        ev.free()


def motion_search_wavefront(dev, params, nframes, cur, planes, nrefs=1, lambda_ref=0.0, steps=(16, 8, 4),
                            sadfp=False, snap=True, cur_stride=None, planes_stride=None, out=None, **kw):
    """Every macroblock of nframes P frames decided in macroblock order with its 16x16
    best_part result over nrefs references (rh264_motion_search_wavefront;
    mb16.rs:4921-4959, predictor predict_mv of the decided neighbours, inter.rs:36-73).
    planes: nframes + nrefs - 1 consecutive plane sets, frame f's reference r being
    set f + nrefs - 1 - r.  out: (ref, mv, cost) device buffers or None; returns them
    (asynchronous).  kw: the other DiamondCfg knobs."""
    d = DiamondCfg(dev, steps=steps, sadfp=sadfp, snap=snap, **kw)
    dc = d.dc
    n = (params.cw // 16) * (params.ch // 16) * nframes
    if out is None:
        out = (dev.alloc(4 * n), dev.alloc(8 * n), dev.alloc(4 * n))
    ref, mv, cost = out
    _need(ref, 4 * n, "ref")
    _need(mv, 8 * n, "mv")
    _need(cost, 4 * n, "cost")
    _check(dev.lib.rh264_motion_search_wavefront(dev.ctx, C.byref(params), C.byref(dc), nframes, nrefs,
                                                 float(lambda_ref), int(cur), cur_stride or params.cw * params.ch,
                                                 int(planes),
                                                 planes_stride or planes_bytes(params.cw, params.ch, params.pad),
                                                 ref.ptr, mv.ptr, cost.ptr), "rh264_motion_search_wavefront")
    if d.tab is not None:
        dev.sync()  # the table is freed once the launches that read it are done
    d.close()
    return out


def split_gate(qp):
    """The partition split gate of mb16.rs:8866-8868 (rh264_split_gate)."""
    return int(_lib.load().rh264_split_gate(int(qp)))


def partition_decision(dev, params, nframes, cur, planes, nrefs=1, lambda_ref=0.0, gate=0, p8x8=True,
                       cur_stride=None, planes_stride=None, out=None, **kw):
    """The P macroblock decision with partitions in macroblock order
    (rh264_partition_decision; mb16.rs:8849-9010 in the SATD currency); planes as
    motion_search_wavefront; kw: the other DiamondCfg knobs.  Returns the device
    buffer of 14 int32 per macroblock {mode, cost, (ref, mvx, mvy) x 4 quadrants}."""
    n = (params.cw // 16) * (params.ch // 16) * nframes
    if out is None:
        out = dev.alloc(4 * 14 * n)
    _need(out, 4 * 14 * n, "rec")
    d = DiamondCfg(dev, **kw)
    try:
        _check(dev.lib.rh264_partition_decision(dev.ctx, C.byref(params), C.byref(d.dc), nframes, nrefs,
                                                float(lambda_ref), int(gate), int(p8x8), int(cur),
                                                cur_stride or params.cw * params.ch, int(planes),
                                                planes_stride or planes_bytes(params.cw, params.ch, params.pad),
                                                out.ptr), "rh264_partition_decision")
        dev.sync()  # the rate table, if any, is freed below
    finally:
        d.close()
    return out


def decision_predict(dev, cw, ch, pad, nframes, nrefs, planes, ref, mv, planes_stride=None):
    """rh264_decision_predict over decided frames (ref, mv device buffers as
    motion_search_wavefront returns them): (skip_mv [n][2], pred [f][ch][cw],
    skip_pred [f][ch][cw]) as host arrays (P_Skip vector mb16.rs:1841-1853, luma MC
    inter.rs:1193-1262)."""
    n = (cw // 16) * (ch // 16) * nframes
    smv, pr, sp = dev.alloc(8 * n), dev.alloc(cw * ch * nframes), dev.alloc(cw * ch * nframes)
    try:
        _check(dev.lib.rh264_decision_predict(dev.ctx, cw, ch, pad, nframes, nrefs, int(planes),
                                              planes_stride or planes_bytes(cw, ch, pad), ref.ptr, mv.ptr, smv.ptr,
                                              pr.ptr, sp.ptr), "rh264_decision_predict")
        return (dev.download(smv, (n, 2), np.int32), dev.download(pr, (nframes, ch, cw), np.uint8),
                dev.download(sp, (nframes, ch, cw), np.uint8))
    finally:
        for x in (smv, pr, sp):
            x.free()


def grain_signature(lib, cur, ref):
    """FrameSignals::grain_signature (signals.rs:528-533) of host frames cur against
    ref, through the library's host function: (grain bool, {median_var, floor, mgain})."""
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    ch, cw = cur.shape
    g = C.c_int32()
    st = np.zeros(3, np.float64)
    _check(lib.rh264_grain_signature(cur.ctypes.data, ref.ctypes.data, cw, ch, C.byref(g), st.ctypes.data),
           "rh264_grain_signature")
    return bool(g.value), {"median_var": float(st[0]), "floor": float(st[1]), "mgain": float(st[2])}


def mbtree_qp_offsets(lib, intra, inter, mv, strength=0.9, res_min=0.03, sd_min=0.0, grain=False):
    """gop_qp_offsets' policy (mbtree.rs:369-593) on host arrays of frame costs
    ([n][mb_h][mb_w] intra/inter, [n][mb_h][mb_w][2] mv), through the library's host
    function rh264_mbtree_qp_offsets; grain: the GOP's grain signature (the latch
    zeroes every offset).  Returns (offsets int32 like intra, stats {residual_frac,
    spread, eff_strength})."""
    intra = np.ascontiguousarray(intra, np.int32)
    inter = np.ascontiguousarray(inter, np.int32)
    mv = np.ascontiguousarray(mv, np.int32)
    n, mb_h, mb_w = intra.shape
    off = np.empty_like(intra)
    st = np.zeros(3, np.float64)
    _check(lib.rh264_mbtree_qp_offsets(n, mb_w, mb_h, intra.ctypes.data, inter.ctypes.data, mv.ctypes.data,
                                       float(strength), float(res_min), float(sd_min), int(grain),
                                       off.ctypes.data, st.ctypes.data), "rh264_mbtree_qp_offsets")
    return off, {"residual_frac": float(st[0]), "spread": float(st[1]), "eff_strength": float(st[2])}


def seq(width, height, qp=26, num_ref_frames=1, level_idc=30):
    """rh264_seq: the sequence the bitstream entries describe (the reference's CAVLC
    configuration, params.rs:31-172)."""
    return _lib.Seq(width, height, qp, num_ref_frames, level_idc)


def _write(fn, what, *args, guess=0):
    n = C.c_size_t()
    if guess:  # one call when the unit fits the guess
        buf = np.empty(guess, np.uint8)
        if fn(*args, buf.ctypes.data, buf.size, C.byref(n)) == 0:
            return bytes(buf[:n.value])
    fn(*args, None, 0, C.byref(n))  # size query (fails with the needed size set)
    buf = np.empty(max(int(n.value), 1), np.uint8)
    _check(fn(*args, buf.ctypes.data, buf.size, C.byref(n)), what)
    return bytes(buf[:n.value])


def write_headers(lib, s):
    """SPS + PPS as Annex-B bytes (rh264_write_headers; Sps/Pps::write_rbsp,
    params.rs:56-101, 148-172)."""
    return _write(lib.rh264_write_headers, "rh264_write_headers", C.byref(s))


def write_idr_pcm(lib, s, y, u, v):
    """The IDR picture of I_PCM macroblocks from host planes: y (16 mb_h) x (16 mb_w),
    u and v half size (rh264_write_idr_pcm)."""
    y, u, v = (np.ascontiguousarray(a, np.uint8) for a in (y, u, v))
    return _write(lib.rh264_write_idr_pcm, "rh264_write_idr_pcm", C.byref(s), y.ctypes.data, u.ctypes.data,
                  v.ctypes.data)


def write_p_slice(lib, s, frame_num, poc_lsb, nrefs, rec):
    """A P picture from decision records ([mb] x 14 int32, rh264_partition_decision's
    layout) as Annex-B bytes (rh264_write_p_slice)."""
    rec = np.ascontiguousarray(rec, np.int32)
    return _write(lib.rh264_write_p_slice, "rh264_write_p_slice", C.byref(s), int(frame_num), int(poc_lsb),
                  int(nrefs), rec.ctypes.data)


def write_p_slice_residual(lib, s, frame_num, poc_lsb, nrefs, rec, src, pred):
    """A P picture with its luma residual (rh264_write_p_slice_residual): src and pred
    are the current picture and its inter prediction at the records' vectors, both
    (16 mb_h) x (16 mb_w) host luma planes; the residual src - pred is coded at s.qp.
    Returns (Annex-B bytes, the reconstructed luma plane)."""
    rec = np.ascontiguousarray(rec, np.int32)
    src = np.ascontiguousarray(src, np.uint8)
    pred = np.ascontiguousarray(pred, np.uint8)
    if src.shape != pred.shape or src.shape != (16 * ((s.height + 15) // 16), 16 * ((s.width + 15) // 16)):
        raise ValueError("write_p_slice_residual: src and pred must be the coded-size luma plane")
    recon = np.empty_like(src)
    data = _write(lib.rh264_write_p_slice_residual, "rh264_write_p_slice_residual", C.byref(s), int(frame_num),
                  int(poc_lsb), int(nrefs), rec.ctypes.data, src.ctypes.data, pred.ctypes.data, recon.ctypes.data,
                  guess=src.size + 65536)
    return data, recon


def encode_p_frames(dev, frames, mv, qp, pad=HPEL_PAD_DEFAULT):
    """The bitstream of a clip decided at 16x16, one reference: frames[0] (host luma,
    [n][ch][cw]) as the I_PCM IDR, each later frame f as a P picture whose macroblocks
    take the vectors mv[f - 1] ([nblk][2] quarter pel) against the previous
    reconstruction, with the luma residual at qp.  Per picture the reference's plane
    set is built from the previous reconstruction on the GPU, rh264_decision_predict
    predicts the picture and rh264_write_p_slice_residual codes it (host).  Returns
    (Annex-B bytes, bytes per picture, reconstructed luma pictures)."""
    frames = np.ascontiguousarray(frames, np.uint8)
    nf, ch, cw = frames.shape
    nb = (cw // 16) * (ch // 16)
    mv = np.ascontiguousarray(mv, np.int32).reshape(nf - 1, nb, 2)
    lib = dev.lib
    s = seq(cw, ch, qp=qp, num_ref_frames=1)
    chroma = np.full((ch // 2, cw // 2), 128, np.uint8)
    head, idr = write_headers(lib, s), write_idr_pcm(lib, s, frames[0], chroma, chroma)
    rbuf, pl = dev.alloc(cw * ch), dev.alloc(planes_bytes(cw, ch, pad))
    refb, mvb = dev.upload(np.zeros(nb, np.int32)), dev.alloc(8 * nb)
    zeros = np.zeros(nb, np.int32)
    out, sizes, recon = [head, idr], [len(idr)], [frames[0]]
    try:
        for f in range(1, nf):
            _check(lib.rh264_h2d(dev.ctx, rbuf.ptr, recon[-1].ctypes.data, cw * ch), "rh264_h2d")
            _check(lib.rh264_h2d(dev.ctx, mvb.ptr, mv[f - 1].ctypes.data, 8 * nb), "rh264_h2d")
            build_planes(dev, rbuf, 1, cw, ch, pad, out=pl)
            _, pred, _ = decision_predict(dev, cw, ch, pad, 1, 1, pl, refb, mvb)
            data, rc = write_p_slice_residual(lib, s, f, 2 * f, 1, records_16x16(zeros, mv[f - 1]), frames[f], pred[0])
            out.append(data)
            sizes.append(len(data))
            recon.append(rc)
    finally:
        for x in (rbuf, pl, refb, mvb):
            x.free()
    return b"".join(out), sizes, recon


def records_16x16(ref, mv):
    """The 16x16 decision's (ref [mb], mv [mb][2]) as partition-decision records (mode
    0, cost 0, four equal quadrants)."""
    ref = np.asarray(ref, np.int32).reshape(-1)
    mv = np.asarray(mv, np.int32).reshape(-1, 2)
    rec = np.zeros((ref.size, 14), np.int32)
    for q in range(4):
        rec[:, 2 + 3 * q] = ref
        rec[:, 3 + 3 * q] = mv[:, 0]
        rec[:, 4 + 3 * q] = mv[:, 1]
    return rec


def rc_complexity(dev, cur, ref, nframes, cw, ch, frame_stride=None, out=None):
    """lookahead::complexity of nframes device-resident frames (ref None = IDR); returns
    the nframes scores as float64 (lookahead.rs:38-101)."""
    if out is None:
        out = dev.alloc(8 * nframes)
    _check(dev.lib.rh264_rc_complexity(dev.ctx, int(cur), int(ref) if ref is not None else None,
                                       frame_stride or cw * ch, nframes, cw, ch, out.ptr), "rh264_rc_complexity")
    return dev.download(out, (nframes,), np.float64)


def ref_bits(r, nrefs):
    """ref_idx rate for the motion-search cost (mb16.rs:5967-5980)."""
    if nrefs <= 1:
        return 0
    if nrefs == 2:
        return 1
    n, length = r + 1, 1
    while n > 1:
        n >>= 1
        length += 2
    return length


class Batch:
    """Batched multi-stream search (rh264_batch_*): streams of uneven size, 1..4
    references each, several partition shapes, one launch per stage per shape.

    streams: list of dicts with device buffers `cur` (cw x ch), `planes` (nrefs
    plane sets back to back), optional `mvps`, and ints cw, ch, nrefs."""

    def __init__(self, dev, streams, shapes, pad=HPEL_PAD_DEFAULT, search_range=32, subpel_pat=0, lam=0.0,
                 lam_fp=None):
        self.dev, self.shapes = dev, [tuple(s) for s in shapes]
        lib = dev.lib
        sh = np.ascontiguousarray(np.array(self.shapes, np.int32).reshape(-1))
        self._sh = sh
        arr = (_lib.Stream * len(streams))()
        self.outs = []
        for i, st in enumerate(streams):
            cw, ch, nr = st["cw"], st["ch"], st["nrefs"]
            _need(st["cur"], cw * ch, f"stream {i} cur")
            _need(st["planes"], nr * planes_bytes(cw, ch, pad), f"stream {i} planes")
            n = int(lib.rh264_batch_out_len(cw, ch, sh.ctypes.data_as(_lib.c_i32p), len(self.shapes)))
            mvps = st.get("mvps")
            _need(mvps, 4 * int(lib.rh264_batch_mvp_len(cw, ch, sh.ctypes.data_as(_lib.c_i32p), len(self.shapes))),
                  f"stream {i} mvps")
            out = dev.alloc(4 * n)
            self.outs.append((out, cw, ch))
            arr[i].cur, arr[i].planes = int(st["cur"]), int(st["planes"])
            arr[i].mvps = _ptr(mvps)
            arr[i].cw, arr[i].ch, arr[i].nrefs, arr[i].out = cw, ch, nr, out.ptr
        lr = np.ascontiguousarray(lambda_table(lam))
        lr_fp = np.ascontiguousarray(lambda_table(lam * 0.5 if lam_fp is None else lam_fp))
        h = C.c_void_p()
        _check(lib.rh264_batch_create(dev.ctx, arr, len(streams), sh.ctypes.data_as(_lib.c_i32p), len(self.shapes),
                                      pad, search_range, subpel_pat, lr_fp.ctypes.data_as(_lib.c_i32p),
                                      lr.ctypes.data_as(_lib.c_i32p), float(lam), C.byref(h)), "rh264_batch_create")
        self.h = h.value

    def run(self):
        _check(self.dev.lib.rh264_batch_run(self.dev.ctx, self.h), "rh264_batch_run")

    def search(self):
        """The integer searches alone (for timing that phase)."""
        _check(self.dev.lib.rh264_batch_search(self.dev.ctx, self.h), "rh264_batch_search")

    def autotune(self, reps=2):
        _check(self.dev.lib.rh264_batch_autotune(self.dev.ctx, self.h, reps), "rh264_batch_autotune")
        out = []
        for i in range(len(self.shapes)):
            v, p = C.c_int(), C.c_int()
            _check(self.dev.lib.rh264_batch_choice(self.h, i, C.byref(v), C.byref(p)), "rh264_batch_choice")
            out.append((v.value, p.value))
        return out

    def capture(self):
        _check(self.dev.lib.rh264_batch_capture(self.dev.ctx, self.h), "rh264_batch_capture")

    def replay(self):
        _check(self.dev.lib.rh264_batch_replay(self.dev.ctx, self.h), "rh264_batch_replay")

    def results(self):
        """Per stream: {shape: (nblk, 4) int32 records {ref, mvx, mvy, cost}}."""
        res = []
        for out, cw, ch in self.outs:
            flat = self.dev.download(out, (out.nbytes // 4,), np.int32)
            d, off = {}, 0
            for bw, bh in self.shapes:
                nb = (cw // bw) * (ch // bh)
                d[(bw, bh)] = flat[off:off + 4 * nb].reshape(nb, 4)
                off += 4 * nb
            res.append(d)
        return res

    def close(self):
        if self.h:
            self.dev.lib.rh264_batch_destroy(self.h)
            self.h = None
        for out, _, _ in self.outs:
            out.free()
        self.outs = []
