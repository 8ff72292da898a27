"""ctypes binding of the CPU oracle (oracle/me_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the parity checker / the timed CPU baseline.
The product (rusty_h264_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LR_N = 128

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)


class SearchCfg(C.Structure):
    _fields_ = [("cw", C.c_int), ("ch", C.c_int), ("pad", C.c_int), ("bw", C.c_int), ("bh", C.c_int),
                ("range", C.c_int), ("subpel_pat", C.c_int), ("lr_fp", i64p), ("lr", i64p),
                ("hadamard", C.c_int)]


class DiamondCfg(C.Structure):
    _fields_ = [("steps", i32p), ("nsteps", C.c_int), ("sadfp", C.c_int), ("snap", C.c_int), ("fc", C.c_int),
                ("sad_only", C.c_int), ("mv_tab", C.POINTER(C.c_uint16)), ("lam", C.c_double),
                ("lam_fp", C.c_double), ("rescue", C.c_int), ("rescue_range", C.c_int),
                ("rescue_thresh", C.c_int), ("rescue_fast", C.c_int), ("rescue_var", C.c_int),
                ("rescue_learn", C.c_int), ("rescue_payoff", C.c_int)]


u16p = C.POINTER(C.c_uint16)


def build(fast=False):
    """Compile the oracle with its Makefile (gcc; seconds)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def load(fast=False, path=None):
    """The oracle library (path: another build of the same source, e.g. a test's)."""
    name = path or ("liboracle_v3.so" if fast else "liboracle.so")
    if name in _libs:
        return _libs[name]
    if path is None:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
    lib = C.CDLL(path)
    _libs[name] = lib
    return lib


def _p(a, t=u8p):
    return a.ctypes.data_as(t)


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


class Oracle:
    """Thin numpy-facing wrapper over the C restatement."""

    def __init__(self, fast=False, path=None):
        self.lib = load(fast, path)
        L = self.lib
        for n in ("orc_sad", "orc_satd4x4_wels", "orc_satd_wels", "orc_intra_cost"):
            getattr(L, n).restype = C.c_int32
        for n in ("orc_satd_px", "orc_full_search_block", "orc_satd_cost", "orc_subpel_refine_block",
                  "orc_motion_search_ref", "orc_motion_search_ex", "orc_mbtree_mc_satd", "orc_satd8_px",
                  "orc_la_block_multi", "orc_la_block_learned"):
            getattr(L, n).restype = C.c_int64
        L.orc_satd_avg.restype = C.c_uint32
        L.orc_mvbits.restype = C.c_uint32
        L.orc_inter_cost.restype = C.c_int32
        L.orc_inter_cost_ex.restype = C.c_int32
        L.orc_rc_complexity.restype = C.c_double
        L.orc_split_gate.restype = C.c_int64

    # ---- block metrics
    def sad(self, a, sa, b, sb, w, h):
        return self.lib.orc_sad(_p(_u8(a)), sa, _p(_u8(b)), sb, w, h)

    def satd_wels(self, a, sa, b, sb, w, h):
        return self.lib.orc_satd_wels(_p(_u8(a)), sa, _p(_u8(b)), sb, w, h)

    def satd_px(self, a, sa, b, sb, w, h):
        return self.lib.orc_satd_px(_p(_u8(a)), sa, _p(_u8(b)), sb, w, h)

    def satd8_px(self, src, ss, pred, ps, w, h):
        return int(self.lib.orc_satd8_px(_p(_u8(src)), ss, _p(_u8(pred)), ps, w, h))

    def satd_avg(self, src, ss, a, b, stride, w, h):
        return self.lib.orc_satd_avg(_p(_u8(src)), ss, _p(_u8(a)), _p(_u8(b)), stride, w, h)

    def hadamard_4x4(self, blk):
        i = np.ascontiguousarray(blk, np.int32).reshape(16)
        o = np.empty(16, np.int32)
        self.lib.orc_hadamard_4x4(_p(i, i32p), _p(o, i32p))
        return o

    # ---- interpolation
    def luma_sample(self, ref, cw, ch, ix, iy, fx, fy):
        return self.lib.orc_luma_sample(_p(_u8(ref)), cw, ch, ix, iy, fx, fy)

    def mc_luma(self, ref, cw, ch, x0, y0, bw, bh, mvx, mvy):
        out = np.empty((bh, bw), np.uint8)
        self.lib.orc_mc_luma(_p(_u8(ref)), int(cw), int(ch), int(x0), int(y0), int(bw), int(bh), int(mvx), int(mvy),
                             _p(out))
        return out

    def planes(self, y, cw, ch, pad=32, fused=False):
        pw, ph = cw + 2 * pad, ch + 2 * pad
        out = np.empty(4 * pw * ph, np.uint8)
        self.lib.orc_build_hpel_planes(_p(_u8(y)), cw, ch, pad, int(fused), _p(out))
        return out.reshape(4, ph, pw)

    def hpel_block(self, planes, cw, ch, pad, x0, y0, bw, bh, mvx, mvy):
        out = np.zeros((bh, bw), np.uint8)
        ok = self.lib.orc_hpel_block(_p(_u8(planes)), int(cw), int(ch), int(pad), int(x0), int(y0), int(bw), int(bh),
                                     int(mvx), int(mvy), _p(out))
        return out if ok else None

    def pred_block(self, planes, cw, ch, pad, x0, y0, bw, bh, mvx, mvy):
        out = np.empty((bh, bw), np.uint8)
        self.lib.orc_pred_block(_p(_u8(planes)), int(cw), int(ch), int(pad), int(x0), int(y0), int(bw), int(bh),
                                int(mvx), int(mvy), _p(out))
        return out

    # ---- search
    def mvbits(self, d):
        return self.lib.orc_mvbits(int(d))

    def lambda_table(self, lam):
        lr = np.empty(LR_N, np.int64)
        self.lib.orc_lambda_table(C.c_double(lam), LR_N, _p(lr, i64p))
        return lr

    def cfg(self, cw, ch, pad=32, bw=16, bh=16, search_range=32, subpel_pat=0, lam=0.0, lam_fp=None, hadamard=4):
        lr = self.lambda_table(lam)
        lr_fp = self.lambda_table(lam * 0.5 if lam_fp is None else lam_fp)
        c = SearchCfg(cw, ch, pad, bw, bh, search_range, subpel_pat, _p(lr_fp, i64p), _p(lr, i64p), hadamard)
        c._keep = (lr, lr_fp)
        c._lam, c._lam_fp = float(lam), float(lam * 0.5 if lam_fp is None else lam_fp)
        return c

    def diamond(self, cfg, steps=(16, 8, 4), sadfp=False, snap=True, fc=False, sad_only=False, mv_tab=None,
                rescue=False, rescue_range=24, rescue_thresh=3, rescue_fast=True, rescue_var=800, rescue_learn=40,
                rescue_payoff=15):
        """motion_search's knobs (orc_diamond_cfg); lambdas from cfg (lam, lam_fp); the
        learning window (rescue_learn, rescue_payoff) applies to the frame decision."""
        st = np.ascontiguousarray(steps, np.int32)
        tab = None if mv_tab is None else np.ascontiguousarray(mv_tab, np.uint16)
        dc = DiamondCfg(_p(st, i32p), len(st), int(sadfp), int(snap), int(fc), int(sad_only),
                        None if tab is None else _p(tab, u16p), cfg._lam, cfg._lam_fp, int(rescue),
                        int(rescue_range), int(rescue_thresh), int(rescue_fast), int(rescue_var), int(rescue_learn),
                        int(rescue_payoff))
        dc._keep = (st, tab)
        return dc

    def full_search_block(self, cfg, cur, planes, lx, ly, cx=0, cy=0, mvpx=0, mvpy=0):
        mv = np.empty(2, np.int32)
        c = self.lib.orc_full_search_block(C.byref(cfg), _p(_u8(cur)), _p(_u8(planes)), int(lx), int(ly), int(cx),
                                           int(cy), int(mvpx), int(mvpy), _p(mv, i32p))
        return (int(mv[0]), int(mv[1])), int(c)

    def satd_cost(self, cfg, cur, planes, lx, ly, mvx, mvy, mvpx=0, mvpy=0):
        return int(self.lib.orc_satd_cost(C.byref(cfg), _p(_u8(cur)), _p(_u8(planes)), int(lx), int(ly), int(mvx),
                                          int(mvy), int(mvpx), int(mvpy)))

    def subpel_refine_block(self, cfg, cur, planes, lx, ly, mvx, mvy, mvpx=0, mvpy=0):
        mv = np.empty(2, np.int32)
        ev = C.c_int32()
        c = self.lib.orc_subpel_refine_block(C.byref(cfg), _p(_u8(cur)), _p(_u8(planes)), int(lx), int(ly),
                                             int(mvx), int(mvy), int(mvpx), int(mvpy), _p(mv, i32p), C.byref(ev))
        return (int(mv[0]), int(mv[1])), int(c), ev.value

    def plain_search_frame(self, cfg, cur, planes, centres=None, mvps=None, nthreads=1):
        nb = (cfg.cw // cfg.bw) * (cfg.ch // cfg.bh)
        imv = np.empty((nb, 2), np.int32)
        ic = np.empty(nb, np.int64)
        mv = np.empty((nb, 2), np.int32)
        c = np.empty(nb, np.int64)
        ce = _p(np.ascontiguousarray(centres, np.int32), i32p) if centres is not None else None
        mp = _p(np.ascontiguousarray(mvps, np.int32), i32p) if mvps is not None else None
        self.lib.orc_plain_search_frame(C.byref(cfg), _p(_u8(cur)), _p(_u8(planes)), ce, mp, _p(imv, i32p),
                                        _p(ic, i64p), _p(mv, i32p), _p(c, i64p), nthreads)
        return {"int_mv": imv, "int_cost": ic, "mv": mv, "cost": c}

    def motion_search_ref(self, cfg, cur, planes, lx, ly, preds, steps=(16, 8, 4), sadfp=False, snap=True, **kw):
        mv, c, _, _ = self.motion_search_trace(cfg, cur, planes, lx, ly, preds, steps=steps, sadfp=sadfp, snap=snap,
                                               cap=0, **kw)
        return mv, c

    def motion_search_trace(self, cfg, cur, planes, lx, ly, preds, cap=4096, dc=None, **kw):
        """(mv, cost, events (ran, paid), trace [(phase, mvx, mvy, cost)]) of one search."""
        dc = self.diamond(cfg, **kw) if dc is None else dc
        pr = np.ascontiguousarray(preds, np.int32).reshape(-1)
        mv = np.empty(2, np.int32)
        ev = np.zeros(2, np.int32)
        tr = np.zeros((max(cap, 1), 4), np.int64)
        n = C.c_int()
        c = self.lib.orc_motion_search_ex(C.byref(cfg), C.byref(dc), _p(_u8(cur)), _p(_u8(planes)), int(lx), int(ly),
                                          _p(pr, i32p), len(pr) // 2, _p(mv, i32p), _p(ev, i32p), _p(tr, i64p), cap,
                                          C.byref(n))
        trace = [tuple(int(v) for v in row) for row in tr[:min(n.value, cap)]]
        return (int(mv[0]), int(mv[1])), int(c), (int(ev[0]), int(ev[1])), trace

    def predict_mv(self, a, b, c, cur_ref):
        """predict_mv (inter.rs:36-73); a, b, c = (available, (mvx, mvy), ref_idx) or None."""
        nb = np.zeros(12, np.int32)
        for k, n in enumerate((a, b, c)):
            nb[4 * k:4 * k + 4] = (0, 0, 0, -1) if n is None else (int(n[0]), n[1][0], n[1][1], n[2])
        out = np.empty(2, np.int32)
        self.lib.orc_predict_mv(_p(nb, i32p), int(cur_ref), _p(out, i32p))
        return int(out[0]), int(out[1])

    def motion_search_frame(self, cfg, cur, planes, lambda_ref=0.0, steps=(16, 8, 4), sadfp=False, snap=True, **kw):
        """A P frame's 16x16 best_part decisions in raster order with predict_mv
        predictors; planes: one plane set (reference 0) or a list, most recent first.
        Returns (ref (nblk,) int32, mv (nblk, 2) int32, cost (nblk,) int64)."""
        dc = self.diamond(cfg, steps=steps, sadfp=sadfp, snap=snap, **kw)
        pls = [planes] if isinstance(planes, np.ndarray) else list(planes)
        pls = [_u8(x) for x in pls]
        arr = (C.c_void_p * len(pls))(*[x.ctypes.data for x in pls])
        nb = (cfg.cw // 16) * (cfg.ch // 16)
        ref = np.empty(nb, np.int32)
        mv = np.empty((nb, 2), np.int32)
        c = np.empty(nb, np.int64)
        self.lib.orc_motion_search_frame(C.byref(cfg), C.byref(dc), _p(_u8(cur)), arr, len(pls),
                                         C.c_double(lambda_ref), _p(ref, i32p), _p(mv, i32p), _p(c, i64p))
        return ref, mv, c

    def partition_decision_frame(self, cfg, cur, planes, lme, split_gate, p8x8=True, **kw):
        """The P macroblock decision with partitions (SATD currency) in raster order;
        planes: a list of plane sets, most recent first; kw: diamond knobs.  Returns
        rec (nblk, 14) int32: mode, cost, then (ref, mvx, mvy) per quadrant."""
        dc = self.diamond(cfg, **kw)
        pls = [_u8(x) for x in planes]
        arr = (C.c_void_p * len(pls))(*[x.ctypes.data for x in pls])
        nb = (cfg.cw // 16) * (cfg.ch // 16)
        rec = np.zeros((nb, 14), np.int32)
        self.lib.orc_partition_decision_frame(C.byref(cfg), C.byref(dc), _p(_u8(cur)), arr, len(pls),
                                              C.c_double(lme), C.c_int64(split_gate), int(p8x8), _p(rec, i32p))
        return rec

    def split_gate(self, qp):
        return int(self.lib.orc_split_gate(int(qp)))

    def skip_mv(self, a, b, c):
        """skip_mv (mb16.rs:1841-1853); neighbours as predict_mv's."""
        nb = np.zeros(12, np.int32)
        for k, n in enumerate((a, b, c)):
            nb[4 * k:4 * k + 4] = (0, 0, 0, -1) if n is None else (int(n[0]), n[1][0], n[1][1], n[2])
        out = np.empty(2, np.int32)
        self.lib.orc_skip_mv(_p(nb, i32p), _p(out, i32p))
        return int(out[0]), int(out[1])

    def decision_predict(self, cw, ch, refs, ref, mv):
        """(skip_mv (nblk, 2), pred (ch, cw), skip_pred (ch, cw)) of one decided frame;
        refs: unpadded reference frames, most recent first."""
        rs = [_u8(x) for x in refs]
        arr = (C.c_void_p * len(rs))(*[x.ctypes.data for x in rs])
        nb = (cw // 16) * (ch // 16)
        smv = np.empty((nb, 2), np.int32)
        pred = np.empty((ch, cw), np.uint8)
        sp = np.empty((ch, cw), np.uint8)
        self.lib.orc_decision_predict(cw, ch, len(rs), arr, _p(np.ascontiguousarray(ref, np.int32), i32p),
                                      _p(np.ascontiguousarray(mv, np.int32), i32p), _p(smv, i32p), _p(pred), _p(sp))
        return smv, pred, sp

    # ---- mb-tree
    def grain_signature(self, cur, ref):
        """grain_signature (signals.rs:528-533): (grain bool, {median_var, floor, mgain})."""
        cur, ref = _u8(cur), _u8(ref)
        ch, cw = cur.shape
        st = np.zeros(3, np.float64)
        g = self.lib.orc_grain_signature(_p(cur), _p(ref), cw, ch, st.ctypes.data_as(C.POINTER(C.c_double)))
        return bool(g), {"median_var": float(st[0]), "floor": float(st[1]), "mgain": float(st[2])}

    def mbtree_qp_offsets(self, intra, inter, mv, strength=0.9, res_min=0.03, sd_min=0.0, grain=False):
        """gop_qp_offsets' policy on frame costs [n][mb_h][mb_w]: (offsets, stats)."""
        intra = np.ascontiguousarray(intra, np.int32)
        inter = np.ascontiguousarray(inter, np.int32)
        mv = np.ascontiguousarray(mv, np.int32)
        n, mb_h, mb_w = intra.shape
        off = np.empty_like(intra)
        st = np.zeros(3, np.float64)
        self.lib.orc_mbtree_qp_offsets(n, mb_w, mb_h, _p(intra, i32p), _p(inter, i32p), _p(mv, i32p),
                                       C.c_double(strength), C.c_double(res_min), C.c_double(sd_min), int(grain),
                                       _p(off, i32p), st.ctypes.data_as(C.POINTER(C.c_double)))
        return off, {"residual_frac": float(st[0]), "spread": float(st[1]), "eff_strength": float(st[2])}

    def downsample2x(self, y, cw, ch):
        out = np.empty((ch // 2, cw // 2), np.uint8)
        self.lib.orc_downsample2x(_p(_u8(y)), cw, ch, _p(out))
        return out

    def intra_cost(self, sy, cw, bx0, by0, bs):
        return self.lib.orc_intra_cost(_p(_u8(sy)), cw, bx0, by0, bs)

    def inter_cost(self, sy, cw, ch, ref, bx0, by0, bs, seed=(0, 0), max_step=8):
        mv = np.empty(2, np.int32)
        c = self.lib.orc_inter_cost(_p(_u8(sy)), cw, ch, _p(_u8(ref)), bx0, by0, bs, seed[0], seed[1], max_step,
                                    _p(mv, i32p))
        return int(c), (int(mv[0]), int(mv[1]))

    def inter_cost_trace(self, sy, cw, ch, ref, bx0, by0, bs, seed=(0, 0), max_step=8, cap=512):
        """(cost, mv, trace [(mvx, mvy, satd)]) of inter_cost."""
        mv = np.empty(2, np.int32)
        tr = np.zeros((cap, 3), np.int64)
        n = C.c_int()
        c = self.lib.orc_inter_cost_ex(_p(_u8(sy)), cw, ch, _p(_u8(ref)), bx0, by0, bs, seed[0], seed[1], max_step,
                                       _p(mv, i32p), _p(tr, i64p), cap, C.byref(n))
        return int(c), (int(mv[0]), int(mv[1])), [tuple(int(v) for v in r) for r in tr[:min(n.value, cap)]]

    def rc_complexity(self, y, ref, cw, ch):
        """lookahead::complexity of one coded-size frame (ref None = IDR)."""
        rp = _p(_u8(ref)) if ref is not None else None
        return float(self.lib.orc_rc_complexity(_p(_u8(y)), rp, cw, ch))

    def frame_costs_half(self, half, ref_half, mb_w, mb_h):
        n = mb_w * mb_h
        intra = np.empty(n, np.int32)
        inter = np.empty(n, np.int32)
        mv = np.zeros((n, 2), np.int32)
        rp = _p(_u8(ref_half)) if ref_half is not None else None
        self.lib.orc_frame_costs_half(_p(_u8(half)), rp, mb_w, mb_h, _p(intra, i32p), _p(inter, i32p),
                                      _p(mv, i32p))
        return intra, inter, mv

    def mbtree_costs_multi(self, halves, mb_w, mb_h, nrefs=3, search_range=64, nthreads=1):
        """Build-defined multi-reference costs of a stack of half-res frames."""
        n = halves.shape[0]
        nm = mb_w * mb_h
        intra = np.empty((n, nm), np.int32)
        inter = np.empty((n, nm), np.int32)
        mv = np.empty((n, nm, 2), np.int32)
        ref = np.empty((n, nm), np.int32)
        self.lib.orc_mbtree_costs_multi(_p(_u8(halves)), n, mb_w, mb_h, nrefs, search_range, _p(intra, i32p),
                                        _p(inter, i32p), _p(mv, i32p), _p(ref, i32p), nthreads)
        return intra, inter, mv, ref

    def la_block_multi(self, halves, cw, ch, f, nrefs, search_range, bx0, by0):
        mv = np.empty(2, np.int32)
        k = C.c_int32()
        c = self.lib.orc_la_block_multi(_p(_u8(halves)), int(cw), int(ch), int(f), int(nrefs), int(search_range),
                                        int(bx0), int(by0), _p(mv, i32p), C.byref(k))
        return int(c), (int(mv[0]), int(mv[1])), k.value

    def la_block_learned(self, halves, cw, ch, f, nrefs, search_range, local_range, fb_sad, centres, bx0, by0):
        """centres: [nrefs][8][2] half-res full-pel proposals in rank order."""
        mv = np.empty(2, np.int32)
        k, fb = C.c_int32(), C.c_int32()
        cen = np.ascontiguousarray(centres, np.int32).reshape(-1)
        c = self.lib.orc_la_block_learned(_p(_u8(halves)), int(cw), int(ch), int(f), int(nrefs), int(search_range),
                                          int(local_range), int(fb_sad), _p(cen, i32p), int(bx0), int(by0),
                                          _p(mv, i32p), C.byref(k), C.byref(fb))
        return int(c), (int(mv[0]), int(mv[1])), k.value, fb.value

    def propagate_multi_i32(self, intra, inter, mv, ref, mb_w, mb_h):
        n = intra.shape[0]
        out = np.empty((n, mb_w * mb_h), np.int32)
        rp = _p(np.ascontiguousarray(ref, np.int32), i32p) if ref is not None else None
        self.lib.orc_propagate_multi_i32(n, mb_w, mb_h, _p(np.ascontiguousarray(intra, np.int32), i32p),
                                         _p(np.ascontiguousarray(inter, np.int32), i32p),
                                         _p(np.ascontiguousarray(mv, np.int32), i32p), rp, _p(out, i32p))
        return out

    def propagate_f64(self, intra, inter, mv, mb_w, mb_h):
        n = intra.shape[0]
        out = np.empty((n, mb_w * mb_h), np.float64)
        self.lib.orc_propagate_f64(n, mb_w, mb_h, _p(np.ascontiguousarray(intra, np.int32), i32p),
                                   _p(np.ascontiguousarray(inter, np.int32), i32p),
                                   _p(np.ascontiguousarray(mv, np.int32), i32p),
                                   out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def propagate_i32(self, intra, inter, mv, mb_w, mb_h):
        n = intra.shape[0]
        out = np.empty((n, mb_w * mb_h), np.int32)
        self.lib.orc_propagate_i32(n, mb_w, mb_h, _p(np.ascontiguousarray(intra, np.int32), i32p),
                                   _p(np.ascontiguousarray(inter, np.int32), i32p),
                                   _p(np.ascontiguousarray(mv, np.int32), i32p), _p(out, i32p))
        return out
