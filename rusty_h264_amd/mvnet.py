"""The learned motion-vector proposal: host side of the BF16 CNN in
csrc/rh264_mvnet.hip and of rh264_learned_search (include/rh264_me.h).

The network proposes 8 integer-MV centres per 16x16 macroblock (bins of 4 pels,
-32..32 pels per axis); every proposal is then re-costed with the exact integer
search kernels and refined with the sub-pel ring walk, so the learned part only
decides where to look, never what a candidate costs.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from .accel import DevBuf, _check, _need, _ptr

NB, K, CH = 17, 289, 32
LOCAL_RANGE = 2       # integer search of +-2 pels around each proposed centre
FALLBACK_COST = 6000  # best candidate's integer cost above which a block is fully searched
REFINE_PAT = 4        # sub-pel refinement of the proposal: the +-2 quarter-pel window (subpel_pat 4)
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "mvnet.npz")

# flat parameter layout (csrc/rh264_mvnet.hip)
SIZES = [("W1", (CH, 9, 2)), ("b1", (CH,)), ("W2", (CH, 9, CH)), ("b2", (CH,)), ("W3", (CH, 9, CH)),
         ("b3", (CH,)), ("W4", (CH, 9, CH)), ("b4", (CH,)), ("Wh", (K, CH)), ("bh", (K,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def init_params(seed=0):
    """He-normal weights (fan-in = 9 * input channels, or 32 for the head), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        if name.startswith("W"):
            fan_in = int(np.prod(shape[1:]))
            parts.append(rng.normal(0.0, np.sqrt(2.0 / fan_in), shape).reshape(-1))
        else:
            parts.append(np.zeros(int(np.prod(shape))))
    return np.concatenate(parts).astype(np.float32)


def mv_bins(int_mv):
    """Training label of an integer MV (quarter pel): the nearest 4-pel bin, clamped."""
    px = np.asarray(int_mv, np.int64) // 4
    b = np.clip(np.floor_divide(px + 2, 4), -(NB // 2), NB // 2) + NB // 2
    return (b[..., 1] * NB + b[..., 0]).astype(np.int32)


def load_weights(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as d:
        return np.ascontiguousarray(d["params"], np.float32)


class MVNet:
    """One network instance on a Device, with activation buffers for up to
    max_n frame pairs of max_w x max_h."""

    def __init__(self, dev, max_n, max_w, max_h, params=None):
        self.dev = dev
        self.lib = dev.lib
        assert int(self.lib.rh264_mvnet_param_count()) == NPARAM
        h = C.c_void_p()
        _check(self.lib.rh264_mvnet_create(dev.ctx, max_n, max_w, max_h, C.byref(h)), "rh264_mvnet_create")
        self.h = h.value
        self.set_params(init_params() if params is None else params)

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        _check(self.lib.rh264_mvnet_set_params(self.dev.ctx, self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size),
               "rh264_mvnet_set_params")

    def params(self):
        p = np.empty(NPARAM, np.float32)
        _check(self.lib.rh264_mvnet_get_params(self.dev.ctx, self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size),
               "rh264_mvnet_get_params")
        return p

    def grads(self):
        g = np.empty(NPARAM, np.float32)
        _check(self.lib.rh264_mvnet_get_grads(self.dev.ctx, self.h, g.ctypes.data_as(C.POINTER(C.c_float)), g.size),
               "rh264_mvnet_get_grads")
        return g

    def forward(self, cur, ref, stride, n, w, h, logits=None, top8=None):
        _check(self.lib.rh264_mvnet_forward(self.dev.ctx, self.h, int(cur), int(ref), stride, n, w, h, _ptr(logits),
                                            _ptr(top8)), "rh264_mvnet_forward")

    def backward(self, cur, ref, stride, n, w, h, labels):
        loss = C.c_float()
        _check(self.lib.rh264_mvnet_backward(self.dev.ctx, self.h, int(cur), int(ref), stride, n, w, h, int(labels),
                                             C.byref(loss)), "rh264_mvnet_backward")
        return loss.value

    def adam(self, lr):
        _check(self.lib.rh264_mvnet_adam(self.dev.ctx, self.h, float(lr)), "rh264_mvnet_adam")

    def close(self):
        if self.h:
            self.lib.rh264_mvnet_destroy(self.h)
            self.h = None


def learned_search(dev, net, params, nframes, cur, ref, frame_stride, planes, planes_stride, local_range=LOCAL_RANGE,
                   fallback_cost=FALLBACK_COST, out=None):
    """rh264_learned_search; returns (mv, cost, fallback) device buffers (asynchronous)."""
    nb = (params.cw // 16) * (params.ch // 16) * nframes
    if out is None:
        out = (dev.alloc(8 * nb), dev.alloc(4 * nb), dev.alloc(4 * nb))
    mv, cost, fb = out
    _need(mv, 8 * nb, "mv")
    _check(dev.lib.rh264_learned_search(dev.ctx, C.byref(params), net.h, nframes, int(cur), int(ref), frame_stride,
                                        int(planes), planes_stride, local_range, int(fallback_cost), mv.ptr,
                                        cost.ptr, fb.ptr), "rh264_learned_search")
    return out
