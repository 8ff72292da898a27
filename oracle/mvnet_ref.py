"""numpy restatement of the learned MV-proposal network (rusty_h264_amd/csrc/
rh264_mvnet.hip): forward, backward and the bin encoding.

TEST INFRASTRUCTURE ONLY: the CPU reference the GPU kernels are checked against
(with `bf16=True` it rounds activations, data gradients and the four conv
layers' weights to BF16 at the same points as the GPU), and in float64 (`bf16=False`) the function
whose gradients are checked by finite differences.  The network has no
counterpart in the reference codec.
"""
import numpy as np

C, K, NB = 32, 289, 17
SIZES = [("W1", (C, 9, 2)), ("b1", (C,)), ("W2", (C, 9, C)), ("b2", (C,)), ("W3", (C, 9, C)), ("b3", (C,)),
         ("W4", (C, 9, C)), ("b4", (C,)), ("Wh", (K, C)), ("bh", (K,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)
TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]  # (dy, dx)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round to bfloat16 (nearest even) and back, as f2bf/bf2f in the kernels."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def init_params(seed=0):
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if name.startswith("W"):
            fan_in = int(np.prod(shape[1:]))
            d[name] = rng.normal(0.0, np.sqrt(2.0 / fan_in), shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def prep(cur, ref):
    """(n,h,w) uint8 pairs -> (n, h/4, w/4, 2): ((sum of 4x4) + 8) >> 4, (v-128)/64."""
    def ds(y):
        n, h, w = y.shape
        s = y.reshape(n, h // 4, 4, w // 4, 4).astype(np.int64).sum(axis=(2, 4))
        return (((s + 8) >> 4).astype(np.float32) - 128.0) / 64.0
    return np.stack([ds(cur), ds(ref)], axis=-1).astype(np.float32)


def _conv(x, W, b, stride):
    """x (n,H,W,ci), W (co,9,ci) -> (n,Ho,Wo,co), zero padding 1."""
    n, H, Wd, ci = x.shape
    Ho, Wo = H // stride, Wd // stride
    xp = np.zeros((n, H + 2, Wd + 2, ci), x.dtype)
    xp[:, 1:H + 1, 1:Wd + 1] = x
    out = np.zeros((n, Ho, Wo, W.shape[0]), np.float64)
    for t, (dy, dx) in enumerate(TAPS):
        patch = xp[:, 1 + dy:1 + dy + stride * Ho:stride, 1 + dx:1 + dx + stride * Wo:stride, :]
        out += patch.astype(np.float64) @ W[:, t, :].T.astype(np.float64)
    return out + b


def _conv_wgrad(x, dz, stride):
    """dW[co,t,ci] = sum_p dz[p,co] x[stride p + tap, ci]."""
    n, H, Wd, ci = x.shape
    _, Ho, Wo, co = dz.shape
    xp = np.zeros((n, H + 2, Wd + 2, ci), np.float64)
    xp[:, 1:H + 1, 1:Wd + 1] = x
    dW = np.zeros((co, 9, ci))
    dzf = dz.reshape(-1, co).astype(np.float64)
    for t, (dy, dx) in enumerate(TAPS):
        patch = xp[:, 1 + dy:1 + dy + stride * Ho:stride, 1 + dx:1 + dx + stride * Wo:stride, :].reshape(-1, ci)
        dW[:, t, :] = dzf.T @ patch
    return dW, dzf.sum(0)


def _conv_dgrad(dz, W):
    """stride-1 data gradient: dx[q,ci] = sum_t sum_co W[co,t,ci] dz[q - tap, co]."""
    n, H, Wd, co = dz.shape
    dzp = np.zeros((n, H + 2, Wd + 2, co), np.float64)
    dzp[:, 1:H + 1, 1:Wd + 1] = dz
    dx = np.zeros((n, H, Wd, W.shape[2]))
    for t, (dy, dx_) in enumerate(TAPS):
        patch = dzp[:, 1 - dy:1 - dy + H, 1 - dx_:1 - dx_ + Wd, :]
        dx += patch @ W[:, t, :].astype(np.float64)
    return dx


def forward(p, cur, ref, bf=True):
    P = unpack(np.asarray(p, np.float64))
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    x0 = r(prep(cur, ref))
    a1 = r(np.maximum(_conv(x0, r(P["W1"]), P["b1"], 2), 0))
    a2 = r(np.maximum(_conv(a1, r(P["W2"]), P["b2"], 1), 0))
    a3 = r(np.maximum(_conv(a2, r(P["W3"]), P["b3"], 1), 0))
    a4 = r(np.maximum(_conv(a3, r(P["W4"]), P["b4"], 1), 0))
    n, H, W, _ = a4.shape
    f = a4.astype(np.float64).reshape(n, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))
    logits = r(f).astype(np.float64) @ r(P["Wh"]).astype(np.float64).T + P["bh"]
    return {"x0": x0, "a1": a1, "a2": a2, "a3": a3, "a4": a4, "f": f, "logits": logits}


def loss_and_grads(p, cur, ref, labels, bf=True):
    """Mean softmax cross-entropy over cells and its gradient (flat, like p)."""
    P = unpack(np.asarray(p, np.float64))
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    F = forward(p, cur, ref, bf)
    z = F["logits"].reshape(-1, K)
    lab = np.asarray(labels).reshape(-1)
    m = z.max(1, keepdims=True)
    e = np.exp(z - m)
    se = e.sum(1, keepdims=True)
    ncell = z.shape[0]
    loss = float(np.mean(np.log(se[:, 0]) + m[:, 0] - z[np.arange(ncell), lab]))
    dl = e / se
    dl[np.arange(ncell), lab] -= 1.0
    dl /= ncell
    G = {}
    f = F["f"].reshape(-1, C)
    G["Wh"] = dl.T @ f
    G["bh"] = dl.sum(0)
    df = (dl @ P["Wh"]).reshape(F["f"].shape)
    n, Hm, Wm, _ = df.shape
    a4 = F["a4"]
    d4 = np.repeat(np.repeat(df, 2, axis=1), 2, axis=2) * 0.25
    d4 = r(np.where(a4 != 0, d4, 0.0))
    acts = [F["a1"], F["a2"], F["a3"], a4]
    dz = d4
    for layer in (4, 3, 2):
        G[f"W{layer}"], G[f"b{layer}"] = _conv_wgrad(acts[layer - 2], dz, 1)
        dx = _conv_dgrad(dz, r(P[f"W{layer}"]))
        dz = r(np.where(acts[layer - 2] != 0, dx, 0.0))
    G["W1"], G["b1"] = _conv_wgrad(F["x0"], dz, 2)
    return loss, pack(G)


def rank_keys(logits):
    """32-bit ranking key of every logit (bins along the last axis): the
    order-preserving image of the FP32 bits with its 9 lowest bits replaced by
    511 - bin.  The top-8 proposals are the bins of the 8 largest keys."""
    z = np.ascontiguousarray(logits, np.float32)
    u = z.view(np.uint32)
    o = np.where(u & np.uint32(0x80000000), ~u, u | np.uint32(0x80000000)).astype(np.uint32)
    bins = np.arange(z.shape[-1], dtype=np.uint32)
    return (o & np.uint32(0xFFFFFE00)) | (np.uint32(511) - bins)


def top8(logits):
    k = rank_keys(logits)
    return np.argsort(-k.astype(np.int64), axis=-1, kind="stable")[..., :8].astype(np.int32)


def labels_from_mv(int_mv):
    """Bin index of an integer MV (quarter pel): nearest 4-pel bin in -8..8 per axis."""
    px = np.asarray(int_mv, np.int64) // 4
    b = np.clip(np.floor_divide(px + 2, 4), -(NB // 2), NB // 2) + NB // 2
    return (b[..., 1] * NB + b[..., 0]).astype(np.int32)


def bin_centre(k):
    """Quarter-pel MV of bin k's centre."""
    k = np.asarray(k)
    return np.stack([((k % NB) - NB // 2) * 16, ((k // NB) - NB // 2) * 16], -1)
