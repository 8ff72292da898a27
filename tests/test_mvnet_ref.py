"""The numpy reference of the learned-proposal network: its float64 gradients are
checked by central finite differences; the product's bin encoding and
initialisation agree with it.  CPU only."""
import numpy as np

from oracle import mvnet_ref as R
from rusty_h264_amd import mvnet


def _data(seed, n=1, w=32, h=32):
    rng = np.random.default_rng(seed)
    cur = rng.integers(0, 256, (n, h, w), dtype=np.uint8)
    ref = rng.integers(0, 256, (n, h, w), dtype=np.uint8)
    labels = rng.integers(0, R.K, (n, h // 16, w // 16)).astype(np.int32)
    return cur, ref, labels


def test_layout_and_encoding_agree():
    assert R.NPARAM == mvnet.NPARAM
    assert np.array_equal(R.init_params(3), mvnet.init_params(3))
    mv = np.random.default_rng(0).integers(-200, 201, (500, 2)) * 4
    assert np.array_equal(R.labels_from_mv(mv), mvnet.mv_bins(mv))
    # bin centres round-trip
    k = np.arange(R.K)
    assert np.array_equal(R.labels_from_mv(R.bin_centre(k)), k)


def test_finite_difference_gradients():
    cur, ref, labels = _data(1)
    p = R.init_params(7).astype(np.float64)
    P = R.unpack(p.copy())
    for name in ("b1", "b2", "b3", "b4"):   # non-zero biases so every ReLU path is exercised
        P[name][:] = 0.05
    p = R.pack(P)
    loss, g = R.loss_and_grads(p, cur, ref, labels, bf=False)
    rng = np.random.default_rng(2)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in R.SIZES])
    idx = []
    for i in range(len(R.SIZES)):   # a few parameters of every tensor
        idx += list(rng.integers(offs[i], offs[i + 1], 4))
    eps = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = R.loss_and_grads(pp, cur, ref, labels, bf=False)
        lm, _ = R.loss_and_grads(pm, cur, ref, labels, bf=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.1415926], np.float32)
    y = R.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == np.float32(1.0078125)
    assert y[3] == -2.5 and abs(y[4] - 3.140625) < 1e-7
