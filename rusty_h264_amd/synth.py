"""Seeded synthetic luma clips for the motion-search workloads.

Only the luma plane is produced (motion search reads luma only).  All content
is deterministic for a given seed (numpy PCG64).

  cif_clip    352x288: vertical luma gradient scrolling 1 px/frame, a 64x64
              box of uniform-random texture moving (+3,+2) px/frame, Gaussian
              noise sigma=2, clipped to 0..255.
  hd_clip     WxH (default 1920x1088): a globally panning fBm background with
              8 uniform-random textured rectangles (32..256 px) moving at
              random sub-pel velocities |v| <= 12 px/frame, noise sigma=3.
"""
import numpy as np


def _finish(frames, rng, sigma):
    out = np.empty(frames.shape, np.uint8)
    for i in range(frames.shape[0]):
        f = frames[i] + rng.normal(0.0, sigma, frames[i].shape)
        out[i] = np.clip(np.rint(f), 0, 255).astype(np.uint8)
    return out


def cif_clip(nframes=60, seed=0, w=352, h=288):
    rng = np.random.default_rng(seed)
    bs = min(64, w // 2, h // 2)  # 64x64 at CIF; smaller test frames get a smaller box
    tex = rng.integers(0, 256, (bs, bs)).astype(np.float64)
    yy = np.arange(h, dtype=np.float64)[:, None]
    frames = np.empty((nframes, h, w), np.float64)
    for t in range(nframes):
        # vertical gradient scrolling 1 px/frame
        frames[t] = np.broadcast_to(((yy + t) % h) * (235.0 / h) + 16.0, (h, w))
        bx, by = 40 + 3 * t, 30 + 2 * t
        x0, y0 = bx % (w - bs), by % (h - bs)
        frames[t, y0:y0 + bs, x0:x0 + bs] = tex
    return _finish(frames, rng, 2.0)


def _fbm(rng, h, w, beta=2.2):
    """Fractional-Brownian-motion texture via spectral synthesis, scaled 0..255."""
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    f = np.sqrt(fx * fx + fy * fy)
    f[0, 0] = 1.0
    amp = f ** (-beta / 2.0)
    amp[0, 0] = 0.0
    phase = rng.uniform(0, 2 * np.pi, amp.shape)
    spec = amp * np.exp(1j * phase)
    img = np.fft.irfft2(spec, s=(h, w))
    img -= img.min()
    img *= 255.0 / max(img.max(), 1e-9)
    return img


def _bilinear(img, x0, y0, w, h):
    """img sampled at (x0 + i, y0 + j), i<w, j<h, clamped at the edges."""
    H, W = img.shape
    xs = x0 + np.arange(w, dtype=np.float64)
    ys = y0 + np.arange(h, dtype=np.float64)
    xi = np.floor(xs).astype(np.int64)
    yi = np.floor(ys).astype(np.int64)
    fx = (xs - xi)[None, :]
    fy = (ys - yi)[:, None]
    xa, xb = np.clip(xi, 0, W - 1), np.clip(xi + 1, 0, W - 1)
    ya, yb = np.clip(yi, 0, H - 1), np.clip(yi + 1, 0, H - 1)
    a = img[np.ix_(ya, xa)]
    b = img[np.ix_(ya, xb)]
    c = img[np.ix_(yb, xa)]
    d = img[np.ix_(yb, xb)]
    return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy


def hd_clip(nframes=16, seed=0, w=1920, h=1088, nrects=8, vmax=12.0, pan_max=4.0, sigma=3.0):
    rng = np.random.default_rng(seed)
    pan = rng.uniform(-pan_max, pan_max, 2)
    margin = int(np.ceil(pan_max * nframes)) + 4
    bg = _fbm(rng, h + 2 * margin, w + 2 * margin)
    rects = []
    hi = min(257, w // 2 + 1, h // 2 + 1)  # 32..256 px; small test frames get smaller rectangles
    for _ in range(nrects):
        rw, rh = (int(v) for v in rng.integers(min(32, hi - 1), hi, 2))
        tex = rng.integers(0, 256, (rh + 2, rw + 2)).astype(np.float64)
        x, y = rng.uniform(0, w - rw), rng.uniform(0, h - rh)
        ang, speed = rng.uniform(0, 2 * np.pi), rng.uniform(0, vmax)
        rects.append((rw, rh, tex, x, y, speed * np.cos(ang), speed * np.sin(ang)))
    frames = np.empty((nframes, h, w), np.float64)
    for t in range(nframes):
        frames[t] = _bilinear(bg, margin + pan[0] * t, margin + pan[1] * t, w, h)
        for rw, rh, tex, x, y, vx, vy in rects:
            px, py = x + vx * t, y + vy * t
            ix, iy = int(np.floor(px)), int(np.floor(py))
            # texture sampled at sub-pel phase of the rectangle's position
            patch = _bilinear(tex, 1.0 - (px - ix), 1.0 - (py - iy), rw, rh)
            x0, y0 = max(ix, 0), max(iy, 0)
            x1, y1 = min(ix + rw, w), min(iy + rh, h)
            if x1 > x0 and y1 > y0:
                frames[t, y0:y1, x0:x1] = patch[y0 - iy:y1 - iy, x0 - ix:x1 - ix]
    return _finish(frames, rng, sigma)


def lcg_u8(seed, n):
    """The reference's test LCG (crates/rusty_h264-accel/src/satd_sad.rs:449-460):
    state = state*6364136223846793005 + 1442695040888963407; byte = (state >> 33) as u8."""
    out = np.empty(n, np.uint8)
    st = seed & 0xFFFFFFFFFFFFFFFF
    mul, inc, mask = 6364136223846793005, 1442695040888963407, 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        st = (st * mul + inc) & mask
        out[i] = (st >> 33) & 0xFF
    return out


def corpus_pair(seed):
    """The reference's closed-form block corpus (satd_sad.rs:390-399)."""
    i = np.arange(16)[:, None]
    j = np.arange(16)[None, :]
    a = ((i * 37 + j * 101 + seed * 3) & 0xFF).astype(np.uint8)
    b = ((i * 53 + j * 17 + seed * 29 + 7) & 0xFF).astype(np.uint8)
    return a, b
