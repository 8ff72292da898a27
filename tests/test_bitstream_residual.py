"""The luma residual of P pictures (CPU: the library's writer is host code).

The oracle's residual (oracle/bitstream.py) is pinned by textbook known answers for
CAVLC, by the code tables' prefix property, by a write/read round trip over every nC
context, by closed forms of the 8.5.12 inverse transform and by the reconstruction
error of the quantiser at every QP; the library's rh264_write_p_slice_residual must
write the oracle's bytes and reconstruction, and the oracle's parser and decoder must
rebuild the library's reconstruction from the stream alone."""
import numpy as np
import pytest

from oracle import bitstream as B
from oracle.oracle import Oracle
from rusty_h264_amd import _lib, accel


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def _scan(blk):
    return [int(np.asarray(blk).reshape(16)[B.ZIGZAG[i]]) for i in range(16)]


# ---- CAVLC ----
@pytest.mark.parametrize("blk,bits", [
    # I. Richardson, "H.264 and MPEG-4 Video Compression" (Wiley, 2003), 6.4.13.2,
    # examples 1 and 2 (nC = 0)
    ([[0, 3, -1, 0], [0, -1, 1, 0], [1, 0, 0, 0], [0, 0, 0, 0]], "000010001110010111101101"),
    ([[-2, 4, 0, -1], [3, 0, 0, 0], [-3, 0, 0, 0], [0, 0, 0, 0]], "000000011010001001000010111001100"),
])
def test_cavlc_textbook_examples(blk, bits):
    w = B.BitWriter()
    B.write_block(w, _scan(blk), 0)
    assert w.bitstring() == bits
    w.bits += [1] + [0] * (-(len(w.bits) + 1) % 8)
    lv, n = B.read_block(B.BitReader(w.bytes()), 0)
    assert lv == _scan(blk)


def _prefix_free(codes):
    codes = [c for c in codes if c]
    return all(not b.startswith(a) for a in codes for b in codes if a is not b)


def test_vlc_tables_are_prefix_free():
    for tab in B.COEFF_TOKEN:
        flat = [tab[t][k] for k in range(17) for t in range(4)]
        assert len([c for c in flat if c]) == 62 and _prefix_free(flat)
    for i, row in enumerate(B.TOTAL_ZEROS):
        assert len(row) == 16 - i and _prefix_free(row)
    for i, row in enumerate(B.RUN_BEFORE):
        assert len(row) == (i + 2 if i < 6 else 15) and _prefix_free(row)


def test_cavlc_round_trip_every_context():
    rng = np.random.default_rng(1)
    for t in range(1500):
        lv = [0] * 16
        k = int(rng.integers(0, 17))
        for p in rng.choice(16, k, replace=False):
            mag = int(rng.choice([1, 1, 1, 2, 3, int(rng.integers(1, 2000))]))
            lv[p] = mag if rng.random() < 0.5 else -mag
        nc = int(rng.integers(0, 17))
        w = B.BitWriter()
        assert B.write_block(w, lv, nc) == k
        w.bits += [1, 0, 1]
        w.align()
        r = B.BitReader(w.bytes())
        got, n = B.read_block(r, nc)
        assert got == lv and n == k, (lv, nc)
        assert [r.u(1) for _ in range(3)] == [1, 0, 1]


# ---- transform and quantiser ----
def test_inverse_transform_closed_forms():
    for c in (-4096, -64, -33, 0, 31, 32, 64, 1000):
        d = np.zeros((4, 4), np.int64)
        d[0, 0] = c
        assert (B.inverse4x4(d) == (c + 32) >> 6).all()
    # the first AC basis of the rows: d01 = c gives columns (c, c/2, -c/2, -c) before rounding
    d = np.zeros((4, 4), np.int64)
    d[0, 1] = 128
    assert (B.inverse4x4(d) == np.array([[2, 1, -1, -2]] * 4)).all()
    # forward core of a constant block is its DC only: 16 x the value
    w = B.forward4x4(np.full((4, 4), 7))
    assert w[0, 0] == 112 and (w.reshape(16)[1:] == 0).all()


def _recon_rms(qp, rng, vtab=None):
    errs = []
    for _ in range(60):
        x = rng.integers(-255, 256, (4, 4))
        z = B.quant4x4(B.forward4x4(x), qp)
        rr = B.inverse4x4(B.dequant4x4(z, qp))
        errs.append(np.sqrt(np.mean((rr - x) ** 2.0)))
    return float(np.mean(errs))


def test_quantiser_reconstruction_error_every_qp(monkeypatch):
    rng = np.random.default_rng(3)
    for qp in range(52):
        qstep = (0.625, 0.6875, 0.8125, 0.875, 1.0, 1.125)[qp % 6] * 2 ** (qp // 6)
        assert _recon_rms(qp, rng) <= 0.5 * qstep + 0.3, qp
    # a wrong scale in the dequantiser table is caught
    bad = tuple(tuple(2 * v if k == 0 and i == 1 else v for i, v in enumerate(row)) for k, row in enumerate(B.V))
    monkeypatch.setattr(B, "V", bad)
    assert _recon_rms(24, rng) > 0.5 * 1.0 * 16 + 0.3


# ---- the library against the oracle ----
def _records(rng, mbw, mbh, nrefs, mvr=24):
    rec = np.zeros((mbw * mbh, 14), np.int32)
    for m in range(mbw * mbh):
        mode = int(rng.integers(0, 4)) if rng.random() < 0.4 else 0
        base = [0, 0, 0] if rng.random() < 0.3 else [int(rng.integers(0, nrefs)), int(rng.integers(-mvr, mvr + 1)),
                                                     int(rng.integers(-mvr, mvr + 1))]
        quads = [base] + [[int(rng.integers(0, nrefs)), int(rng.integers(-mvr, mvr + 1)),
                           int(rng.integers(-mvr, mvr + 1))] for _ in range(3)]
        same = {0: [0, 0, 0, 0], 1: [0, 0, 2, 2], 2: [0, 1, 0, 1], 3: [0, 1, 2, 3]}[mode]
        rec[m, 0] = mode
        for q in range(4):
            rec[m, 2 + 3 * q:5 + 3 * q] = quads[same[q]]
    return rec


def _predict(o, refs, rec, w, h):
    """The inter prediction of every macroblock at its records' vectors (per quadrant)."""
    pred = np.zeros((h, w), np.uint8)
    mbw = w // 16
    for m, r in enumerate(rec):
        mx, my = m % mbw, m // mbw
        for q in range(4):
            ref, x, y = (int(v) for v in r[2 + 3 * q:5 + 3 * q])
            x0, y0 = mx * 16 + (q % 2) * 8, my * 16 + (q // 2) * 8
            pred[y0:y0 + 8, x0:x0 + 8] = o.mc_luma(refs[ref], w, h, x0, y0, 8, 8, x, y)
    return pred


@pytest.mark.parametrize("w,h,nref,qp", [(64, 48, 1, 26), (96, 64, 2, 14), (80, 48, 3, 38), (48, 32, 4, 0),
                                         (64, 32, 1, 51)])
def test_library_writes_the_oracle_residual(lib, w, h, nref, qp):
    rng = np.random.default_rng(w + h + nref + qp)
    o = Oracle()
    s = accel.seq(w, h, qp=qp, num_ref_frames=nref)
    src = rng.integers(0, 256, (h, w)).astype(np.uint8)
    # a prediction near the source (small residuals, many zero blocks) and a far one
    for spread in (3, 255):
        pred = np.clip(src.astype(int) + rng.integers(-spread, spread + 1, (h, w)), 0, 255).astype(np.uint8)
        rec = _records(rng, w // 16, h // 16, nref)
        got, recon = accel.write_p_slice_residual(lib, s, 1, 2, nref, rec, src, pred)
        want, wrecon = B.p_slice(w, h, nref, 1, 2, nref, rec, src=src, pred=pred, qp=qp)
        assert got == want and np.array_equal(recon, wrecon), spread


def test_clip_decodes_to_the_reconstruction(lib):
    """An IDR and four P pictures with residual; each P picture predicts from the
    previous reconstruction.  The oracle parses the stream and decodes the luma from
    the bitstream alone (8.4.2.2 prediction at the parsed vectors + 8.5.12 residual):
    every picture equals the library's reconstruction."""
    rng = np.random.default_rng(7)
    o = Oracle()
    w, h, nref, qp = 96, 64, 2, 28
    mbw, mbh = w // 16, h // 16
    yy, xx = np.mgrid[0:h, 0:w]
    src = [np.clip(128 + 60 * np.sin((xx + 3 * f) / 7.0) * np.cos((yy - 2 * f) / 9.0) + rng.normal(0, 4, (h, w)),
                   0, 255).astype(np.uint8) for f in range(5)]
    s = accel.seq(w, h, qp=qp, num_ref_frames=nref)
    chroma = np.full((h // 2, w // 2), 128, np.uint8)
    stream = accel.write_headers(lib, s) + accel.write_idr_pcm(lib, s, src[0], chroma, chroma)
    recons = [src[0]]
    for f in range(1, 5):
        nact = min(f, nref)
        dpb = recons[::-1][:nact]  # ref_idx 0 = the previous picture
        rec = _records(rng, mbw, mbh, nact, mvr=12)
        pred = _predict(o, dpb, rec, w, h)
        data, recon = accel.write_p_slice_residual(lib, s, f, 2 * f, nact, rec, src[f], pred)
        stream += data
        recons.append(recon)
    units = B.parse_stream(stream)
    assert [u[0] for u in units] == ["sps", "pps"] + ["slice"] * 5
    decoded = [np.block([[units[2][4][my * mbw + mx][1] for mx in range(mbw)] for my in range(mbh)])]
    coded = 0
    for f in range(1, 5):
        hdr, mbs = units[2 + f][3], units[2 + f][4]
        dpb = decoded[::-1][:hdr["num_ref_idx_active"]]
        rec = np.zeros((mbw * mbh, 14), np.int32)
        for m, (mode, quads, _) in enumerate(mbs):
            for q in range(4):
                rec[m, 2 + 3 * q:5 + 3 * q] = quads[q]
        pred = _predict(o, dpb, rec, w, h)
        out = pred.copy()
        for m, lv in enumerate(hdr["levels"]):
            if lv is not None:
                B.reconstruct(pred, lv, m % mbw, m // mbw, qp + hdr["slice_qp_delta"], out)
                coded += 1
        assert all(d == 0 for d in hdr.get("mb_qp_delta", []))
        assert hdr["disable_deblocking_filter_idc"] == 1 and "alpha" not in hdr
        decoded.append(out)
        assert np.array_equal(out, recons[f]), f
    assert coded > 0


def test_residual_writer_rejects_bad_pictures(lib):
    s = accel.seq(32, 32)
    rec = np.zeros((4, 14), np.int32)
    with pytest.raises(ValueError):
        accel.write_p_slice_residual(lib, s, 1, 2, 1, rec, np.zeros((16, 32), np.uint8), np.zeros((16, 32), np.uint8))


def _nc_no_round(tc, gx, gy):
    ha, hb = gx > 0, gy > 0
    na = int(tc[gy, gx - 1]) if ha else 0
    nb = int(tc[gy - 1, gx]) if hb else 0
    return (na + nb) >> 1 if ha and hb else na if ha else nb


@pytest.mark.parametrize("kind", ["nc_without_rounding", "transposed_scan", "quant_intra_dead_zone"])
def test_planted_residual_mistakes_are_caught(lib, monkeypatch, kind):
    """A copy of the oracle with one plausible mistake no longer writes the library's
    bytes (so the byte comparison above would catch it in either)."""
    rng = np.random.default_rng(21)
    w, h, qp = 96, 64, 20
    s = accel.seq(w, h, qp=qp)
    src = rng.integers(0, 256, (h, w)).astype(np.uint8)
    # residual amplitudes varying over the picture, so that nC takes every table
    amp = np.repeat(np.repeat(rng.integers(1, 40, (h // 8, w // 8)), 8, 0), 8, 1)
    pred = np.clip(src.astype(int) + (rng.standard_normal((h, w)) * amp).astype(int), 0, 255).astype(np.uint8)
    rec = _records(rng, w // 16, h // 16, 1)
    got, _ = accel.write_p_slice_residual(lib, s, 1, 2, 1, rec, src, pred)
    assert got == B.p_slice(w, h, 1, 1, 2, 1, rec, src=src, pred=pred, qp=qp)[0]
    if kind == "nc_without_rounding":
        monkeypatch.setattr(B, "_nc", _nc_no_round)
    elif kind == "transposed_scan":
        monkeypatch.setattr(B, "ZIGZAG", (0, 4, 1, 2, 5, 8, 12, 9, 6, 3, 7, 10, 13, 14, 11, 15))
    else:
        orig = B.quant4x4

        def intra_dz(wc, q):  # 2^qbits / 3, the intra offset
            qbits = 15 + q // 6
            z = orig(wc, q)
            for i in range(4):
                for j in range(4):
                    a = abs(int(wc[i][j]))
                    lvl = (a * B.MF[q % 6][B._cls(i, j)] + (1 << qbits) // 3) >> qbits
                    z[i][j] = -lvl if wc[i][j] < 0 else lvl
            return z
        monkeypatch.setattr(B, "quant4x4", intra_dz)
    assert got != B.p_slice(w, h, 1, 1, 2, 1, rec, src=src, pred=pred, qp=qp)[0]
