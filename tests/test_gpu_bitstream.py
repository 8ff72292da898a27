"""The bitstream of GPU-decided frames: the same bytes as from the oracle's decisions,
the decisions read back by the oracle's parser, P_Skip exactly where the GPU's own
P_Skip vector (rh264_decision_predict) equals a reference-0 decision, and the first P
frame's decoded luma (motion compensation of the parsed vectors on the decoded I_PCM
picture) equal to the GPU's prediction of the decided frame."""
import numpy as np
import pytest

from oracle import bitstream as B
from oracle.oracle import Oracle
from rusty_h264_amd import _lib, accel, synth
from tests.test_oracle_walkers import partition_clip

pytestmark = pytest.mark.gpu
O = Oracle(fast=True)
PAD = 32


@pytest.fixture(scope="module")
def dev():
    d = accel.Device(0)
    print("device:", d.name(), flush=True)
    yield d
    d.close()


@pytest.mark.parametrize("nrefs,qp", [(1, 26), (2, 20), (3, 26)])
def test_partition_decisions_write_the_same_stream(dev, nrefs, qp):
    cw, ch = 176, 144
    lib = _lib.load()
    cur, refs = partition_clip(cw, ch, nrefs, 300 + qp + nrefs)
    dref = dev.upload(np.stack(refs[::-1]))
    pl = accel.build_planes(dev, dref, nrefs, cw, ch, PAD)
    dcur = dev.upload(cur)
    lam, gate = accel.lambda_me(qp), accel.split_gate(qp)
    p = accel.search_params(cw, ch, PAD, 16, 16, 16, 0, lam)
    print("partition stream", nrefs, qp, flush=True)
    out = accel.partition_decision(dev, p, 1, dcur, pl, nrefs=nrefs, lambda_ref=lam, gate=gate, fc=True)
    got = dev.download(out, ((cw // 16) * (ch // 16), 14), np.int32)
    want = O.partition_decision_frame(O.cfg(cw, ch, PAD, 16, 16, 16, 0, lam), cur,
                                      [O.planes(r, cw, ch, PAD) for r in refs], lam, gate, True, fc=True)
    s = accel.seq(cw, ch, qp=qp, num_ref_frames=nrefs)
    gpu_bytes = accel.write_p_slice(lib, s, 1, 2, nrefs, got)
    assert gpu_bytes == accel.write_p_slice(lib, s, 1, 2, nrefs, want)
    assert gpu_bytes == B.p_slice(cw, ch, nrefs, 1, 2, nrefs, got)
    units = B.parse_stream(accel.write_headers(lib, s) + gpu_bytes)
    mbs = units[2][4]
    assert len(set(int(m) for m in got[:, 0])) >= 3  # several shapes in the stream
    for m, (mode, quads, skipped) in enumerate(mbs):
        assert quads == [tuple(int(x) for x in got[m, 2 + 3 * q:5 + 3 * q]) for q in range(4)], m
        assert skipped or mode == got[m, 0]
    for x in (dref, pl, dcur, out):
        x.free()


def test_decided_clip_decodes_to_the_prediction(dev):
    cw, ch, nf = 352, 288, 3
    lib = _lib.load()
    clip = synth.cif_clip(nf + 1, seed=91, w=cw, h=ch)
    lam = accel.lambda_me(26)
    cur, refs = dev.upload(clip[1:]), dev.upload(clip[:-1])
    pl = accel.build_planes(dev, refs, nf, cw, ch, PAD)
    nb = (cw // 16) * (ch // 16)
    p = accel.search_params(cw, ch, PAD, 16, 16, 32, 0, lam)
    print("decided clip stream", cw, ch, nf, flush=True)
    ref_d, mv_d, cost_d = accel.motion_search_wavefront(dev, p, nf, cur, pl, nrefs=1, fc=True)
    ref = dev.download(ref_d, (nf, nb), np.int32)
    mv = dev.download(mv_d, (nf, nb, 2), np.int32)
    smv, pred, _ = accel.decision_predict(dev, cw, ch, PAD, nf, 1, pl, ref_d, mv_d)
    s = accel.seq(cw, ch, qp=26, num_ref_frames=1)
    chroma = np.full((ch // 2, cw // 2), 128, np.uint8)
    stream = accel.write_headers(lib, s) + accel.write_idr_pcm(lib, s, clip[0], chroma, chroma)
    cfg = O.cfg(cw, ch, PAD, 16, 16, 32, 0, lam)
    for f in range(nf):
        rec = accel.records_16x16(ref[f], mv[f])
        bytes_f = accel.write_p_slice(lib, s, f + 1, 2 * (f + 1), 1, rec)
        # the oracle's raster-order decision of the same frame writes the same bytes
        w_ref, w_mv, _ = O.motion_search_frame(cfg, clip[f + 1], O.planes(clip[f], cw, ch, PAD), fc=True)
        assert bytes_f == accel.write_p_slice(lib, s, f + 1, 2 * (f + 1), 1, accel.records_16x16(w_ref, w_mv)), f
        stream += bytes_f
    units = B.parse_stream(stream)
    assert [u[0] for u in units] == ["sps", "pps"] + ["slice"] * (nf + 1)
    idr = units[2][4]
    dec0 = np.block([[idr[my * (cw // 16) + mx][1] for mx in range(cw // 16)] for my in range(ch // 16)])
    assert np.array_equal(dec0, clip[0])
    for f in range(nf):
        mbs = units[3 + f][4]
        sl = slice(f * nb, (f + 1) * nb)
        skipped = np.array([m[2] for m in mbs])
        # P_Skip exactly where the decision is reference 0 at the GPU's P_Skip vector
        assert np.array_equal(skipped, (ref[f] == 0) & (mv[f] == smv[sl]).all(1)), f
        parsed = np.array([m[1][0][1:] for m in mbs], np.int32)
        assert np.array_equal(parsed, mv[f]), f
    # frame 1 predicts from the I_PCM picture, which decodes to the source: its decoded
    # luma (no residual coded) is the GPU's prediction of the decided frame
    dec1 = np.zeros((ch, cw), np.uint8)
    for m, (_, quads, _) in enumerate(units[3][4]):
        by, bx = (m // (cw // 16)) * 16, (m % (cw // 16)) * 16
        dec1[by:by + 16, bx:bx + 16] = O.mc_luma(dec0, cw, ch, bx, by, 16, 16, quads[0][1], quads[0][2])
    assert np.array_equal(dec1, pred[0])
    assert skipped.sum() > 0
    for x in (cur, refs, pl, ref_d, mv_d, cost_d):
        x.free()


@pytest.mark.parametrize("qp", [22, 32])
def test_residual_coded_clip_decodes(dev, qp):
    """A CIF clip searched on the GPU (plain path, 16x16) and coded with the luma
    residual against the previous reconstruction (accel.encode_p_frames: GPU plane
    build and prediction, host residual writer): the GPU's prediction of every picture
    equals the oracle's mc_luma on the reconstruction, the library's bytes equal the
    oracle writer's, and the oracle decodes the stream alone to the reconstructions."""
    cw, ch, nf = 352, 288, 4
    lib = _lib.load()
    clip = synth.cif_clip(nf + 1, seed=93 + qp, w=cw, h=ch)
    lam = accel.lambda_me(qp)
    cur, refs = dev.upload(clip[1:]), dev.upload(clip[:-1])
    pl = accel.build_planes(dev, refs, nf, cw, ch, PAD)
    ms = accel.MotionSearch(dev, accel.search_params(cw, ch, PAD, 16, 16, 32, 0, lam))
    print("residual clip", cw, ch, nf, "qp", qp, flush=True)
    ms.run_plain(cur, pl, nf)
    mv = ms.results(nf)["mv"]
    stream, sizes, recon = accel.encode_p_frames(dev, clip, mv, qp, pad=PAD)
    nb, mbw = (cw // 16) * (ch // 16), cw // 16
    units = B.parse_stream(stream)
    assert [u[0] for u in units] == ["sps", "pps"] + ["slice"] * (nf + 1)
    decoded = [np.block([[units[2][4][my * mbw + mx][1] for mx in range(mbw)] for my in range(ch // 16)])]
    assert np.array_equal(decoded[0], clip[0])
    s = accel.seq(cw, ch, qp=qp)
    coded = 0
    for f in range(1, nf + 1):
        hdr, mbs = units[2 + f][3], units[2 + f][4]
        fmv = mv[(f - 1) * nb:f * nb]
        # the GPU's prediction on the reconstruction, through the oracle's mc_luma
        pred = np.zeros((ch, cw), np.uint8)
        for m in range(nb):
            by, bx = (m // mbw) * 16, (m % mbw) * 16
            pred[by:by + 16, bx:bx + 16] = O.mc_luma(decoded[-1], cw, ch, bx, by, 16, 16, int(fmv[m, 0]), int(fmv[m, 1]))
        rec = accel.records_16x16(np.zeros(nb, np.int32), fmv)
        want, wrecon = B.p_slice(cw, ch, 1, f, 2 * f, 1, rec, src=clip[f], pred=pred, qp=qp)
        off = len(stream) - sum(sizes[f:])  # the library's bytes of picture f
        assert stream[off:off + sizes[f]] == want, f
        assert np.array_equal(wrecon, recon[f]), f
        out = pred.copy()
        for m, (mode, quads, _) in enumerate(mbs):
            assert quads[0][1:] == (int(fmv[m, 0]), int(fmv[m, 1])), (f, m)
            if hdr["levels"][m] is not None:
                B.reconstruct(pred, hdr["levels"][m], m % mbw, m // mbw, qp, out)
                coded += 1
        assert np.array_equal(out, recon[f]), f
        decoded.append(out)
    assert coded > 0
    for x in (cur, refs, pl):
        x.free()
