"""The bitstream of decided frames (CPU: the library's writer is host code).

The oracle (oracle/bitstream.py) is pinned by the reference's own known answers and by
hand-derived clause 8.4.1.3 cases; the library's rh264_write_* entries must write the
oracle's bytes for the same decisions, and the oracle's parser must read back the
decisions, the parameter sets and the I_PCM samples."""
import numpy as np
import pytest

from oracle import bitstream as B
from rusty_h264_amd import _lib, accel


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


# ---- the oracle against the reference's known answers ----
def test_exp_golomb_known_strings():
    # H.264 Table 9-2 / 9-3 strings, as bit_writer.rs:190-228 lists them
    for v, s in [(0, "1"), (1, "010"), (2, "011"), (3, "00100"), (4, "00101"), (5, "00110"), (6, "00111"),
                 (7, "0001000"), (8, "0001001")]:
        w = B.BitWriter()
        w.ue(v)
        assert w.bitstring() == s
    for v, s in [(0, "1"), (1, "010"), (-1, "011"), (2, "00100"), (-2, "00101"), (3, "00110"), (-3, "00111")]:
        w = B.BitWriter()
        w.se(v)
        assert w.bitstring() == s
    w = B.BitWriter()
    w.u(0b101, 3)
    w.trailing()
    assert w.bytes() == bytes([0b1011_0000])  # bit_writer.rs:239-245


def test_emulation_prevention_known_answer():
    # nal.rs:148-153: five zeros then a one escape twice
    assert B.escape(bytes([0, 0, 0, 0, 0, 1])) == bytes([0, 0, 3, 0, 0, 3, 0, 1])
    for p in ([0, 0, 0, 1], [0, 0, 1, 2, 3], [1, 2, 3, 4, 5], [0] * 6, [0, 0, 3, 0, 0, 3]):
        assert B.unescape(B.escape(bytes(p))) == bytes(p)  # nal.rs:156-170
    assert B.nal(3, 7, b"")[4] == 0b0110_0111  # nal.rs:143-146


def test_parameter_sets_known_fields():
    # params.rs:226-259 (1920x1080, Constrained Baseline, CAVLC)
    sps = B.parse_sps(B.split_annex_b(B.headers(1920, 1080))[0][1])
    assert sps["profile_idc"] == 66 and (sps["constraints"] >> 6) & 1 == 1 and sps["level_idc"] == 30
    assert (sps["sps_id"], sps["log2_max_frame_num_minus4"], sps["poc_type"], sps["log2_max_poc_lsb_minus4"]) == (0,) * 4
    assert sps["max_num_ref_frames"] == 1 and sps["gaps"] == 0
    assert (sps["width_mbs_minus1"], sps["height_map_units_minus1"]) == (119, 67)
    assert sps["frame_mbs_only"] == 1 and sps["direct_8x8_inference"] == 1
    assert sps["cropping"] == 1 and sps["crop"] == (0, 0, 0, 4)
    # params.rs:261-276 (640x480)
    pps = B.parse_pps(B.split_annex_b(B.headers(640, 480))[1][1])
    assert (pps["pps_id"], pps["sps_id"], pps["entropy_coding_mode"], pps["bottom_field_poc"]) == (0, 0, 0, 0)
    assert (pps["num_slice_groups_minus1"], pps["num_ref_idx_l0_minus1"], pps["num_ref_idx_l1_minus1"]) == (0, 0, 0)
    assert (pps["weighted_pred"], pps["weighted_bipred_idc"], pps["pic_init_qp_minus26"]) == (0, 0, 0)


def _grid_with(mbw, mbh, blocks):
    g = B.MotionGrid(mbw, mbh)
    for (bx, by, w4, h4, val) in blocks:
        g.put(bx, by, w4, h4, val)
    return g


def test_mvp_16x16_reference_known_answers():
    # inter.rs:1993-2034: single neighbour, median of three, one matching reference
    g = _grid_with(3, 2, [(0, 4, 4, 4, (0, 4, -8))])  # only A (the left macroblock) decoded
    assert B.mvp(g, 1, 1, 0, 0, 0) == (4, -8)  # B and C unavailable -> A
    g = _grid_with(3, 2, [(0, 4, 4, 4, (0, 2, 0)), (4, 0, 4, 4, (0, 6, 5)), (8, 0, 4, 4, (0, -4, 9))])
    assert B.mvp(g, 1, 1, 0, 0, 0) == (2, 5)  # median per component
    g = _grid_with(3, 2, [(0, 4, 4, 4, (1, 2, 0)), (4, 0, 4, 4, (0, 6, 5)), (8, 0, 4, 4, (1, -4, 9))])
    assert B.mvp(g, 1, 1, 0, 0, 0) == (6, 5)  # exactly one neighbour of reference 0
    assert B.mvp(g, 1, 1, 0, 0, 1) == (2, 5)  # two of reference 1 -> the median


def test_mvp_directional_and_c_to_d():
    # 16x8: the upper half takes B of its reference, the lower half A; 8x16: the left
    # half takes A, the right half C (8.4.1.3); a C below the decoded area is replaced by D
    g = _grid_with(3, 3, [(0, 4, 4, 4, (0, 10, 1)), (4, 0, 4, 4, (0, 20, 2)), (8, 0, 4, 4, (0, 30, 3)),
                          (0, 0, 4, 4, (0, 40, 4))])
    assert B.mvp(g, 1, 1, 0, 1, 0) == (20, 2)  # 16x8 top: B
    g.put(4, 4, 4, 2, (1, 7, 7))  # the top half decided with reference 1
    assert B.mvp(g, 1, 1, 1, 1, 0) == (10, 1)  # 16x8 bottom: A
    # reference 1: A (ref 0), B = the top half (ref 1), C right of the macroblock is not
    # decoded -> D = left macroblock (ref 0): one match -> B
    assert B.mvp(g, 1, 1, 1, 1, 1) == (7, 7)
    g = _grid_with(3, 3, [(0, 4, 4, 4, (0, 10, 1)), (4, 0, 4, 4, (0, 20, 2)), (8, 0, 4, 4, (0, 30, 3)),
                          (0, 0, 4, 4, (0, 40, 4))])
    assert B.mvp(g, 1, 1, 0, 2, 0) == (10, 1)  # 8x16 left: A
    g.put(4, 4, 2, 4, (1, 5, 5))
    assert B.mvp(g, 1, 1, 1, 2, 0) == (30, 3)  # 8x16 right: C (above-right macroblock)
    # 8x8: the bottom-right quadrant's C lies right of the macroblock (not decoded): D
    g = _grid_with(3, 3, [(4, 4, 2, 2, (0, 1, 1)), (6, 4, 2, 2, (0, 3, 3)), (4, 6, 2, 2, (0, 5, 5))])
    assert B.mvp(g, 1, 1, 3, 3, 0) == (3, 3)  # A (5,5), B (3,3), D (1,1) -> median (3, 3)


def test_skip_mv_rules():
    g = B.MotionGrid(2, 2)
    assert B.skip_mv(g, 0, 0) == (0, 0)  # no neighbours
    g.put(0, 0, 4, 4, (0, 0, 0))
    g.put(4, 0, 4, 4, (0, 8, 8))
    g.put(0, 4, 4, 4, (0, 4, 4))
    # A = (4, 4), B = (8, 8), C outside -> D = the above-left zero vector: the zero test
    # is on A and B only, so the median (4, 4)
    assert B.skip_mv(g, 1, 1) == (4, 4)
    g.put(0, 0, 4, 4, (0, 20, 20))
    assert B.skip_mv(g, 1, 1) == (8, 8)
    g.put(0, 4, 4, 4, (0, 0, 0))  # A a zero vector of reference 0 -> (0, 0)
    assert B.skip_mv(g, 1, 1) == (0, 0)
    g.put(0, 4, 4, 4, (1, 0, 0))  # of reference 1: not the rule
    assert B.skip_mv(g, 1, 1) == (8, 8)


# ---- the library against the oracle ----
def _random_records(rng, mbw, mbh, nrefs, p_split=0.6, mvr=40, p_skiplike=0.2):
    n = mbw * mbh
    rec = np.zeros((n, 14), np.int32)
    for m in range(n):
        mode = int(rng.integers(0, 4)) if rng.random() < p_split else 0
        quads = []
        base = [int(rng.integers(0, nrefs)), int(rng.integers(-mvr, mvr + 1)), int(rng.integers(-mvr, mvr + 1))]
        if rng.random() < p_skiplike:
            base = [0, 0, 0]
        for q in range(4):
            quads.append([int(rng.integers(0, nrefs)), int(rng.integers(-mvr, mvr + 1)), int(rng.integers(-mvr, mvr + 1))])
        same = {0: [0, 0, 0, 0], 1: [0, 0, 2, 2], 2: [0, 1, 0, 1], 3: [0, 1, 2, 3]}[mode]
        quads[0] = base
        rec[m, 0] = mode
        for q in range(4):
            rec[m, 2 + 3 * q:5 + 3 * q] = quads[same[q]]
    return rec


@pytest.mark.parametrize("w,h,nref,nact", [(64, 48, 1, 1), (176, 144, 2, 2), (208, 160, 4, 3), (352, 288, 4, 4),
                                           (100, 70, 3, 1)])
def test_library_writes_the_oracle_bytes(lib, w, h, nref, nact):
    rng = np.random.default_rng(w * 7 + h + nref)
    s = accel.seq(w, h, qp=26, num_ref_frames=nref)
    assert accel.write_headers(lib, s) == B.headers(w, h, 26, nref)
    mbw, mbh = (w + 15) // 16, (h + 15) // 16
    y = rng.integers(0, 256, (mbh * 16, mbw * 16)).astype(np.uint8)
    u = rng.integers(0, 256, (mbh * 8, mbw * 8)).astype(np.uint8)
    v = rng.integers(0, 256, (mbh * 8, mbw * 8)).astype(np.uint8)
    assert accel.write_idr_pcm(lib, s, y, u, v) == B.idr_pcm(w, h, y, u, v)
    for fn in range(1, 4):
        rec = _random_records(rng, mbw, mbh, nact)
        got = accel.write_p_slice(lib, s, fn, 2 * fn, nact, rec)
        assert got == B.p_slice(w, h, nref, fn, 2 * fn, nact, rec), (fn, nact)


def test_parser_reads_back_the_decisions(lib):
    rng = np.random.default_rng(5)
    w, h, nref = 144, 112, 3
    mbw, mbh = w // 16, h // 16
    s = accel.seq(w, h, qp=30, num_ref_frames=nref)
    y = rng.integers(0, 256, (h, w)).astype(np.uint8)
    u = rng.integers(0, 256, (h // 2, w // 2)).astype(np.uint8)
    v = rng.integers(0, 256, (h // 2, w // 2)).astype(np.uint8)
    stream = accel.write_headers(lib, s) + accel.write_idr_pcm(lib, s, y, u, v)
    recs = [_random_records(rng, mbw, mbh, k) for k in (1, 2, 3)]
    for i, rec in enumerate(recs):
        stream += accel.write_p_slice(lib, s, i + 1, 2 * (i + 1), i + 1, rec)
    units = B.parse_stream(stream)
    assert [u_[0] for u_ in units] == ["sps", "pps", "slice", "slice", "slice", "slice"]
    assert units[1][2]["pic_init_qp_minus26"] == 4 and units[0][2]["max_num_ref_frames"] == 3
    hdr, mbs = units[2][3], units[2][4]
    assert units[2][2] == 5 and hdr["slice_type"] == 7
    pcm_y = np.block([[mbs[my * mbw + mx][1] for mx in range(mbw)] for my in range(mbh)])
    assert np.array_equal(pcm_y, y)
    assert np.array_equal(np.block([[mbs[my * mbw + mx][2] for mx in range(mbw)] for my in range(mbh)]), u)
    nskip = 0
    for i, rec in enumerate(recs):
        hdr, mbs = units[3 + i][3], units[3 + i][4]
        assert hdr["frame_num"] == i + 1 and hdr["poc_lsb"] == 2 * (i + 1) and hdr["num_ref_idx_active"] == i + 1
        for m, (mode, quads, skipped) in enumerate(mbs):
            want = [tuple(int(x) for x in rec[m, 2 + 3 * q:5 + 3 * q]) for q in range(4)]
            assert quads == want, (i, m)
            assert mode == (0 if skipped else rec[m, 0])
            nskip += skipped
    assert nskip > 0  # the P_Skip path is exercised


def test_writer_rejects_malformed_records(lib):
    s = accel.seq(32, 32, num_ref_frames=2)
    rec = np.zeros((4, 14), np.int32)
    rec[0, 0] = 5
    with pytest.raises(RuntimeError):
        accel.write_p_slice(lib, s, 1, 2, 2, rec)
    rec[0, 0] = 1
    rec[0, 2 + 3 * 2] = 1  # bottom half of a 16x8 whose quadrants differ
    with pytest.raises(RuntimeError):
        accel.write_p_slice(lib, s, 1, 2, 2, rec)
    rec[0, 2 + 3 * 2] = 0
    rec[1, 2] = 2  # reference 2 with 2 active
    with pytest.raises(RuntimeError):
        accel.write_p_slice(lib, s, 1, 2, 2, rec)


def _mvp_mutant(kind):
    def mvp(grid, mx, my, part, shape, ref):
        x4, y4, w4, _ = B.SHAPES[shape][part]
        bx, by = mx * 4 + x4, my * 4 + y4
        A, Bn, Cn = grid.get(bx - 1, by), grid.get(bx, by - 1), grid.get(bx + w4, by - 1)
        if Cn is None and kind != "no_c_to_d":
            Cn = grid.get(bx - 1, by - 1)
        na = (-1, 0, 0)
        rA, rB, rC = (A or na), (Bn or na), (Cn or na)
        if shape == 1:
            if part == 0 and rB[0] == ref:
                return rB[1], rB[2]
            lower = rB if kind == "16x8_lower_takes_B" else rA
            if part == 1 and lower[0] == ref:
                return lower[1], lower[2]
        if shape == 2:
            if part == 0 and rA[0] == ref:
                return rA[1], rA[2]
            if part == 1 and rC[0] == ref:
                return rC[1], rC[2]
        if Bn is None and Cn is None and A is not None and kind != "no_copy_of_a":
            rB = rC = rA
        match = [n for n in (rA, rB, rC) if n[0] == ref]
        if len(match) == 1:
            return match[0][1], match[0][2]
        return B.median(rA[1], rB[1], rC[1]), B.median(rA[2], rB[2], rC[2])
    return mvp


@pytest.mark.parametrize("kind", ["16x8_lower_takes_B", "no_c_to_d", "no_copy_of_a"])
def test_planted_oracle_mistakes_are_caught(lib, monkeypatch, kind):
    """A copy of the oracle's predictor with one plausible mistake no longer writes the
    library's bytes (so the byte comparison above would catch it in either)."""
    rng = np.random.default_rng(11)
    w, h, nref = 176, 144, 2
    s = accel.seq(w, h, num_ref_frames=nref)
    recs = [_random_records(rng, w // 16, h // 16, nref) for _ in range(3)]
    lib_bytes = [accel.write_p_slice(lib, s, 1, 2, nref, r) for r in recs]
    assert lib_bytes == [B.p_slice(w, h, nref, 1, 2, nref, r) for r in recs]
    monkeypatch.setattr(B, "mvp", _mvp_mutant(kind))
    assert lib_bytes != [B.p_slice(w, h, nref, 1, 2, nref, r) for r in recs]
