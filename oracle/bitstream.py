"""Oracle for the bitstream of decided frames (test infrastructure; only tests/ use it).

A plain-Python restatement of the H.264 syntax the library's rh264_write_* entries
write -- the reference's CAVLC configuration -- and a parser for that subset (the
decode side of the check):

  SPS / PPS             crates/rusty_h264-encoder/src/params.rs:56-101, 148-172
  slice headers         crates/rusty_h264-encoder/src/slice.rs:49-107
  P macroblock layer    crates/rusty_h264-encoder/src/mb16.rs:5380-5776, 4416-4435,
                        5957-5964; crates/rusty_h264-common/src/cavlc.rs:136-138
  exp-Golomb, NAL       crates/rusty_h264-common/src/bit_writer.rs, nal.rs:78-117

Motion vector prediction is restated from ITU-T H.264 8.4.1.3 (with 6.4.11.7 for the
neighbouring partitions) on a 4x4-block motion grid, and P_Skip from 8.4.1.1.

The luma residual (forward core transform, inter quantiser, CAVLC residual blocks and
the 8.5.12 reconstruction) is restated from the same sources and H.264 clause 9.2;
its pins are in tests/test_bitstream_residual.py (two textbook CAVLC known answers,
prefix-free tables, write/read round trips in every nC context, inverse-transform
closed forms, the quantiser's reconstruction error at every QP, planted mistakes).

Pins (tests/test_bitstream.py): the reference's own known answers (Table 9-2/9-3
exp-Golomb strings, bit_writer.rs:187-228; emulation prevention and the NAL header
byte, nal.rs:143-160; the SPS/PPS field values of params.rs:226-276), the predict_mv
known answers of inter.rs:1993-2034 for 16x16, and hand-derived 8.4.1.3 cases for the
16x8 / 8x16 directional predictors and the C -> D substitution.
"""
import numpy as np


class BitWriter:
    def __init__(self):
        self.bits = []

    def u(self, v, n):
        self.bits += [(v >> (n - 1 - i)) & 1 for i in range(n)]

    def ue(self, v):
        c = v + 1
        n = c.bit_length()
        self.bits += [0] * (n - 1) + [(c >> (n - 1 - i)) & 1 for i in range(n)]

    def se(self, v):
        self.ue(2 * v - 1 if v > 0 else -2 * v)

    def align(self):
        while len(self.bits) % 8:
            self.bits.append(0)

    def trailing(self):
        self.bits.append(1)
        self.align()

    def bytes(self):
        assert len(self.bits) % 8 == 0
        return bytes(int("".join(map(str, self.bits[i:i + 8])), 2) for i in range(0, len(self.bits), 8))

    def bitstring(self):
        return "".join(map(str, self.bits))


def escape(rbsp):
    """Emulation prevention (nal.rs:99-117): 03 before a byte <= 3 after two zeros."""
    out, zeros = bytearray(), 0
    for b in rbsp:
        if zeros >= 2 and b <= 3:
            out.append(3)
            zeros = 0
        out.append(b)
        zeros = zeros + 1 if b == 0 else 0
    return bytes(out)


def unescape(ebsp):
    out, zeros = bytearray(), 0
    i = 0
    while i < len(ebsp):
        b = ebsp[i]
        if zeros >= 2 and b == 3:
            zeros = 0
            i += 1
            continue
        out.append(b)
        zeros = zeros + 1 if b == 0 else 0
        i += 1
    return bytes(out)


def nal(ref_idc, nal_type, rbsp):
    return b"\x00\x00\x00\x01" + bytes([(ref_idc << 5) | nal_type]) + escape(rbsp)


def split_annex_b(stream):
    """NAL units of an Annex-B stream with 4-byte start codes: (header byte, rbsp)."""
    parts = stream.split(b"\x00\x00\x00\x01")
    assert parts[0] == b""
    return [(p[0], unescape(p[1:])) for p in parts[1:]]


# ---------------------------------------------------------------------------- writer
def headers(width, height, qp=26, num_ref_frames=1, level_idc=30):
    mbw, mbh = (width + 15) // 16, (height + 15) // 16
    cr, cb = (mbw * 16 - width) // 2, (mbh * 16 - height) // 2
    w = BitWriter()
    w.u(66, 8)
    w.u(1 << 6, 8)
    w.u(level_idc, 8)
    for v in (0, 0, 0, 0, num_ref_frames):
        w.ue(v)
    w.u(0, 1)
    w.ue(mbw - 1)
    w.ue(mbh - 1)
    w.u(1, 1)
    w.u(1, 1)
    crop = 1 if (cr or cb) else 0
    w.u(crop, 1)
    if crop:
        for v in (0, cr, 0, cb):
            w.ue(v)
    w.u(0, 1)
    w.trailing()
    p = BitWriter()
    p.ue(0)
    p.ue(0)
    p.u(0, 1)
    p.u(0, 1)
    p.ue(0)
    p.ue(num_ref_frames - 1)
    p.ue(0)
    p.u(0, 1)
    p.u(0, 2)
    p.se(qp - 26)
    p.se(0)
    p.se(0)
    p.u(1, 1)
    p.u(0, 1)
    p.u(0, 1)
    p.trailing()
    return nal(3, 7, w.bytes()) + nal(3, 8, p.bytes())


def idr_pcm(width, height, y, u, v):
    mbw, mbh = (width + 15) // 16, (height + 15) // 16
    w = BitWriter()
    for val in (0, 7, 0):
        w.ue(val)
    w.u(0, 4)
    w.ue(0)
    w.u(0, 4)
    w.u(0, 1)
    w.u(0, 1)
    w.se(0)
    w.ue(0)
    w.se(0)
    w.se(0)
    for my in range(mbh):
        for mx in range(mbw):
            w.ue(25)
            w.align()
            for s in y[my * 16:my * 16 + 16, mx * 16:mx * 16 + 16].reshape(-1):
                w.u(int(s), 8)
            for pl in (u, v):
                for s in pl[my * 8:my * 8 + 8, mx * 8:mx * 8 + 8].reshape(-1):
                    w.u(int(s), 8)
    w.trailing()
    return nal(3, 5, w.bytes())


# partition shapes: (x4, y4, w4, h4) of each partition in 4x4 blocks
SHAPES = {0: [(0, 0, 4, 4)], 1: [(0, 0, 4, 2), (0, 2, 4, 2)], 2: [(0, 0, 2, 4), (2, 0, 2, 4)],
          3: [(0, 0, 2, 2), (2, 0, 2, 2), (0, 2, 2, 2), (2, 2, 2, 2)]}


class MotionGrid:
    """Decoded (ref, mvx, mvy) per 4x4 block of a picture; None = not decoded."""

    def __init__(self, mbw, mbh):
        self.mbw, self.mbh = mbw, mbh
        self.g = [[None] * (mbw * 4) for _ in range(mbh * 4)]

    def get(self, bx, by):
        """Block (bx, by) in picture 4x4 units: its motion or None (outside the picture
        or not decoded yet: 6.4.11.7 / 6.4.12, availability in decoding order)."""
        if bx < 0 or by < 0 or bx >= self.mbw * 4 or by >= self.mbh * 4:
            return None
        return self.g[by][bx]

    def put(self, bx, by, w4, h4, val):
        for j in range(h4):
            for i in range(w4):
                self.g[by + j][bx + i] = val


def median(a, b, c):
    return sorted((a, b, c))[1]


def mvp(grid, mx, my, part, shape, ref):
    """8.4.1.3 for partition `part` of `shape` in macroblock (mx, my)."""
    x4, y4, w4, _ = SHAPES[shape][part]
    bx, by = mx * 4 + x4, my * 4 + y4
    A = grid.get(bx - 1, by)
    B = grid.get(bx, by - 1)
    C = grid.get(bx + w4, by - 1)
    if C is None:
        C = grid.get(bx - 1, by - 1)  # D
    na = (-1, 0, 0)
    rA, rB, rC = (A or na), (B or na), (C or na)
    if shape == 1:
        if part == 0 and rB[0] == ref:
            return rB[1], rB[2]
        if part == 1 and rA[0] == ref:
            return rA[1], rA[2]
    if shape == 2:
        if part == 0 and rA[0] == ref:
            return rA[1], rA[2]
        if part == 1 and rC[0] == ref:
            return rC[1], rC[2]
    if B is None and C is None and A is not None:
        rB = rC = rA
    match = [n for n in (rA, rB, rC) if n[0] == ref]
    if len(match) == 1:
        return match[0][1], match[0][2]
    return median(rA[1], rB[1], rC[1]), median(rA[2], rB[2], rC[2])


def skip_mv(grid, mx, my):
    """8.4.1.1: (0, 0) when A or B is not available or is a zero vector of reference 0."""
    A, B = grid.get(mx * 4 - 1, my * 4), grid.get(mx * 4, my * 4 - 1)
    if A is None or B is None or A == (0, 0, 0) or B == (0, 0, 0):
        return 0, 0
    return mvp(grid, mx, my, 0, 0, 0)


def _quad_of(shape, part):
    x4, y4, _, _ = SHAPES[shape][part]
    return (y4 // 2) * 2 + x4 // 2


def p_slice(width, height, num_ref_frames, frame_num, poc_lsb, nrefs, rec, src=None, pred=None, qp=26):
    """The P picture of decision records rec [mb][14] (mode, cost, TL/TR/BL/BR ref mvx mvy).
    With src and pred (coded-size luma planes) the residual src - pred is coded at qp and
    (bytes, reconstructed luma) is returned; without them no residual (bytes only)."""
    mbw, mbh = (width + 15) // 16, (height + 15) // 16
    res = src is not None
    recon = np.zeros((mbh * 16, mbw * 16), np.uint8) if res else None
    tc = np.zeros((mbh * 4, mbw * 4), np.int64)
    rec = np.asarray(rec).reshape(mbh * mbw, 14)
    w = BitWriter()
    for val in (0, 5, 0):
        w.ue(val)
    w.u(frame_num & 15, 4)
    w.u(poc_lsb & 15, 4)
    if num_ref_frames > 1:
        w.u(1, 1)
        w.ue(nrefs - 1)
    else:
        w.u(0, 1)
    w.u(0, 1)
    w.u(0, 1)
    w.se(0)
    if res:
        w.ue(1)  # disable_deblocking_filter_idc 1: the reconstruction is the decoder's output
    else:
        w.ue(0)
        w.se(0)
        w.se(0)
    nact = nrefs if num_ref_frames > 1 else 1
    grid = MotionGrid(mbw, mbh)
    run = 0
    for my in range(mbh):
        for mx in range(mbw):
            r = [int(v) for v in rec[my * mbw + mx]]
            mode = r[0]
            q = lambda k: r[2 + 3 * k:5 + 3 * k]
            lv, cbp = mb_residual(src, pred, mx, my, qp) if res else (None, 0)
            if mode == 0 and q(0)[0] == 0 and cbp == 0 and tuple(q(0)[1:]) == skip_mv(grid, mx, my):
                run += 1
                grid.put(mx * 4, my * 4, 4, 4, (0, q(0)[1], q(0)[2]))
                if res:
                    recon[my * 16:my * 16 + 16, mx * 16:mx * 16 + 16] = pred[my * 16:my * 16 + 16, mx * 16:mx * 16 + 16]
                continue
            w.ue(run)
            run = 0
            w.ue(mode)
            if mode == 3:
                for _ in range(4):
                    w.ue(0)
            parts = range(len(SHAPES[mode]))
            if nact > 1:
                for k in parts:
                    ref = q(_quad_of(mode, k))[0]
                    if nact == 2:
                        w.u(1 - ref, 1)
                    else:
                        w.ue(ref)
            for k in parts:
                ref, x, y = q(_quad_of(mode, k))
                px, py = mvp(grid, mx, my, k, mode, ref)
                w.se(x - px)
                w.se(y - py)
                x4, y4, w4, h4 = SHAPES[mode][k]
                grid.put(mx * 4 + x4, my * 4 + y4, w4, h4, (ref, x, y))
            if res:
                write_residual(w, lv, cbp, mx, my, tc)
                reconstruct(pred, lv, mx, my, qp, recon)
            else:
                w.ue(0)
    if run:
        w.ue(run)
    w.trailing()
    return (nal(3, 1, w.bytes()), recon) if res else nal(3, 1, w.bytes())


# ---------------------------------------------------------------------------- residual
# The luma residual of a P macroblock (7.3.5.3, 8.5.12, 9.2): the reference's CAVLC
# residual (crates/rusty_h264-common/src/cavlc.rs) and its inter quantiser
# (crates/rusty_h264-common/src/transform.rs:92, 121, 172-174), restated.

# H.264 Table 9-5 (coeff_token), [table][TrailingOnes][TotalCoeff] for 0<=nC<2, 2<=nC<4, 4<=nC<8
COEFF_TOKEN = (
    (
        ('1', '000101', '00000111', '000000111', '0000000111', '00000000111', '0000000001111', '0000000001011', '0000000001000', '00000000001111', '00000000001011', '000000000001111', '000000000001011', '0000000000001111', '0000000000001011', '0000000000000111', '0000000000000100'),
        ('', '01', '000100', '00000110', '000000110', '0000000110', '00000000110', '0000000001110', '0000000001010', '00000000001110', '00000000001010', '000000000001110', '000000000001010', '000000000000001', '0000000000001110', '0000000000001010', '0000000000000110'),
        ('', '', '001', '0000101', '00000101', '000000101', '0000000101', '00000000101', '0000000001101', '0000000001001', '00000000001101', '00000000001001', '000000000001101', '000000000001001', '0000000000001101', '0000000000001001', '0000000000000101'),
        ('', '', '', '00011', '000011', '0000100', '00000100', '000000100', '0000000100', '00000000100', '0000000001100', '00000000001100', '00000000001000', '000000000001100', '000000000001000', '0000000000001100', '0000000000001000'),
    ),
    (
        ('11', '001011', '000111', '0000111', '00000111', '00000100', '000000111', '00000001111', '00000001011', '000000001111', '000000001011', '000000001000', '0000000001111', '0000000001011', '0000000000111', '00000000001001', '00000000000111'),
        ('', '10', '00111', '001010', '000110', '0000110', '00000110', '000000110', '00000001110', '00000001010', '000000001110', '000000001010', '0000000001110', '0000000001010', '00000000001011', '00000000001000', '00000000000110'),
        ('', '', '011', '001001', '000101', '0000101', '00000101', '000000101', '00000001101', '00000001001', '000000001101', '000000001001', '0000000001101', '0000000001001', '0000000000110', '00000000001010', '00000000000101'),
        ('', '', '', '0101', '0100', '00110', '001000', '000100', '0000100', '000000100', '00000001100', '00000001000', '000000001100', '0000000001100', '0000000001000', '0000000000001', '00000000000100'),
    ),
    (
        ('1111', '001111', '001011', '001000', '0001111', '0001011', '0001001', '0001000', '00001111', '00001011', '000001111', '000001011', '000001000', '0000001101', '0000001001', '0000000101', '0000000001'),
        ('', '1110', '01111', '01100', '01010', '01000', '001110', '001010', '0001110', '00001110', '00001010', '000001110', '000001010', '000000111', '0000001100', '0000001000', '0000000100'),
        ('', '', '1101', '01110', '01011', '01001', '001101', '001001', '0001101', '0001010', '00001101', '00001001', '000001101', '000001001', '0000001011', '0000000111', '0000000011'),
        ('', '', '', '1100', '1011', '1010', '1001', '1000', '01101', '001100', '0001100', '00001100', '00001000', '000001100', '0000001010', '0000000110', '0000000010'),
    ),
)
# Tables 9-7 / 9-8 (total_zeros of a 4x4 block), [TotalCoeff - 1][total_zeros]
TOTAL_ZEROS = (
    ('1', '011', '010', '0011', '0010', '00011', '00010', '000011', '000010', '0000011', '0000010', '00000011', '00000010', '000000011', '000000010', '000000001'),
    ('111', '110', '101', '100', '011', '0101', '0100', '0011', '0010', '00011', '00010', '000011', '000010', '000001', '000000'),
    ('0101', '111', '110', '101', '0100', '0011', '100', '011', '0010', '00011', '00010', '000001', '00001', '000000'),
    ('00011', '111', '0101', '0100', '110', '101', '100', '0011', '011', '0010', '00010', '00001', '00000'),
    ('0101', '0100', '0011', '111', '110', '101', '100', '011', '0010', '00001', '0001', '00000'),
    ('000001', '00001', '111', '110', '101', '100', '011', '010', '0001', '001', '000000'),
    ('000001', '00001', '101', '100', '011', '11', '010', '0001', '001', '000000'),
    ('000001', '0001', '00001', '011', '11', '10', '010', '001', '000000'),
    ('000001', '000000', '0001', '11', '10', '001', '01', '00001'),
    ('00001', '00000', '001', '11', '10', '01', '0001'),
    ('0000', '0001', '001', '010', '1', '011'),
    ('0000', '0001', '01', '1', '001'),
    ('000', '001', '1', '01'),
    ('00', '01', '1'),
    ('0', '1'),
)
# Table 9-10 (run_before), [min(zerosLeft, 7) - 1][run_before]
RUN_BEFORE = (
    ('1', '0',),
    ('1', '01', '00',),
    ('11', '10', '01', '00',),
    ('11', '10', '01', '001', '000',),
    ('11', '10', '011', '010', '001', '000',),
    ('11', '000', '001', '011', '010', '101', '100',),
    ('111', '110', '101', '100', '011', '010', '001', '0001', '00001', '000001', '0000001', '00000001', '000000001', '0000000001', '00000000001',),
)

ZIGZAG = (0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15)  # Table 8-13: scan -> raster
# Table 9-4, inter column: codeNum -> coded_block_pattern (chroma_format_idc 1)
CBP_INTER = (0, 16, 1, 2, 4, 8, 32, 3, 5, 10, 12, 15, 47, 7, 11, 13, 14, 6, 9, 31, 35, 37, 42, 44, 33, 34, 36, 40,
             39, 43, 45, 46, 17, 18, 20, 24, 19, 21, 26, 28, 23, 27, 29, 30, 22, 25, 38, 41)
CF = np.array([[1, 1, 1, 1], [2, 1, -1, -2], [1, -1, -1, 1], [1, -2, 2, -1]], np.int64)
MF = ((13107, 5243, 8066), (11916, 4660, 7490), (10082, 4194, 6554), (9362, 3647, 5825), (8192, 3355, 5243),
      (7282, 2893, 4559))
V = ((10, 16, 13), (11, 18, 14), (13, 20, 16), (14, 23, 18), (16, 25, 20), (18, 29, 23))


def _cls(i, j):
    """Position class of coefficient (i, j): 0 both even, 1 both odd, 2 mixed."""
    return 0 if i % 2 == 0 and j % 2 == 0 else 1 if i % 2 and j % 2 else 2


def forward4x4(x):
    """W = Cf X Cf^T (the textbook matrix form of the core transform)."""
    return CF @ np.asarray(x, np.int64).reshape(4, 4) @ CF.T


def quant4x4(w, qp):
    """Inter dead zone: |level| = (|W| MF + 2^qbits / 6) >> qbits, the sign of W."""
    qbits = 15 + qp // 6
    f = (1 << qbits) // 6
    z = np.zeros((4, 4), np.int64)
    for i in range(4):
        for j in range(4):
            a = abs(int(w[i][j]))
            lvl = (a * MF[qp % 6][_cls(i, j)] + f) >> qbits
            z[i][j] = -lvl if w[i][j] < 0 else lvl
    return z


def dequant4x4(z, qp):
    """8.5.12.1 with flat weights: d = c * LevelScale4x4 << (qp/6 - 4) for qp >= 24, else
    (c * LevelScale4x4 + 2^(3 - qp/6)) >> (4 - qp/6), LevelScale4x4 = 16 v."""
    d = np.zeros((4, 4), np.int64)
    for i in range(4):
        for j in range(4):
            ls = 16 * V[qp % 6][_cls(i, j)]
            c = int(z[i][j])
            d[i][j] = (c * ls) << (qp // 6 - 4) if qp >= 24 else (c * ls + (1 << (3 - qp // 6))) >> (4 - qp // 6)
    return d


def inverse4x4(d):
    """8.5.12.2: the row and column butterflies (with the >> 1 of the odd terms), then
    (x + 32) >> 6."""
    d = np.asarray(d, np.int64)
    f = np.zeros((4, 4), np.int64)
    for i in range(4):
        e0, e1 = d[i][0] + d[i][2], d[i][0] - d[i][2]
        e2, e3 = (d[i][1] >> 1) - d[i][3], d[i][1] + (d[i][3] >> 1)
        f[i] = (e0 + e3, e1 + e2, e1 - e2, e0 - e3)
    h = np.zeros((4, 4), np.int64)
    for j in range(4):
        g0, g1 = f[0][j] + f[2][j], f[0][j] - f[2][j]
        g2, g3 = (f[1][j] >> 1) - f[3][j], f[1][j] + (f[3][j] >> 1)
        h[0][j], h[1][j], h[2][j], h[3][j] = g0 + g3, g1 + g2, g1 - g2, g0 - g3
    return (h + 32) >> 6


def write_block(w, lv, nc):
    """residual_block_cavlc of 16 levels in scan order with context nc; returns TotalCoeff."""
    nz = [i for i in range(16) if lv[i]]
    n = len(nz)
    t1 = 0
    for k in range(n - 1, -1, -1):
        if t1 == 3 or abs(lv[nz[k]]) != 1:
            break
        t1 += 1
    if nc >= 8:
        w.u(3 if n == 0 else ((n - 1) << 2) | t1, 6)
    else:
        w.bits += [int(b) for b in COEFF_TOKEN[0 if nc < 2 else 1 if nc < 4 else 2][t1][n]]
    if n == 0:
        return 0
    for k in range(n - 1, n - 1 - t1, -1):
        w.u(1 if lv[nz[k]] < 0 else 0, 1)
    sl = 1 if n > 10 and t1 < 3 else 0
    for k in range(n - 1 - t1, -1, -1):
        v = lv[nz[k]]
        code = 2 * v - 2 if v > 0 else -2 * v - 1
        if k == n - 1 - t1 and t1 < 3:
            code -= 2
        if sl == 0 and code < 14:
            prefix, size, suffix = code, 0, 0
        elif sl == 0 and code < 30:
            prefix, size, suffix = 14, 4, code - 14
        elif sl == 0:
            prefix, size, suffix = 15, 12, code - 30
        elif code < (15 << sl):
            prefix, size, suffix = code >> sl, sl, code & ((1 << sl) - 1)
        else:
            prefix, size, suffix = 15, 12, code - (15 << sl)
        assert suffix < 4096, "level beyond the escape range"
        w.u(1, prefix + 1)
        if size:
            w.u(suffix, size)
        sl = max(sl, 1)
        if abs(v) > (3 << (sl - 1)) and sl < 6:
            sl += 1
    tz = nz[-1] + 1 - n
    if n < 16:
        w.bits += [int(b) for b in TOTAL_ZEROS[n - 1][tz]]
    left = tz
    for k in range(n - 1, 0, -1):
        if left <= 0:
            break
        run = nz[k] - nz[k - 1] - 1
        w.bits += [int(b) for b in RUN_BEFORE[min(left, 7) - 1][run]]
        left -= run
    return n


def _read_vlc(r, codes):
    """Index of the code word of `codes` (a prefix-free sequence, "" = unused) next in r."""
    cur = ""
    while True:
        cur += str(r.u(1))
        if cur in codes:
            return codes.index(cur)
        assert len(cur) <= 16, "no code word"


def read_block(r, nc):
    """The 16 levels (scan order) of residual_block_cavlc with context nc, and TotalCoeff."""
    if nc >= 8:
        v = r.u(6)
        n, t1 = (0, 0) if v == 3 else ((v >> 2) + 1, v & 3)
    else:
        tab = COEFF_TOKEN[0 if nc < 2 else 1 if nc < 4 else 2]
        flat = [tab[t][k] for k in range(17) for t in range(4)]
        i = _read_vlc(r, flat)
        n, t1 = i // 4, i % 4
    lv = [0] * 16
    if n == 0:
        return lv, 0
    levels = []
    for _ in range(t1):
        levels.append(-1 if r.u(1) else 1)
    sl = 1 if n > 10 and t1 < 3 else 0
    for i in range(t1, n):
        prefix = 0
        while r.u(1) == 0:
            prefix += 1
        size = 4 if (prefix == 14 and sl == 0) else 12 if prefix >= 15 else sl
        suffix = r.u(size) if size else 0
        code = (min(15, prefix) << sl) + suffix
        if prefix >= 15 and sl == 0:
            code += 15
        assert prefix <= 15, "escape beyond Baseline"
        if i == t1 and t1 < 3:
            code += 2
        v = (code + 2) >> 1 if code % 2 == 0 else (-code - 1) >> 1
        levels.append(v)
        sl = max(sl, 1)
        if abs(v) > (3 << (sl - 1)) and sl < 6:
            sl += 1
    tz = _read_vlc(r, list(TOTAL_ZEROS[n - 1])) if n < 16 else 0
    runs, left = [], tz
    for _ in range(n - 1):
        run = _read_vlc(r, list(RUN_BEFORE[min(left, 7) - 1])) if left > 0 else 0
        runs.append(run)
        left -= run
    runs.append(left)
    pos = tz + n - 1  # scan position of the highest-frequency level
    for k in range(n):
        lv[pos] = levels[k]
        pos -= 1 + runs[k]
    return lv, n


def _nc(tc, gx, gy):
    """nC of the 4x4 block at (gx, gy) of the picture's 4x4 grid (9.2.1, one slice)."""
    ha, hb = gx > 0, gy > 0
    na = int(tc[gy, gx - 1]) if ha else 0
    nb = int(tc[gy - 1, gx]) if hb else 0
    return (na + nb + 1) >> 1 if ha and hb else na if ha else nb


def mb_residual(src, pred, mx, my, qp):
    """The quantised levels [block raster 0..15][4x4 raster] of macroblock (mx, my) and
    its luma coded_block_pattern."""
    lv = np.zeros((16, 4, 4), np.int64)
    cbp = 0
    for b in range(16):
        y0, x0 = my * 16 + (b // 4) * 4, mx * 16 + (b % 4) * 4
        d = src[y0:y0 + 4, x0:x0 + 4].astype(np.int64) - pred[y0:y0 + 4, x0:x0 + 4].astype(np.int64)
        lv[b] = quant4x4(forward4x4(d), qp)
        if lv[b].any():
            cbp |= 1 << ((b // 8) * 2 + (b % 4) // 2)
    return lv, cbp


def reconstruct(pred, lv, mx, my, qp, out):
    """out[macroblock] = clip(pred + the decoded residual of lv)."""
    for b in range(16):
        y0, x0 = my * 16 + (b // 4) * 4, mx * 16 + (b % 4) * 4
        r = inverse4x4(dequant4x4(lv[b], qp)) if lv[b].any() else 0
        out[y0:y0 + 4, x0:x0 + 4] = np.clip(pred[y0:y0 + 4, x0:x0 + 4].astype(np.int64) + r, 0, 255)


def write_residual(w, lv, cbp, mx, my, tc):
    """coded_block_pattern, mb_qp_delta and residual_luma of a coded P macroblock."""
    w.ue(CBP_INTER.index(cbp))
    if cbp:
        w.se(0)
    for q8 in range(4):
        for q4 in range(4):
            x4, y4 = (q8 % 2) * 2 + q4 % 2, (q8 // 2) * 2 + q4 // 2
            gx, gy = mx * 4 + x4, my * 4 + y4
            n = 0
            if cbp & (1 << q8):
                blk = lv[y4 * 4 + x4].reshape(16)
                n = write_block(w, [int(blk[ZIGZAG[i]]) for i in range(16)], _nc(tc, gx, gy))
            tc[gy, gx] = n


# ---------------------------------------------------------------------------- parser
class BitReader:
    def __init__(self, data):
        self.data, self.pos = data, 0

    def u(self, n):
        v = 0
        for _ in range(n):
            v = (v << 1) | ((self.data[self.pos >> 3] >> (7 - (self.pos & 7))) & 1)
            self.pos += 1
        return v

    def ue(self, ):
        z = 0
        while self.u(1) == 0:
            z += 1
        return (1 << z) - 1 + self.u(z)

    def se(self):
        k = self.ue()
        return (k + 1) // 2 if k & 1 else -(k // 2)

    def align(self):
        while self.pos & 7:
            assert self.u(1) == 0, "non-zero alignment bit"

    def trailing(self):
        assert self.u(1) == 1, "rbsp_stop_one_bit"
        self.align()
        assert self.pos == 8 * len(self.data), "bytes after the trailing bits"


def parse_sps(rbsp):
    r = BitReader(rbsp)
    s = {"profile_idc": r.u(8), "constraints": r.u(8), "level_idc": r.u(8), "sps_id": r.ue(),
         "log2_max_frame_num_minus4": r.ue(), "poc_type": r.ue()}
    s["log2_max_poc_lsb_minus4"] = r.ue()
    s["max_num_ref_frames"] = r.ue()
    s["gaps"] = r.u(1)
    s["width_mbs_minus1"] = r.ue()
    s["height_map_units_minus1"] = r.ue()
    s["frame_mbs_only"] = r.u(1)
    s["direct_8x8_inference"] = r.u(1)
    s["cropping"] = r.u(1)
    s["crop"] = tuple(r.ue() for _ in range(4)) if s["cropping"] else (0, 0, 0, 0)
    s["vui"] = r.u(1)
    r.trailing()
    return s


def parse_pps(rbsp):
    r = BitReader(rbsp)
    p = {"pps_id": r.ue(), "sps_id": r.ue(), "entropy_coding_mode": r.u(1), "bottom_field_poc": r.u(1),
         "num_slice_groups_minus1": r.ue(), "num_ref_idx_l0_minus1": r.ue(), "num_ref_idx_l1_minus1": r.ue(),
         "weighted_pred": r.u(1), "weighted_bipred_idc": r.u(2), "pic_init_qp_minus26": r.se(),
         "pic_init_qs_minus26": r.se(), "chroma_qp_index_offset": r.se(), "deblocking_control": r.u(1),
         "constrained_intra_pred": r.u(1), "redundant_pic_cnt": r.u(1)}
    r.trailing()
    return p


def parse_slice(rbsp, nal_type, sps, pps):
    """Slice header and macroblocks of this subset: returns (header, per-MB list) where
    an I_PCM macroblock is ("pcm", y16x16, u8x8, v8x8) and a P macroblock is
    (mode, [(ref, mvx, mvy) per quadrant TL TR BL BR], skipped)."""
    mbw, mbh = sps["width_mbs_minus1"] + 1, sps["height_map_units_minus1"] + 1
    r = BitReader(rbsp)
    h = {"first_mb": r.ue(), "slice_type": r.ue(), "pps_id": r.ue(), "frame_num": r.u(4)}
    idr = nal_type == 5
    if idr:
        h["idr_pic_id"] = r.ue()
    h["poc_lsb"] = r.u(4)
    nact = pps["num_ref_idx_l0_minus1"] + 1
    if h["slice_type"] % 5 == 0:
        h["override"] = r.u(1)
        if h["override"]:
            nact = r.ue() + 1
        h["ref_pic_list_modification"] = r.u(1)
    if idr:
        h["no_output_of_prior_pics"], h["long_term_reference"] = r.u(1), r.u(1)
    else:
        h["adaptive_ref_pic_marking"] = r.u(1)
    h["slice_qp_delta"] = r.se()
    h["disable_deblocking_filter_idc"] = r.ue()
    if h["disable_deblocking_filter_idc"] != 1:
        h["alpha"], h["beta"] = r.se(), r.se()
    h["num_ref_idx_active"] = nact
    mbs = []
    if h["slice_type"] % 5 == 2:
        for _ in range(mbw * mbh):
            assert r.ue() == 25, "I_PCM"
            r.align()
            yb = np.array([r.u(8) for _ in range(256)], np.uint8).reshape(16, 16)
            ub = np.array([r.u(8) for _ in range(64)], np.uint8).reshape(8, 8)
            vb = np.array([r.u(8) for _ in range(64)], np.uint8).reshape(8, 8)
            mbs.append(("pcm", yb, ub, vb))
        r.trailing()
        return h, mbs
    grid = MotionGrid(mbw, mbh)
    tc = np.zeros((mbh * 4, mbw * 4), np.int64)
    h["levels"] = [None] * (mbw * mbh)  # per macroblock: [16 blocks][4][4] levels, or None
    n, skip = 0, 0
    while n < mbw * mbh:
        if skip == 0:
            skip = r.ue()
            if skip == 0:
                skip = -1  # a coded macroblock follows
        while skip > 0 and n < mbw * mbh:
            mx, my = n % mbw, n // mbw
            sx, sy = skip_mv(grid, mx, my)
            grid.put(mx * 4, my * 4, 4, 4, (0, sx, sy))
            mbs.append((0, [(0, sx, sy)] * 4, True))
            n += 1
            skip -= 1
        if n == mbw * mbh:
            break
        # coded macroblock (the run before it, possibly 0, has been read)
        mx, my = n % mbw, n // mbw
        mode = r.ue()
        assert 0 <= mode <= 3, f"mb_type {mode}"
        if mode == 3:
            assert [r.ue() for _ in range(4)] == [0, 0, 0, 0], "sub_mb_type P_L0_8x8"
        parts = range(len(SHAPES[mode]))
        refs = []
        for _ in parts:
            if nact == 1:
                refs.append(0)
            elif nact == 2:
                refs.append(1 - r.u(1))
            else:
                refs.append(r.ue())
        quads = [None] * 4
        for k in parts:
            px, py = mvp(grid, mx, my, k, mode, refs[k])
            x, y = px + r.se(), py + r.se()
            x4, y4, w4, h4 = SHAPES[mode][k]
            grid.put(mx * 4 + x4, my * 4 + y4, w4, h4, (refs[k], x, y))
            for qq in range(4):
                qx, qy = (qq % 2) * 2, (qq // 2) * 2
                if x4 <= qx < x4 + w4 and y4 <= qy < y4 + h4:
                    quads[qq] = (refs[k], x, y)
        cbp = CBP_INTER[r.ue()]
        assert cbp < 16, "chroma residual"
        if cbp:
            h.setdefault("mb_qp_delta", []).append(r.se())
            lv = np.zeros((16, 4, 4), np.int64)
            for q8 in range(4):
                for q4 in range(4):
                    x4, y4 = (q8 % 2) * 2 + q4 % 2, (q8 // 2) * 2 + q4 // 2
                    gx, gy = mx * 4 + x4, my * 4 + y4
                    if cbp & (1 << q8):
                        scan, tc[gy, gx] = read_block(r, _nc(tc, gx, gy))
                        for i in range(16):
                            lv[y4 * 4 + x4].reshape(16)[ZIGZAG[i]] = scan[i]
            h["levels"][n] = lv
        mbs.append((mode, quads, False))
        n += 1
        skip = 0
    r.trailing()
    return h, mbs


def parse_stream(stream):
    """Every NAL unit of an Annex-B stream of this subset, in order."""
    sps = pps = None
    out = []
    for hdr, rbsp in split_annex_b(stream):
        ref_idc, t = hdr >> 5, hdr & 31
        if t == 7:
            sps = parse_sps(rbsp)
            out.append(("sps", ref_idc, sps))
        elif t == 8:
            pps = parse_pps(rbsp)
            out.append(("pps", ref_idc, pps))
        elif t in (1, 5):
            out.append(("slice", ref_idc, t) + parse_slice(rbsp, t, sps, pps))
        else:
            raise AssertionError(f"NAL type {t}")
    return out
