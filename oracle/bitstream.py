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


def p_slice(width, height, num_ref_frames, frame_num, poc_lsb, nrefs, rec):
    """The P picture of decision records rec [mb][14] (mode, cost, TL/TR/BL/BR ref mvx mvy)."""
    mbw, mbh = (width + 15) // 16, (height + 15) // 16
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
            if mode == 0 and q(0)[0] == 0 and tuple(q(0)[1:]) == skip_mv(grid, mx, my):
                run += 1
                grid.put(mx * 4, my * 4, 4, 4, (0, q(0)[1], q(0)[2]))
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
            w.ue(0)
    if run:
        w.ue(run)
    w.trailing()
    return nal(3, 1, w.bytes())


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
        assert r.ue() == 0, "coded_block_pattern 0"
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
