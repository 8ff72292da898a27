// rh264_bitstream.hip -- the H.264 bitstream of decided frames (host code, no kernels):
// the SPS and PPS of the reference's CAVLC configuration, an IDR picture of I_PCM
// macroblocks, and P slices carrying the decided motion (P_Skip; P_L0 16x16, 16x8,
// 8x16; P_8x8 with every sub_mb_type P_L0_8x8) with no residual (coded_block_pattern 0).
//
// Syntax follows ITU-T H.264 clause 7.3 as the reference writes it:
//   SPS / PPS fields      crates/rusty_h264-encoder/src/params.rs:31-101, 122-172
//                         (Constrained Baseline: profile 66, constraint_set1, CAVLC)
//   slice headers         crates/rusty_h264-encoder/src/slice.rs:49-107
//   P macroblock layer    crates/rusty_h264-encoder/src/mb16.rs:5380-5776 (mb_skip_run
//                         before each coded macroblock and at the end), 4416-4435
//                         (mb_type, four sub_mb_type for P_8x8, every ref_idx, every
//                         mvd), 5957-5964 (ref_idx: one inverted bit with two active
//                         references, else ue), crates/rusty_h264-common/src/cavlc.rs:
//                         136-138 (coded_block_pattern: inter codeNum 0 is cbp 0)
//   NAL units             crates/rusty_h264-common/src/nal.rs:78-117 (4-byte start
//                         code, header byte, emulation prevention after two zeros)
// Motion vector prediction is the clause's own: 8.4.1.3 (16x8 and 8x16 directional
// predictors, one matching reference, the median; C replaced by D where C is not
// available; B and C taken from A when only A is) and 8.4.1.1 (P_Skip).

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "rh264_internal.h"

namespace {

class Bits {
  public:
    void bit(int b)
    {
        cur_ = (uint8_t)((cur_ << 1) | (b & 1));
        if (++n_ == 8) {
            buf_.push_back(cur_);
            cur_ = 0;
            n_ = 0;
        }
    }
    void bits(uint32_t v, int n)
    {
        for (int i = n - 1; i >= 0; i--) bit((int)((v >> i) & 1u));
    }
    void ue(uint32_t v)
    {
        const uint64_t c = (uint64_t)v + 1;
        int len = 0;
        while ((c >> (len + 1)) != 0) len++;
        bits(0, len);
        for (int i = len; i >= 0; i--) bit((int)((c >> i) & 1u));
    }
    void se(int32_t v) { ue(v > 0 ? 2u * (uint32_t)v - 1u : 2u * (uint32_t)(-(int64_t)v)); }
    bool aligned() const { return n_ == 0; }
    void align_zero()
    {
        while (n_) bit(0);
    }
    void trailing() // rbsp_trailing_bits
    {
        bit(1);
        align_zero();
    }
    const std::vector<uint8_t> &bytes() const { return buf_; }

  private:
    std::vector<uint8_t> buf_;
    uint8_t cur_ = 0;
    int n_ = 0;
};

// Annex-B NAL unit: start code 00 00 00 01, header forbidden_zero_bit | nal_ref_idc |
// nal_unit_type, the RBSP with an emulation-prevention byte before any byte <= 3
// that follows two zero bytes.
void nal(std::vector<uint8_t> &out, int ref_idc, int type, const std::vector<uint8_t> &rbsp)
{
    out.insert(out.end(), {0, 0, 0, 1});
    out.push_back((uint8_t)((ref_idc << 5) | type));
    int zeros = 0;
    for (uint8_t b : rbsp) {
        if (zeros >= 2 && b <= 3) {
            out.push_back(3);
            zeros = 0;
        }
        out.push_back(b);
        zeros = b == 0 ? zeros + 1 : 0;
    }
}

int emit(const std::vector<uint8_t> &v, uint8_t *out, size_t cap, size_t *len)
{
    if (len) *len = v.size();
    if (!out || v.size() > cap) return fail("bitstream: %zu bytes do not fit the %zu-byte buffer", v.size(), cap);
    memcpy(out, v.data(), v.size());
    return 0;
}

bool seq_ok(const rh264_seq *s)
{
    return s && s->width > 0 && s->height > 0 && s->width <= 16384 && s->height <= 16384 && s->qp >= 0 &&
           s->qp <= 51 && s->num_ref_frames >= 1 && s->num_ref_frames <= 16 && s->level_idc > 0 &&
           s->level_idc < 256;
}

struct Mv {
    int ref; // -1: not available
    int x, y;
};

int median3(int a, int b, int c) { return a + b + c - (a < b ? (a < c ? a : c) : (b < c ? b : c)) - (a > b ? (a > c ? a : c) : (b > c ? b : c)); }

// The decoded motion of a picture on the 8x8 grid (two entries per macroblock side),
// in macroblock order.
class Grid {
  public:
    Grid(int mbw, int mbh) : mbw_(mbw), mbh_(mbh), g_((size_t)mbw * mbh * 4, Mv{-1, 0, 0}) {}
    // neighbour at 4x4 position (x4, y4) relative to macroblock (mx, my); `done` is
    // the set of the current macroblock's quadrants already decoded
    Mv at(int mx, int my, int x4, int y4, unsigned done) const
    {
        int nx = mx, ny = my;
        if (x4 < 0) {
            nx--;
            x4 += 4;
        } else if (x4 >= 4) {
            nx++;
            x4 -= 4;
        }
        if (y4 < 0) {
            ny--;
            y4 += 4;
        }
        if (nx < 0 || nx >= mbw_ || ny < 0 || ny >= mbh_) return Mv{-1, 0, 0};
        const int q = (y4 >> 1) * 2 + (x4 >> 1);
        if (nx == mx && ny == my) {
            if (!(done & (1u << q))) return Mv{-1, 0, 0};
        } else if (ny > my || (ny == my && nx > mx)) {
            return Mv{-1, 0, 0}; // not decoded yet
        }
        return g_[((size_t)ny * mbw_ + nx) * 4 + q];
    }
    void set(int mx, int my, int q, Mv v) { g_[((size_t)my * mbw_ + mx) * 4 + q] = v; }

  private:
    int mbw_, mbh_;
    std::vector<Mv> g_;
};

// mvpLX of a partition whose top-left 4x4 block is (x4, y4) and whose width is w4
// 4x4 blocks (8.4.1.3); shape 0 16x16, 1 16x8, 2 8x16, 3 8x8; part = partition index
void predict(const Grid &g, int mx, int my, int x4, int y4, int w4, unsigned done, int shape, int part, int ref,
             int *px, int *py)
{
    Mv A = g.at(mx, my, x4 - 1, y4, done), B = g.at(mx, my, x4, y4 - 1, done);
    Mv C = g.at(mx, my, x4 + w4, y4 - 1, done);
    if (C.ref < 0) C = g.at(mx, my, x4 - 1, y4 - 1, done); // D
    if (shape == 1) {
        if (part == 0 && B.ref == ref) {
            *px = B.x;
            *py = B.y;
            return;
        }
        if (part == 1 && A.ref == ref) {
            *px = A.x;
            *py = A.y;
            return;
        }
    } else if (shape == 2) {
        if (part == 0 && A.ref == ref) {
            *px = A.x;
            *py = A.y;
            return;
        }
        if (part == 1 && C.ref == ref) {
            *px = C.x;
            *py = C.y;
            return;
        }
    }
    if (B.ref < 0 && C.ref < 0 && A.ref >= 0) B = C = A;
    const int n = (A.ref == ref) + (B.ref == ref) + (C.ref == ref);
    if (n == 1) {
        const Mv &m = A.ref == ref ? A : B.ref == ref ? B : C;
        *px = m.x;
        *py = m.y;
        return;
    }
    *px = median3(A.x, B.x, C.x);
    *py = median3(A.y, B.y, C.y);
}

// P_Skip motion vector (8.4.1.1)
void skip_mv(const Grid &g, int mx, int my, int *px, int *py)
{
    const Mv A = g.at(mx, my, -1, 0, 0), B = g.at(mx, my, 0, -1, 0);
    const bool aout = mx == 0, bout = my == 0;
    if (aout || bout || (A.ref == 0 && A.x == 0 && A.y == 0) || (B.ref == 0 && B.x == 0 && B.y == 0)) {
        *px = *py = 0;
        return;
    }
    predict(g, mx, my, 0, 0, 4, 0, 0, 0, 0, px, py);
}

} // namespace

extern "C" {

int rh264_write_headers(const rh264_seq *s, uint8_t *out, size_t cap, size_t *len)
{
    if (!seq_ok(s)) return fail("write_headers: bad sequence description");
    const int mbw = (s->width + 15) / 16, mbh = (s->height + 15) / 16;
    const int crop_r = (mbw * 16 - s->width) / 2, crop_b = (mbh * 16 - s->height) / 2;
    std::vector<uint8_t> v;
    {
        Bits w; // seq_parameter_set_rbsp
        w.bits(66, 8);       // profile_idc: Baseline (Constrained: constraint_set1_flag)
        w.bits(1u << 6, 8);  // constraint_set0..5 flags, reserved_zero_2bits
        w.bits((uint32_t)s->level_idc, 8);
        w.ue(0);             // seq_parameter_set_id
        w.ue(0);             // log2_max_frame_num_minus4
        w.ue(0);             // pic_order_cnt_type
        w.ue(0);             // log2_max_pic_order_cnt_lsb_minus4
        w.ue((uint32_t)s->num_ref_frames);
        w.bit(0);            // gaps_in_frame_num_value_allowed_flag
        w.ue((uint32_t)(mbw - 1));
        w.ue((uint32_t)(mbh - 1));
        w.bit(1);            // frame_mbs_only_flag
        w.bit(1);            // direct_8x8_inference_flag
        const bool crop = crop_r || crop_b;
        w.bit(crop);
        if (crop) {
            w.ue(0);
            w.ue((uint32_t)crop_r);
            w.ue(0);
            w.ue((uint32_t)crop_b);
        }
        w.bit(0);            // vui_parameters_present_flag
        w.trailing();
        nal(v, 3, 7, w.bytes());
    }
    {
        Bits w; // pic_parameter_set_rbsp
        w.ue(0);             // pic_parameter_set_id
        w.ue(0);             // seq_parameter_set_id
        w.bit(0);            // entropy_coding_mode_flag: CAVLC
        w.bit(0);            // bottom_field_pic_order_in_frame_present_flag
        w.ue(0);             // num_slice_groups_minus1
        w.ue((uint32_t)(s->num_ref_frames - 1)); // num_ref_idx_l0_default_active_minus1
        w.ue(0);             // num_ref_idx_l1_default_active_minus1
        w.bit(0);            // weighted_pred_flag
        w.bits(0, 2);        // weighted_bipred_idc
        w.se(s->qp - 26);    // pic_init_qp_minus26
        w.se(0);             // pic_init_qs_minus26
        w.se(0);             // chroma_qp_index_offset
        w.bit(1);            // deblocking_filter_control_present_flag
        w.bit(0);            // constrained_intra_pred_flag
        w.bit(0);            // redundant_pic_cnt_present_flag
        w.trailing();
        nal(v, 3, 8, w.bytes());
    }
    return emit(v, out, cap, len);
}

int rh264_write_idr_pcm(const rh264_seq *s, const uint8_t *y, const uint8_t *u, const uint8_t *vv, uint8_t *out,
                        size_t cap, size_t *len)
{
    if (!seq_ok(s) || !y || !u || !vv) return fail("write_idr_pcm: bad arguments");
    const int mbw = (s->width + 15) / 16, mbh = (s->height + 15) / 16, W = mbw * 16, CWD = mbw * 8;
    Bits w; // IDR slice header (slice.rs:87-107), slice_qp_delta 0
    w.ue(0);          // first_mb_in_slice
    w.ue(7);          // slice_type: I, all slices
    w.ue(0);          // pic_parameter_set_id
    w.bits(0, 4);     // frame_num
    w.ue(0);          // idr_pic_id
    w.bits(0, 4);     // pic_order_cnt_lsb
    w.bit(0);         // no_output_of_prior_pics_flag
    w.bit(0);         // long_term_reference_flag
    w.se(0);          // slice_qp_delta
    w.ue(0);          // disable_deblocking_filter_idc
    w.se(0);          // slice_alpha_c0_offset_div2
    w.se(0);          // slice_beta_offset_div2
    for (int my = 0; my < mbh; my++)
        for (int mx = 0; mx < mbw; mx++) {
            w.ue(25); // mb_type I_PCM
            w.align_zero(); // pcm_alignment_zero_bit
            for (int r = 0; r < 16; r++)
                for (int c = 0; c < 16; c++) w.bits(y[(size_t)(my * 16 + r) * W + mx * 16 + c], 8);
            for (const uint8_t *pl : {u, vv})
                for (int r = 0; r < 8; r++)
                    for (int c = 0; c < 8; c++) w.bits(pl[(size_t)(my * 8 + r) * CWD + mx * 8 + c], 8);
        }
    w.trailing();
    std::vector<uint8_t> v;
    nal(v, 3, 5, w.bytes());
    return emit(v, out, cap, len);
}

int rh264_write_p_slice(const rh264_seq *s, int frame_num, int poc_lsb, int nrefs, const int32_t *rec, uint8_t *out,
                        size_t cap, size_t *len)
{
    if (!seq_ok(s) || !rec || nrefs < 1 || nrefs > s->num_ref_frames) return fail("write_p_slice: bad arguments");
    const int mbw = (s->width + 15) / 16, mbh = (s->height + 15) / 16;
    // validate the records: mode 0..3, references in range, quadrants consistent with
    // the mode's partitions
    for (int m = 0; m < mbw * mbh; m++) {
        const int32_t *r = rec + 14 * (size_t)m;
        const int mode = r[0];
        if (mode < 0 || mode > 3) return fail("write_p_slice: macroblock %d: mode %d", m, mode);
        for (int q = 0; q < 4; q++)
            if (r[2 + 3 * q] < 0 || r[2 + 3 * q] >= nrefs)
                return fail("write_p_slice: macroblock %d: reference %d of %d", m, r[2 + 3 * q], nrefs);
        static const int same[4][4] = {{0, 0, 0, 0}, {0, 0, 2, 2}, {0, 1, 0, 1}, {0, 1, 2, 3}};
        for (int q = 0; q < 4; q++)
            for (int k = 0; k < 3; k++)
                if (r[2 + 3 * q + k] != r[2 + 3 * same[mode][q] + k])
                    return fail("write_p_slice: macroblock %d: quadrant %d differs inside its partition", m, q);
    }
    Bits w; // P slice header (slice.rs:51-85), slice_qp_delta 0
    w.ue(0);                             // first_mb_in_slice
    w.ue(5);                             // slice_type: P, all slices
    w.ue(0);                             // pic_parameter_set_id
    w.bits((uint32_t)(frame_num & 15), 4);
    w.bits((uint32_t)(poc_lsb & 15), 4);
    if (s->num_ref_frames > 1) {
        w.bit(1);                        // num_ref_idx_active_override_flag
        w.ue((uint32_t)(nrefs - 1));
    } else {
        w.bit(0);
    }
    w.bit(0);                            // ref_pic_list_modification_flag_l0
    w.bit(0);                            // adaptive_ref_pic_marking_mode_flag
    w.se(0);                             // slice_qp_delta
    w.ue(0);                             // disable_deblocking_filter_idc
    w.se(0);
    w.se(0);
    const int nactive = s->num_ref_frames > 1 ? nrefs : 1;
    Grid g(mbw, mbh);
    uint32_t skip_run = 0;
    // partitions of each shape: top-left 4x4 block, width in 4x4 blocks, quadrants
    struct Part {
        int x4, y4, w4, quads;
    };
    static const Part parts[4][4] = {{{0, 0, 4, 0xF}},
                                     {{0, 0, 4, 0x3}, {0, 2, 4, 0xC}},
                                     {{0, 0, 2, 0x5}, {2, 0, 2, 0xA}},
                                     {{0, 0, 2, 0x1}, {2, 0, 2, 0x2}, {0, 2, 2, 0x4}, {2, 2, 2, 0x8}}};
    static const int nparts[4] = {1, 2, 2, 4};
    for (int my = 0; my < mbh; my++)
        for (int mx = 0; mx < mbw; mx++) {
            const int32_t *r = rec + 14 * ((size_t)my * mbw + mx);
            const int mode = r[0];
            if (mode == 0 && r[2] == 0) {
                int sx, sy;
                skip_mv(g, mx, my, &sx, &sy);
                if (sx == r[3] && sy == r[4]) { // P_Skip: no residual, the same prediction
                    skip_run++;
                    for (int q = 0; q < 4; q++) g.set(mx, my, q, Mv{0, sx, sy});
                    continue;
                }
            }
            w.ue(skip_run);
            skip_run = 0;
            w.ue((uint32_t)mode); // P_L0_16x16, P_L0_L0_16x8, P_L0_L0_8x16, P_8x8
            if (mode == 3)
                for (int k = 0; k < 4; k++) w.ue(0); // sub_mb_type P_L0_8x8
            if (nactive > 1)
                for (int k = 0; k < nparts[mode]; k++) {
                    const int q = __builtin_ctz((unsigned)parts[mode][k].quads);
                    const int ref = r[2 + 3 * q];
                    if (nactive == 2)
                        w.bit(ref == 0); // te(v) with one bit: !ref_idx
                    else
                        w.ue((uint32_t)ref);
                }
            unsigned done = 0;
            for (int k = 0; k < nparts[mode]; k++) {
                const Part &pt = parts[mode][k];
                const int q = __builtin_ctz((unsigned)pt.quads);
                const int ref = r[2 + 3 * q], mvx = r[3 + 3 * q], mvy = r[4 + 3 * q];
                int px, py;
                predict(g, mx, my, pt.x4, pt.y4, pt.w4, done, mode, k, ref, &px, &py);
                w.se(mvx - px);
                w.se(mvy - py);
                for (int qq = 0; qq < 4; qq++)
                    if (pt.quads & (1 << qq)) g.set(mx, my, qq, Mv{ref, mvx, mvy});
                done |= (unsigned)pt.quads;
            }
            w.ue(0); // coded_block_pattern: inter codeNum 0 = no residual
        }
    if (skip_run) w.ue(skip_run);
    w.trailing();
    std::vector<uint8_t> v;
    nal(v, 3, 1, w.bytes());
    return emit(v, out, cap, len);
}

} // extern "C"
