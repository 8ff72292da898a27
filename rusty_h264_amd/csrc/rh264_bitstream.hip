// rh264_bitstream.hip -- the H.264 bitstream of decided frames (host code, no kernels):
// the SPS and PPS of the reference's CAVLC configuration, an IDR picture of I_PCM
// macroblocks, and P slices carrying the decided motion (P_Skip; P_L0 16x16, 16x8,
// 8x16; P_8x8 with every sub_mb_type P_L0_8x8), either with no residual
// (coded_block_pattern 0) or with the luma residual of the decided prediction:
// forward core transform and inter quantiser as the reference's
// (crates/rusty_h264-common/src/transform.rs:92, 121, 172-174), CAVLC residual blocks
// (crates/rusty_h264-common/src/cavlc.rs; H.264 9.2 and Tables 9-5, 9-7, 9-8, 9-10),
// and the decoder's reconstruction (8.5.12).  Chroma carries no residual.
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


// ---- luma residual (clause 8.5.12 decode side; the encoder side as the reference's) ----

// H.264 Table 9-5: coeff_token code words for 0 <= nC < 2, 2 <= nC < 4 and 4 <= nC < 8,
// indexed [table][TrailingOnes][TotalCoeff] ("" where TrailingOnes > TotalCoeff); nC >= 8
// uses the 6-bit fixed-length code instead.
const char *const kCoeffToken[3][4][17] = {
    {
        {"1", "000101", "00000111", "000000111", "0000000111", "00000000111", "0000000001111", "0000000001011", "0000000001000", "00000000001111", "00000000001011", "000000000001111", "000000000001011", "0000000000001111", "0000000000001011", "0000000000000111", "0000000000000100"},
        {"", "01", "000100", "00000110", "000000110", "0000000110", "00000000110", "0000000001110", "0000000001010", "00000000001110", "00000000001010", "000000000001110", "000000000001010", "000000000000001", "0000000000001110", "0000000000001010", "0000000000000110"},
        {"", "", "001", "0000101", "00000101", "000000101", "0000000101", "00000000101", "0000000001101", "0000000001001", "00000000001101", "00000000001001", "000000000001101", "000000000001001", "0000000000001101", "0000000000001001", "0000000000000101"},
        {"", "", "", "00011", "000011", "0000100", "00000100", "000000100", "0000000100", "00000000100", "0000000001100", "00000000001100", "00000000001000", "000000000001100", "000000000001000", "0000000000001100", "0000000000001000"},
    },
    {
        {"11", "001011", "000111", "0000111", "00000111", "00000100", "000000111", "00000001111", "00000001011", "000000001111", "000000001011", "000000001000", "0000000001111", "0000000001011", "0000000000111", "00000000001001", "00000000000111"},
        {"", "10", "00111", "001010", "000110", "0000110", "00000110", "000000110", "00000001110", "00000001010", "000000001110", "000000001010", "0000000001110", "0000000001010", "00000000001011", "00000000001000", "00000000000110"},
        {"", "", "011", "001001", "000101", "0000101", "00000101", "000000101", "00000001101", "00000001001", "000000001101", "000000001001", "0000000001101", "0000000001001", "0000000000110", "00000000001010", "00000000000101"},
        {"", "", "", "0101", "0100", "00110", "001000", "000100", "0000100", "000000100", "00000001100", "00000001000", "000000001100", "0000000001100", "0000000001000", "0000000000001", "00000000000100"},
    },
    {
        {"1111", "001111", "001011", "001000", "0001111", "0001011", "0001001", "0001000", "00001111", "00001011", "000001111", "000001011", "000001000", "0000001101", "0000001001", "0000000101", "0000000001"},
        {"", "1110", "01111", "01100", "01010", "01000", "001110", "001010", "0001110", "00001110", "00001010", "000001110", "000001010", "000000111", "0000001100", "0000001000", "0000000100"},
        {"", "", "1101", "01110", "01011", "01001", "001101", "001001", "0001101", "0001010", "00001101", "00001001", "000001101", "000001001", "0000001011", "0000000111", "0000000011"},
        {"", "", "", "1100", "1011", "1010", "1001", "1000", "01101", "001100", "0001100", "00001100", "00001000", "000001100", "0000001010", "0000000110", "0000000010"},
    },
};
// Tables 9-7 and 9-8: total_zeros of a 4x4 block, indexed [TotalCoeff - 1][total_zeros].
const char *const kTotalZeros[15][16] = {
    {"1", "011", "010", "0011", "0010", "00011", "00010", "000011", "000010", "0000011", "0000010", "00000011", "00000010", "000000011", "000000010", "000000001"},
    {"111", "110", "101", "100", "011", "0101", "0100", "0011", "0010", "00011", "00010", "000011", "000010", "000001", "000000"},
    {"0101", "111", "110", "101", "0100", "0011", "100", "011", "0010", "00011", "00010", "000001", "00001", "000000"},
    {"00011", "111", "0101", "0100", "110", "101", "100", "0011", "011", "0010", "00010", "00001", "00000"},
    {"0101", "0100", "0011", "111", "110", "101", "100", "011", "0010", "00001", "0001", "00000"},
    {"000001", "00001", "111", "110", "101", "100", "011", "010", "0001", "001", "000000"},
    {"000001", "00001", "101", "100", "011", "11", "010", "0001", "001", "000000"},
    {"000001", "0001", "00001", "011", "11", "10", "010", "001", "000000"},
    {"000001", "000000", "0001", "11", "10", "001", "01", "00001"},
    {"00001", "00000", "001", "11", "10", "01", "0001"},
    {"0000", "0001", "001", "010", "1", "011"},
    {"0000", "0001", "01", "1", "001"},
    {"000", "001", "1", "01"},
    {"00", "01", "1"},
    {"0", "1"},
};
// Table 9-10: run_before, indexed [min(zerosLeft, 7) - 1][run_before].
const char *const kRunBefore[7][15] = {
    {"1", "0"},
    {"1", "01", "00"},
    {"11", "10", "01", "00"},
    {"11", "10", "01", "001", "000"},
    {"11", "10", "011", "010", "001", "000"},
    {"11", "000", "001", "011", "010", "101", "100"},
    {"111", "110", "101", "100", "011", "010", "001", "0001", "00001", "000001", "0000001", "00000001", "000000001", "0000000001", "00000000001"},
};

// 4x4 zig-zag scan (Table 8-13, frame macroblocks): scan position -> raster index
const int kZigzag[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};

// quantisation multipliers MF and dequantisation scales v (8.5.12.1, flat weights) by
// qp % 6 and the position class: 0 = both coordinates even, 1 = both odd, 2 = mixed
const int kQuantMF[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490}, {10082, 4194, 6554},
                            {9362, 3647, 5825},  {8192, 3355, 5243},  {7282, 2893, 4559}};
const int kDequantV[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
inline int pos_class(int i) { return ((i & 1) == 0 && ((i >> 2) & 1) == 0) ? 0 : ((i & 1) && ((i >> 2) & 1)) ? 1 : 2; }

// Forward core transform W = Cf X Cf^T, Cf = [1 1 1 1; 2 1 -1 -2; 1 -1 -1 1; 1 -2 2 -1]
// (the reference's forward_core, crates/rusty_h264-common/src/transform.rs:92), raster order.
void forward4x4(const int x[16], int w[16])
{
    int t[16];
    for (int r = 0; r < 4; r++) { // rows: t = X Cf^T
        const int a = x[4 * r], b = x[4 * r + 1], c = x[4 * r + 2], d = x[4 * r + 3];
        const int s03 = a + d, d03 = a - d, s12 = b + c, d12 = b - c;
        t[4 * r] = s03 + s12;
        t[4 * r + 1] = 2 * d03 + d12;
        t[4 * r + 2] = s03 - s12;
        t[4 * r + 3] = d03 - 2 * d12;
    }
    for (int c = 0; c < 4; c++) { // columns: W = Cf t
        const int a = t[c], b = t[4 + c], cc = t[8 + c], d = t[12 + c];
        const int s03 = a + d, d03 = a - d, s12 = b + cc, d12 = b - cc;
        w[c] = s03 + s12;
        w[4 + c] = 2 * d03 + d12;
        w[8 + c] = s03 - s12;
        w[12 + c] = d03 - 2 * d12;
    }
}

// Inter quantisation: |level| = (|W| * MF + 2^qbits / 6) >> qbits, qbits = 15 + qp / 6, the
// sign of W (transform.rs:121, 172-174: the inter dead zone).  Returns the nonzero count.
int quant4x4(const int w[16], int qp, int z[16])
{
    const int qbits = 15 + qp / 6;
    const int64_t f = ((int64_t)1 << qbits) / 6;
    int nz = 0;
    for (int i = 0; i < 16; i++) {
        const int64_t a = w[i] < 0 ? -(int64_t)w[i] : w[i];
        const int l = (int)((a * kQuantMF[qp % 6][pos_class(i)] + f) >> qbits);
        z[i] = w[i] < 0 ? -l : l;
        nz += l != 0;
    }
    return nz;
}

// Scaling (8.5.12.1 with flat weights: c * v << qp / 6) and the inverse transform
// (8.5.12.2), then (x + 32) >> 6: the residual of one 4x4 block, raster order.
void dequant_inverse4x4(const int z[16], int qp, int r[16])
{
    int d[16], t[16];
    for (int i = 0; i < 16; i++) d[i] = (z[i] * kDequantV[qp % 6][pos_class(i)]) << (qp / 6);
    for (int i = 0; i < 4; i++) { // rows
        const int e0 = d[4 * i] + d[4 * i + 2], e1 = d[4 * i] - d[4 * i + 2];
        const int e2 = (d[4 * i + 1] >> 1) - d[4 * i + 3], e3 = d[4 * i + 1] + (d[4 * i + 3] >> 1);
        t[4 * i] = e0 + e3;
        t[4 * i + 1] = e1 + e2;
        t[4 * i + 2] = e1 - e2;
        t[4 * i + 3] = e0 - e3;
    }
    for (int j = 0; j < 4; j++) { // columns
        const int g0 = t[j] + t[8 + j], g1 = t[j] - t[8 + j];
        const int g2 = (t[4 + j] >> 1) - t[12 + j], g3 = t[4 + j] + (t[12 + j] >> 1);
        r[j] = (g0 + g3 + 32) >> 6;
        r[4 + j] = (g1 + g2 + 32) >> 6;
        r[8 + j] = (g1 - g2 + 32) >> 6;
        r[12 + j] = (g0 - g3 + 32) >> 6;
    }
}

void put_code(Bits &w, const char *code)
{
    for (const char *c = code; *c; c++) w.bit(*c - '0');
}

// residual_block_cavlc (7.3.5.3.2, 9.2) of one 4x4 luma block whose 16 levels are in
// scan order; nC is the coeff_token context (9.2.1).  Returns TotalCoeff.  Fails
// (returns -1) on a level outside the Baseline escape range.
int write_block(Bits &w, const int lv[16], int nC)
{
    int nz[16], n = 0; // scan positions of the nonzero levels, ascending
    for (int i = 0; i < 16; i++)
        if (lv[i]) nz[n++] = i;
    int t1 = 0; // trailing ones: up to three +-1 at the high-frequency end
    for (int k = n - 1; k >= 0 && t1 < 3 && (lv[nz[k]] == 1 || lv[nz[k]] == -1); k--) t1++;
    if (nC >= 8) {
        w.bits(n == 0 ? 3u : (uint32_t)(((n - 1) << 2) | t1), 6);
    } else {
        const int tab = nC < 2 ? 0 : nC < 4 ? 1 : 2;
        put_code(w, kCoeffToken[tab][t1][n]);
    }
    if (n == 0) return 0;
    for (int k = n - 1; k >= n - t1; k--) w.bit(lv[nz[k]] < 0); // trailing_ones_sign_flag
    int sl = (n > 10 && t1 < 3) ? 1 : 0; // suffixLength
    for (int k = n - 1 - t1; k >= 0; k--) {
        const int v = lv[nz[k]];
        int code = v > 0 ? 2 * v - 2 : -2 * v - 1; // levelCode
        if (k == n - 1 - t1 && t1 < 3) code -= 2;   // the first level after fewer than 3 T1s
        int prefix, ssize, suffix;
        if (sl == 0) {
            if (code < 14) {
                prefix = code, ssize = 0, suffix = 0;
            } else if (code < 30) {
                prefix = 14, ssize = 4, suffix = code - 14;
            } else {
                prefix = 15, ssize = 12, suffix = code - 30;
            }
        } else if (code < (15 << sl)) {
            prefix = code >> sl, ssize = sl, suffix = code & ((1 << sl) - 1);
        } else {
            prefix = 15, ssize = 12, suffix = code - (15 << sl);
        }
        if (suffix >= 4096) return -1;
        w.bits(0, prefix);
        w.bit(1);
        if (ssize) w.bits((uint32_t)suffix, ssize);
        if (sl == 0) sl = 1;
        if ((v < 0 ? -v : v) > (3 << (sl - 1)) && sl < 6) sl++;
    }
    const int tz = nz[n - 1] + 1 - n; // zeros before the last nonzero level
    if (n < 16) put_code(w, kTotalZeros[n - 1][tz]);
    int left = tz;
    for (int k = n - 1; k > 0 && left > 0; k--) {
        const int run = nz[k] - nz[k - 1] - 1;
        put_code(w, kRunBefore[(left < 7 ? left : 7) - 1][run]);
        left -= run;
    }
    return n;
}

// Table 9-4, inter column: coded_block_pattern -> codeNum (chroma_format_idc 1)
int cbp_inter_code(int cbp)
{
    static const uint8_t code_to_cbp[48] = {0,  16, 1,  2,  4,  8,  32, 3,  5,  10, 12, 15, 47, 7,  11, 13,
                                            14, 6,  9,  31, 35, 37, 42, 44, 33, 34, 36, 40, 39, 43, 45, 46,
                                            17, 18, 20, 24, 19, 21, 26, 28, 23, 27, 29, 30, 22, 25, 38, 41};
    for (int c = 0; c < 48; c++)
        if (code_to_cbp[c] == cbp) return c;
    return -1;
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

} // extern "C"

namespace {

// A P picture from the decision records; with src/pred (and recon) the luma residual
// src - pred of every macroblock is transformed, quantised at s->qp and coded, and
// recon receives pred + the decoded residual; without them no residual is coded.
int write_p_slice(const rh264_seq *s, int frame_num, int poc_lsb, int nrefs, const int32_t *rec, const uint8_t *src,
                  const uint8_t *pred, uint8_t *recon, uint8_t *out, size_t cap, size_t *len)
{
    if (!seq_ok(s) || !rec || nrefs < 1 || nrefs > s->num_ref_frames) return fail("write_p_slice: bad arguments");
    const bool res = src != nullptr;
    if (res && (!pred || !recon)) return fail("write_p_slice_residual: null picture");
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
    if (res) {
        // disable_deblocking_filter_idc 1: the reconstruction returned (and predicted
        // from) is then exactly what a decoder outputs
        w.ue(1);
    } else {
        w.ue(0);                         // the reference's header: filter on, offsets 0
        w.se(0);
        w.se(0);
    }
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
    const int W = mbw * 16;
    // TotalCoeff of every 4x4 luma block of the picture (0 for skipped and uncoded ones)
    std::vector<uint8_t> tc(res ? (size_t)mbw * mbh * 16 : 0, 0);
    for (int my = 0; my < mbh; my++)
        for (int mx = 0; mx < mbw; mx++) {
            const int32_t *r = rec + 14 * ((size_t)my * mbw + mx);
            const int mode = r[0];
            // the macroblock's quantised residual, blocks in raster order of the 4x4 grid
            int lv[16][16];
            int cbp = 0;
            if (res) {
                for (int b = 0; b < 16; b++) {
                    const int bx = mx * 16 + (b & 3) * 4, by = my * 16 + (b >> 2) * 4;
                    int d[16], wc[16];
                    for (int i = 0; i < 16; i++) {
                        const size_t o = (size_t)(by + (i >> 2)) * W + bx + (i & 3);
                        d[i] = (int)src[o] - (int)pred[o];
                    }
                    forward4x4(d, wc);
                    if (quant4x4(wc, s->qp, lv[b])) cbp |= 1 << (((b >> 3) << 1) | ((b & 3) >> 1));
                }
            }
            if (mode == 0 && r[2] == 0 && cbp == 0) {
                int sx, sy;
                skip_mv(g, mx, my, &sx, &sy);
                if (sx == r[3] && sy == r[4]) { // P_Skip: no residual, the same prediction
                    skip_run++;
                    for (int q = 0; q < 4; q++) g.set(mx, my, q, Mv{0, sx, sy});
                    if (res)
                        for (int y = 0; y < 16; y++)
                            memcpy(recon + (size_t)(my * 16 + y) * W + mx * 16, pred + (size_t)(my * 16 + y) * W + mx * 16, 16);
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
            if (!res) {
                w.ue(0); // coded_block_pattern: inter codeNum 0 = no residual
                continue;
            }
            w.ue((uint32_t)cbp_inter_code(cbp)); // luma bits only: chroma carries no residual
            if (cbp) w.se(0);                     // mb_qp_delta
            // residual_luma: 8x8 quadrants in order, their 4x4 blocks in order (6.4.3)
            for (int q8 = 0; q8 < 4; q8++)
                for (int q4 = 0; q4 < 4; q4++) {
                    const int x4 = (q8 & 1) * 2 + (q4 & 1), y4 = (q8 >> 1) * 2 + (q4 >> 1), b = y4 * 4 + x4;
                    const int gx = mx * 4 + x4, gy = my * 4 + y4; // 4x4 grid of the picture
                    int n = 0;
                    if (cbp & (1 << q8)) {
                        // nC from the left (A) and upper (B) blocks, 9.2.1
                        const bool ha = gx > 0, hb = gy > 0;
                        const int na = ha ? tc[(size_t)gy * mbw * 4 + gx - 1] : 0;
                        const int nb = hb ? tc[(size_t)(gy - 1) * mbw * 4 + gx] : 0;
                        const int nC = ha && hb ? (na + nb + 1) >> 1 : ha ? na : nb;
                        int scan[16];
                        for (int i = 0; i < 16; i++) scan[i] = lv[b][kZigzag[i]];
                        n = write_block(w, scan, nC);
                        if (n < 0) return fail("write_p_slice_residual: macroblock %d: level out of range", my * mbw + mx);
                    }
                    tc[(size_t)gy * mbw * 4 + gx] = (uint8_t)n;
                    // reconstruction: the prediction plus the decoded residual
                    int rr[16] = {0};
                    if (n) dequant_inverse4x4(lv[b], s->qp, rr);
                    for (int i = 0; i < 16; i++) {
                        const size_t o = (size_t)(my * 16 + y4 * 4 + (i >> 2)) * W + mx * 16 + x4 * 4 + (i & 3);
                        const int v = (int)pred[o] + rr[i];
                        recon[o] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
                    }
                }
        }
    if (skip_run) w.ue(skip_run);
    w.trailing();
    std::vector<uint8_t> v;
    nal(v, 3, 1, w.bytes());
    return emit(v, out, cap, len);
}

} // namespace

extern "C" {

int rh264_write_p_slice(const rh264_seq *s, int frame_num, int poc_lsb, int nrefs, const int32_t *rec, uint8_t *out,
                        size_t cap, size_t *len)
{
    return write_p_slice(s, frame_num, poc_lsb, nrefs, rec, nullptr, nullptr, nullptr, out, cap, len);
}

int rh264_write_p_slice_residual(const rh264_seq *s, int frame_num, int poc_lsb, int nrefs, const int32_t *rec,
                                 const uint8_t *src, const uint8_t *pred, uint8_t *recon, uint8_t *out, size_t cap,
                                 size_t *len)
{
    if (!src) return fail("write_p_slice_residual: null source");
    return write_p_slice(s, frame_num, poc_lsb, nrefs, rec, src, pred, recon, out, cap, len);
}

} // extern "C"
