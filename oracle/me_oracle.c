/*
 * me_oracle.c -- CPU restatement of rusty_h264's encoder motion-search path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the timed CPU baseline);
 * the product library never links it.  See me_oracle.h for the contract.
 * Reference citations are file:line in the reference repository.
 */
#include "me_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int clip_u8(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
static inline int iabs(int v) { return v < 0 ? -v : v; }
/* Rust `>>` on i32 is arithmetic (floor); spell it out so no compiler choice leaks in. */
static inline int asr2(int v) { return v >= 0 ? (v >> 2) : -((-v + 3) >> 2); }

/* ---------------------------------------------------------------------------
 * Block metrics
 * ------------------------------------------------------------------------- */

#if defined(__AVX2__)
#include <immintrin.h>
/* The timed CPU baseline (the -march=x86-64-v3 build): the reference's SSE2 sad
 * (satd_sad.rs x86::sad) shape, one psadbw per 16- or 8-byte row; SAD is exact, so
 * the value equals the scalar loop's. */
static int32_t sad_psadbw(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h)
{
    __m128i acc = _mm_setzero_si128();
    for (int i = 0; i < h; i++) {
        const __m128i x = w == 16 ? _mm_loadu_si128((const __m128i *)(a + (size_t)i * sa))
                                  : _mm_loadl_epi64((const __m128i *)(a + (size_t)i * sa));
        const __m128i y = w == 16 ? _mm_loadu_si128((const __m128i *)(b + (size_t)i * sb))
                                  : _mm_loadl_epi64((const __m128i *)(b + (size_t)i * sb));
        acc = _mm_add_epi64(acc, _mm_sad_epu8(x, y));
    }
    return (int32_t)(_mm_cvtsi128_si64(acc) + _mm_extract_epi64(acc, 1));
}
#endif

/* sad_scalar, crates/rusty_h264-accel/src/satd_sad.rs:74-82 */
int32_t orc_sad(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h)
{
#if defined(__AVX2__)
    if (w == 16 || w == 8) return sad_psadbw(a, sa, b, sb, w, h);
#endif
    int32_t s = 0;
    for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++)
            s += iabs((int)a[i * sa + j] - (int)b[i * sb + j]);
    return s;
}

/* satd4x4_scalar, satd_sad.rs:39-58: (sum |H d| + 1) >> 1, row butterfly
 * (s0+s1, s2+s3, s2-s3, s0-s1) with s0=r0+r2, s1=r1+r3, s2=r0-r2, s3=r1-r3. */
int32_t orc_satd4x4_wels(const uint8_t *a, int sa, const uint8_t *b, int sb)
{
    int32_t m[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            m[i][j] = (int32_t)a[i * sa + j] - (int32_t)b[i * sb + j];
    for (int i = 0; i < 4; i++) {
        int32_t s0 = m[i][0] + m[i][2], s1 = m[i][1] + m[i][3];
        int32_t s2 = m[i][0] - m[i][2], s3 = m[i][1] - m[i][3];
        m[i][0] = s0 + s1; m[i][1] = s2 + s3; m[i][2] = s2 - s3; m[i][3] = s0 - s1;
    }
    int32_t sum = 0;
    for (int j = 0; j < 4; j++) {
        int32_t s0 = m[0][j] + m[2][j], s1 = m[1][j] + m[3][j];
        int32_t s2 = m[0][j] - m[2][j], s3 = m[1][j] - m[3][j];
        sum += iabs(s0 + s1) + iabs(s2 + s3) + iabs(s2 - s3) + iabs(s0 - s1);
    }
    return (sum + 1) >> 1;
}

/* satd_region_scalar, satd_sad.rs:60-72 (rounding per 4x4 block) */
int32_t orc_satd_wels(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h)
{
    int32_t s = 0;
    for (int by = 0; by < h; by += 4)
        for (int bx = 0; bx < w; bx += 4)
            s += orc_satd4x4_wels(a + by * sa + bx, sa, b + by * sb + bx, sb);
    return s;
}

/* hadamard_1d / hadamard_4x4, crates/rusty_h264-common/src/transform.rs:631-655 */
void orc_hadamard_4x4(const int32_t in[16], int32_t out[16])
{
    int32_t m[16];
    memcpy(m, in, sizeof(m));
    for (int r = 0; r < 4; r++) {
        int32_t a = m[r * 4], b = m[r * 4 + 1], c = m[r * 4 + 2], d = m[r * 4 + 3];
        m[r * 4] = a + b + c + d; m[r * 4 + 1] = a + b - c - d;
        m[r * 4 + 2] = a - b - c + d; m[r * 4 + 3] = a - b + c - d;
    }
    for (int c = 0; c < 4; c++) {
        int32_t a = m[c], b = m[4 + c], cc = m[8 + c], d = m[12 + c];
        m[c] = a + b + cc + d; m[4 + c] = a + b - cc - d;
        m[8 + c] = a - b - cc + d; m[12 + c] = a - b + cc - d;
    }
    memcpy(out, m, sizeof(m));
}

/* satd_px scalar arm, crates/rusty_h264-encoder/src/mb16.rs:6052-6074 with
 * satd_4x4_sum (transform.rs:701-712): sum over 4x4 blocks of sum |H d|. */
int64_t orc_satd_px(const uint8_t *src, int ss, const uint8_t *pred, int ps, int w, int h)
{
    int64_t total = 0;
    for (int by = 0; by < h / 4; by++)
        for (int bx = 0; bx < w / 4; bx++) {
            int32_t blk[16], co[16];
            for (int dy = 0; dy < 4; dy++)
                for (int dx = 0; dx < 4; dx++)
                    blk[dy * 4 + dx] = (int32_t)src[(by * 4 + dy) * ss + bx * 4 + dx] -
                                       (int32_t)pred[(by * 4 + dy) * ps + bx * 4 + dx];
            orc_hadamard_4x4(blk, co);
            for (int k = 0; k < 16; k++) total += iabs(co[k]);
        }
    return total;
}

/* satd_avg, crates/rusty_h264-accel/src/satd_avg.rs:579-602: sum |H (src - (a+b+1)>>1)|,
 * unhalved (the satd_px value of the materialized average). */
/* 8-point Sylvester Hadamard of 8 values (stride st), in place: three butterfly
 * stages over index bits 2, 1, 0; output k is sum_j (-1)^popcount(j & k) x_j. */
static void had8(int32_t *x, int st)
{
    for (int half = 4; half >= 1; half >>= 1)
        for (int i = 0; i < 8; i++)
            if (!(i & half)) {
                int32_t a = x[i * st], b = x[(i + half) * st];
                x[i * st] = a + b;
                x[(i + half) * st] = a - b;
            }
}

void orc_hadamard_8x8(const int32_t in[64], int32_t out[64])
{
    memcpy(out, in, 64 * sizeof(int32_t));
    for (int r = 0; r < 8; r++) had8(out + 8 * r, 1);
    for (int c = 0; c < 8; c++) had8(out + c, 8);
}

int64_t orc_satd8_px(const uint8_t *src, int ss, const uint8_t *pred, int ps, int w, int h)
{
    int64_t s = 0;
    for (int by = 0; by < h; by += 8)
        for (int bx = 0; bx < w; bx += 8) {
            int32_t d[64], co[64];
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++)
                    d[y * 8 + x] = (int)src[(by + y) * ss + bx + x] - (int)pred[(by + y) * ps + bx + x];
            orc_hadamard_8x8(d, co);
            int64_t t = 0;
            for (int k = 0; k < 64; k++) t += iabs(co[k]);
            s += t / 2;
        }
    return s;
}

uint32_t orc_satd_avg(const uint8_t *src, int ss, const uint8_t *a, const uint8_t *b,
                      int stride, int w, int h)
{
    uint8_t avg[256];
    for (int r = 0; r < h; r++)
        for (int i = 0; i < w; i++)
            avg[r * w + i] = (uint8_t)(((int)a[r * stride + i] + (int)b[r * stride + i] + 1) >> 1);
    return (uint32_t)orc_satd_px(src, ss, avg, w, w, h);
}

/* x4_shape, satd_avg.rs:342-344 */
int orc_x4_shape(int w, int h)
{
    return (w == 16 && h == 16) || (w == 16 && h == 8) || (w == 8 && h == 16) || (w == 8 && h == 8);
}

/* ---------------------------------------------------------------------------
 * Luma interpolation
 * ------------------------------------------------------------------------- */

/* at(), inter.rs:120-124 (edge clamp) */
static inline int at(const uint8_t *ref, int cw, int ch, int x, int y)
{
    return ref[clampi(y, 0, ch - 1) * cw + clampi(x, 0, cw - 1)];
}

/* luma_sample, inter.rs:134-166: the spec sec. 8.4.2.2.1 per-pixel quarter-pel sample. */
int orc_luma_sample(const uint8_t *ref, int cw, int ch, int ix, int iy, int fx, int fy)
{
#define G(dx, dy) at(ref, cw, ch, ix + (dx), iy + (dy))
#define HOR6(dy) (G(-2, dy) - 5 * G(-1, dy) + 20 * G(0, dy) + 20 * G(1, dy) - 5 * G(2, dy) + G(3, dy))
#define VER6(dx) (G(dx, -2) - 5 * G(dx, -1) + 20 * G(dx, 0) + 20 * G(dx, 1) - 5 * G(dx, 2) + G(dx, 3))
    if (fx == 0 && fy == 0) return G(0, 0);
    int b = clip_u8((HOR6(0) + 16) >> 5);
    int hh = clip_u8((VER6(0) + 16) >> 5);
    int m = clip_u8((VER6(1) + 16) >> 5);
    int s = clip_u8((HOR6(1) + 16) >> 5);
    int j1 = HOR6(-2) - 5 * HOR6(-1) + 20 * HOR6(0) + 20 * HOR6(1) - 5 * HOR6(2) + HOR6(3);
    int j = clip_u8((j1 + 512) >> 10);
#undef G
#undef HOR6
#undef VER6
    switch (fx * 4 + fy) {
    case 1 * 4 + 0: return (at(ref, cw, ch, ix, iy) + b + 1) >> 1;
    case 2 * 4 + 0: return b;
    case 3 * 4 + 0: return (at(ref, cw, ch, ix + 1, iy) + b + 1) >> 1;
    case 0 * 4 + 1: return (at(ref, cw, ch, ix, iy) + hh + 1) >> 1;
    case 1 * 4 + 1: return (b + hh + 1) >> 1;
    case 2 * 4 + 1: return (b + j + 1) >> 1;
    case 3 * 4 + 1: return (b + m + 1) >> 1;
    case 0 * 4 + 2: return hh;
    case 1 * 4 + 2: return (hh + j + 1) >> 1;
    case 2 * 4 + 2: return j;
    case 3 * 4 + 2: return (j + m + 1) >> 1;
    case 0 * 4 + 3: return (at(ref, cw, ch, ix, iy + 1) + hh + 1) >> 1;
    case 1 * 4 + 3: return (hh + s + 1) >> 1;
    case 2 * 4 + 3: return (j + s + 1) >> 1;
    default: return (m + s + 1) >> 1; /* (3,3) */
    }
}

#define LUMA_TILE 21

/* luma_tile_into, inter.rs:365-402: (bw+5)x(bh+5) clamped neighbourhood, block at (2,2). */
static int luma_tile_into(uint8_t *t, const uint8_t *ref, int cw, int ch, int ix0, int iy0,
                          int bw, int bh)
{
    int ts = bw + 5;
    for (int ty = 0; ty < bh + 5; ty++) {
        int ry = clampi(iy0 - 2 + ty, 0, ch - 1) * cw;
        for (int tx = 0; tx < ts; tx++)
            t[ty * ts + tx] = ref[ry + clampi(ix0 - 2 + tx, 0, cw - 1)];
    }
    return ts;
}

/* luma_h, inter.rs:406-431 (McHorVer20), block shifted by (dr, dc) */
static void luma_h(const uint8_t *t, int ts, int bw, int bh, int dr, int dc, uint8_t *dst)
{
    for (int r = 0; r < bh; r++) {
        const uint8_t *w = t + (2 + r + dr) * ts + 2 + dc - 2;
        for (int c = 0; c < bw; c++) {
            int f = w[c] - 5 * w[c + 1] + 20 * w[c + 2] + 20 * w[c + 3] - 5 * w[c + 4] + w[c + 5];
            dst[r * bw + c] = (uint8_t)clip_u8((f + 16) >> 5);
        }
    }
}

/* luma_v, inter.rs:434-455 (McHorVer02) */
static void luma_v(const uint8_t *t, int ts, int bw, int bh, int dr, int dc, uint8_t *dst)
{
    for (int r = 0; r < bh; r++) {
        int base = (2 + r + dr) * ts + 2 + dc;
        for (int c = 0; c < bw; c++) {
            int p = base + c;
            int f = t[p - 2 * ts] - 5 * t[p - ts] + 20 * t[p] + 20 * t[p + ts] - 5 * t[p + 2 * ts] +
                    t[p + 3 * ts];
            dst[r * bw + c] = (uint8_t)clip_u8((f + 16) >> 5);
        }
    }
}

/* luma_centre, inter.rs:459-488 (McHorVer22): vertical 6-tap (unrounded i32), then
 * horizontal 6-tap, (+512)>>10. */
static void luma_centre(const uint8_t *t, int ts, int bw, int bh, uint8_t *dst)
{
    int32_t itmp[LUMA_TILE];
    for (int r = 0; r < bh; r++) {
        int base = (2 + r) * ts;
        for (int j = 0; j < bw + 5; j++) {
            int p = base + j;
            itmp[j] = t[p - 2 * ts] - 5 * t[p - ts] + 20 * t[p] + 20 * t[p + ts] -
                      5 * t[p + 2 * ts] + t[p + 3 * ts];
        }
        for (int c = 0; c < bw; c++) {
            const int32_t *w = itmp + c;
            int f = w[0] - 5 * w[1] + 20 * w[2] + 20 * w[3] - 5 * w[4] + w[5];
            dst[r * bw + c] = (uint8_t)clip_u8((f + 512) >> 10);
        }
    }
}

/* pixel_avg, inter.rs:491-520 */
static void pixel_avg(const uint8_t *a, const uint8_t *b, int bw, int bh, uint8_t *dst)
{
    for (int i = 0; i < bw * bh; i++) dst[i] = (uint8_t)((a[i] + b[i] + 1) >> 1);
}

/* avg_full, inter.rs:523-545: full-pel block shifted by (dr, dc) averaged with a half plane */
static void avg_full(const uint8_t *t, int ts, int bw, int bh, int dr, int dc,
                     const uint8_t *half, uint8_t *dst)
{
    for (int r = 0; r < bh; r++) {
        const uint8_t *trow = t + (2 + r + dr) * ts + 2 + dc;
        for (int c = 0; c < bw; c++)
            dst[r * bw + c] = (uint8_t)((trow[c] + half[r * bw + c] + 1) >> 1);
    }
}

/* mc_luma_subpel, inter.rs:660-694, with the compose forms of qpel_one_filter
 * (568-594), qpel_hv (598-619) and qpel_centre_adj (627-654). */
static void mc_luma_subpel(const uint8_t *t, int ts, int bw, int bh, int fx, int fy, uint8_t *out)
{
    uint8_t a[256], b[256];
    switch (fx * 4 + fy) {
    case 2 * 4 + 0: luma_h(t, ts, bw, bh, 0, 0, out); break;
    case 0 * 4 + 2: luma_v(t, ts, bw, bh, 0, 0, out); break;
    case 2 * 4 + 2: luma_centre(t, ts, bw, bh, out); break;
    case 1 * 4 + 0: luma_h(t, ts, bw, bh, 0, 0, a); avg_full(t, ts, bw, bh, 0, 0, a, out); break;
    case 3 * 4 + 0: luma_h(t, ts, bw, bh, 0, 0, a); avg_full(t, ts, bw, bh, 0, 1, a, out); break;
    case 0 * 4 + 1: luma_v(t, ts, bw, bh, 0, 0, a); avg_full(t, ts, bw, bh, 0, 0, a, out); break;
    case 0 * 4 + 3: luma_v(t, ts, bw, bh, 0, 0, a); avg_full(t, ts, bw, bh, 1, 0, a, out); break;
    case 1 * 4 + 1: luma_h(t, ts, bw, bh, 0, 0, a); luma_v(t, ts, bw, bh, 0, 0, b); pixel_avg(a, b, bw, bh, out); break;
    case 3 * 4 + 1: luma_h(t, ts, bw, bh, 0, 0, a); luma_v(t, ts, bw, bh, 0, 1, b); pixel_avg(a, b, bw, bh, out); break;
    case 1 * 4 + 3: luma_h(t, ts, bw, bh, 1, 0, a); luma_v(t, ts, bw, bh, 0, 0, b); pixel_avg(a, b, bw, bh, out); break;
    case 3 * 4 + 3: luma_h(t, ts, bw, bh, 1, 0, a); luma_v(t, ts, bw, bh, 0, 1, b); pixel_avg(a, b, bw, bh, out); break;
    case 2 * 4 + 1: luma_h(t, ts, bw, bh, 0, 0, a); luma_centre(t, ts, bw, bh, b); pixel_avg(a, b, bw, bh, out); break;
    case 2 * 4 + 3: luma_h(t, ts, bw, bh, 1, 0, a); luma_centre(t, ts, bw, bh, b); pixel_avg(a, b, bw, bh, out); break;
    case 1 * 4 + 2: luma_v(t, ts, bw, bh, 0, 0, a); luma_centre(t, ts, bw, bh, b); pixel_avg(a, b, bw, bh, out); break;
    case 3 * 4 + 2: luma_v(t, ts, bw, bh, 0, 1, a); luma_centre(t, ts, bw, bh, b); pixel_avg(a, b, bw, bh, out); break;
    default: break;
    }
}

/* mc_luma, inter.rs:1193-1262 */
void orc_mc_luma(const uint8_t *ref, int cw, int ch, int x0, int y0, int bw, int bh,
                 int mvx, int mvy, uint8_t *out)
{
    int ix0 = x0 + asr2(mvx), iy0 = y0 + asr2(mvy);
    int fx = mvx & 3, fy = mvy & 3;
    if (fx == 0 && fy == 0) {
        for (int dy = 0; dy < bh; dy++)
            for (int dx = 0; dx < bw; dx++)
                out[dy * bw + dx] = (uint8_t)at(ref, cw, ch, ix0 + dx, iy0 + dy);
        return;
    }
    uint8_t tile[LUMA_TILE * LUMA_TILE];
    int ts = luma_tile_into(tile, ref, cw, ch, ix0, iy0, bw, bh);
    mc_luma_subpel(tile, ts, bw, bh, fx, fy, out);
}

/* pad_plane, inter.rs:786-814 (openh264 ExpandPicture) */
void orc_pad_plane(const uint8_t *src, int w, int h, int pad, uint8_t *out)
{
    int pw = w + 2 * pad, ph = h + 2 * pad;
    for (int y = 0; y < ph; y++) {
        const uint8_t *row = src + clampi(y - pad, 0, h - 1) * w;
        uint8_t *d = out + (size_t)y * pw;
        memset(d, row[0], (size_t)pad);
        memcpy(d + pad, row, (size_t)w);
        memset(d + pad + w, row[w - 1], (size_t)pad);
    }
}

/* build_hpel_tiles, inter.rs:868-897: 16x16 tile walk over the padded plane through
 * luma_tile_into + luma_h/luma_v/luma_centre (clamped to the padded plane). */
void orc_build_hpel_tiles(const uint8_t *f, int pw, int ph, uint8_t *h, uint8_t *v, uint8_t *c)
{
    uint8_t tile[LUMA_TILE * LUMA_TILE], blk[256];
    for (int by = 0; by < ph; by += 16)
        for (int bx = 0; bx < pw; bx += 16) {
            int bh = ph - by < 16 ? ph - by : 16;
            int bwid = pw - bx < 16 ? pw - bx : 16;
            int ts = luma_tile_into(tile, f, pw, ph, bx, by, bwid, bh);
            for (int kind = 0; kind < 3; kind++) {
                uint8_t *plane = kind == 0 ? h : (kind == 1 ? v : c);
                if (kind == 0) luma_h(tile, ts, bwid, bh, 0, 0, blk);
                else if (kind == 1) luma_v(tile, ts, bwid, bh, 0, 0, blk);
                else luma_centre(tile, ts, bwid, bh, blk);
                for (int r = 0; r < bh; r++)
                    memcpy(plane + (size_t)(by + r) * pw + bx, blk + r * bwid, (size_t)bwid);
            }
        }
}

/* build_hpel_fused, inter.rs:908-987: single pass, row-clamped vertical taps into
 * i32 intermediates vt (column-clamped halo), then H from the clamped row, V from
 * vt, C from the 6-tap over vt. */
void orc_build_hpel_fused(const uint8_t *f, int pw, int ph, uint8_t *h, uint8_t *v, uint8_t *c)
{
    int32_t *vt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(pw + 5));
    int32_t *hb = (int32_t *)malloc(sizeof(int32_t) * (size_t)(pw + 5));
    for (int y = 0; y < ph; y++) {
        const uint8_t *rows[6];
        for (int k = 0; k < 6; k++) rows[k] = f + (size_t)clampi(y - 2 + k, 0, ph - 1) * pw;
        for (int j = 0; j < pw + 5; j++) {
            int x = clampi(j - 2, 0, pw - 1);
            vt[j] = rows[0][x] - 5 * rows[1][x] + 20 * rows[2][x] + 20 * rows[3][x] - 5 * rows[4][x] +
                    rows[5][x];
            hb[j] = rows[2][x];
        }
        for (int x = 0; x < pw; x++) {
            v[(size_t)y * pw + x] = (uint8_t)clip_u8((vt[x + 2] + 16) >> 5);
            const int32_t *w = vt + x;
            int s = w[0] - 5 * w[1] + 20 * w[2] + 20 * w[3] - 5 * w[4] + w[5];
            c[(size_t)y * pw + x] = (uint8_t)clip_u8((s + 512) >> 10);
            const int32_t *q = hb + x;
            int t = q[0] - 5 * q[1] + 20 * q[2] + 20 * q[3] - 5 * q[4] + q[5];
            h[(size_t)y * pw + x] = (uint8_t)clip_u8((t + 16) >> 5);
        }
    }
    free(vt);
    free(hb);
}

/* build_hpel_planes, inter.rs:816-843 */
void orc_build_hpel_planes(const uint8_t *y, int cw, int ch, int pad, int fused, uint8_t *planes)
{
    int pw = cw + 2 * pad, ph = ch + 2 * pad;
    size_t n = (size_t)pw * ph;
    orc_pad_plane(y, cw, ch, pad, planes);
    if (fused) orc_build_hpel_fused(planes, pw, ph, planes + n, planes + 2 * n, planes + 3 * n);
    else orc_build_hpel_tiles(planes, pw, ph, planes + n, planes + 2 * n, planes + 3 * n);
}

/* Quarter-phase operand table shared by hpel_qpel_refs (inter.rs:1060-1105) and
 * MeCtx::eval (crates/rusty_h264-accel/src/mectx.rs:143-156): plane indices
 * (0=F,1=H,2=V,3=C) and (x,y) offsets of the two averaged operands. */
static const int QTAB[16][6] = {
    /* idx = fx*4+fy : pa, ax, ay, pb, bx, by */
    [1 * 4 + 0] = {0, 0, 0, 1, 0, 0}, [3 * 4 + 0] = {0, 1, 0, 1, 0, 0},
    [0 * 4 + 1] = {0, 0, 0, 2, 0, 0}, [0 * 4 + 3] = {0, 0, 1, 2, 0, 0},
    [1 * 4 + 1] = {1, 0, 0, 2, 0, 0}, [3 * 4 + 1] = {1, 0, 0, 2, 1, 0},
    [1 * 4 + 3] = {1, 0, 1, 2, 0, 0}, [3 * 4 + 3] = {1, 0, 1, 2, 1, 0},
    [2 * 4 + 1] = {1, 0, 0, 3, 0, 0}, [2 * 4 + 3] = {1, 0, 1, 3, 0, 0},
    [1 * 4 + 2] = {2, 0, 0, 3, 0, 0}, [3 * 4 + 2] = {2, 1, 0, 3, 0, 0},
};

/* hpel_block, inter.rs:1107-1162 (with hpel_qpel_refs' guard and table) */
int orc_hpel_block(const uint8_t *planes, int cw, int ch, int pad, int x0, int y0, int bw, int bh,
                   int mvx, int mvy, uint8_t *out)
{
    int pw = cw + 2 * pad, ph = ch + 2 * pad;
    size_t n = (size_t)pw * ph;
    int ix0 = x0 + asr2(mvx), iy0 = y0 + asr2(mvy);
    int fx = mvx & 3, fy = mvy & 3;
    if (fx == 0 && fy == 0) return 0;
    int px = ix0 + pad, py = iy0 + pad;
    if (px < 0 || py < 0 || px + bw + 1 > pw || py + bh + 1 > ph) return 0;
    size_t base = (size_t)py * pw + px;
    if ((fx & 1) == 0 && (fy & 1) == 0) {
        const uint8_t *src = planes + n * (fx == 2 ? (fy == 2 ? 3 : 1) : 2);
        for (int r = 0; r < bh; r++) memcpy(out + r * bw, src + base + (size_t)r * pw, (size_t)bw);
        return 1;
    }
    if (!(bw == 16 || bw == 8 || bw == 4)) return 0;
    const int *q = QTAB[fx * 4 + fy];
    const uint8_t *pa = planes + n * q[0] + base + q[1] + (size_t)q[2] * pw;
    const uint8_t *pb = planes + n * q[3] + base + q[4] + (size_t)q[5] * pw;
    for (int r = 0; r < bh; r++)
        for (int i = 0; i < bw; i++)
            out[r * bw + i] = (uint8_t)((pa[(size_t)r * pw + i] + pb[(size_t)r * pw + i] + 1) >> 1);
    return 1;
}

static inline int plane_at(const uint8_t *planes, size_t n, int pw, int ph, int k, int x, int y)
{
    return planes[n * k + (size_t)clampi(y, 0, ph - 1) * pw + clampi(x, 0, pw - 1)];
}

void orc_pred_block(const uint8_t *planes, int cw, int ch, int pad, int x0, int y0, int bw,
                    int bh, int mvx, int mvy, uint8_t *out)
{
    int pw = cw + 2 * pad, ph = ch + 2 * pad;
    size_t n = (size_t)pw * ph;
    int px = x0 + asr2(mvx) + pad, py = y0 + asr2(mvy) + pad;
    int fx = mvx & 3, fy = mvy & 3;
    for (int r = 0; r < bh; r++)
        for (int i = 0; i < bw; i++) {
            int v;
            if ((fx & 1) == 0 && (fy & 1) == 0) {
                int k = fx == 0 ? (fy == 0 ? 0 : 2) : (fy == 0 ? 1 : 3);
                v = plane_at(planes, n, pw, ph, k, px + i, py + r);
            } else {
                const int *q = QTAB[fx * 4 + fy];
                int a = plane_at(planes, n, pw, ph, q[0], px + i + q[1], py + r + q[2]);
                int b = plane_at(planes, n, pw, ph, q[3], px + i + q[4], py + r + q[5]);
                v = (a + b + 1) >> 1;
            }
            out[r * bw + i] = (uint8_t)v;
        }
}

/* ---------------------------------------------------------------------------
 * Motion search
 * ------------------------------------------------------------------------- */

/* mvbits, mb16.rs:2474-2475: 1 + 2*floor(log2(codenum+1)) */
uint32_t orc_mvbits(int32_t d)
{
    d = clampi(d, -(1 << 24), 1 << 24); /* keeps 2d-1 from wrapping; no coded MV is this large */
    uint32_t codenum = d > 0 ? (uint32_t)(2 * d - 1) : (uint32_t)(-2 * d);
    return 1u + 2u * (31u - (uint32_t)__builtin_clz(codenum + 1u));
}

/* (lambda * rate as f64) as i64, mb16.rs:2563: truncation toward zero */
void orc_lambda_table(double lambda, int n, int64_t *lr)
{
    for (int b = 0; b < n; b++) lr[b] = (int64_t)(lambda * (double)b);
}

static inline int64_t rate_cost(const int64_t *lr, int mvx, int mvy, int px, int py)
{
    uint32_t r = orc_mvbits(mvx - px) + orc_mvbits(mvy - py);
    return lr[r < ORC_LR_N ? r : ORC_LR_N - 1];
}

/* SAD of the block against the full-pel plane F with clamped coordinates
 * (the mc_sad_hp values, mb16.rs:2018-2079, for full-pel candidates). */
static int32_t sad_fullpel(const orc_search_cfg *cfg, const uint8_t *cur, const uint8_t *planes,
                           int lx, int ly, int mvx, int mvy)
{
    int pw = cfg->cw + 2 * cfg->pad, ph = cfg->ch + 2 * cfg->pad;
    int px = lx + asr2(mvx) + cfg->pad, py = ly + asr2(mvy) + cfg->pad;
    int32_t s = 0;
    if (px >= 0 && py >= 0 && px + cfg->bw <= pw && py + cfg->bh <= ph)
        return orc_sad(cur + (size_t)ly * cfg->cw + lx, cfg->cw, planes + (size_t)py * pw + px, pw,
                       cfg->bw, cfg->bh);
    for (int r = 0; r < cfg->bh; r++)
        for (int i = 0; i < cfg->bw; i++)
            s += iabs((int)cur[(size_t)(ly + r) * cfg->cw + lx + i] -
                      plane_at(planes, 0, pw, ph, 0, px + i, py + r));
    return s;
}

int64_t orc_full_search_block(const orc_search_cfg *cfg, const uint8_t *cur,
                              const uint8_t *planes, int lx, int ly, int cx, int cy,
                              int mvpx, int mvpy, int32_t *mv_out)
{
    int R = cfg->range;
    int64_t best = INT64_MAX;
    int bx = cx, by = cy;
    /* raster: dy outer, dx inner, strict < (first minimum wins), mb16.rs:2810-2856 */
    for (int dy = -R; dy <= R; dy++)
        for (int dx = -R; dx <= R; dx++) {
            int mvx = cx + 4 * dx, mvy = cy + 4 * dy;
            int64_t c = sad_fullpel(cfg, cur, planes, lx, ly, mvx, mvy) +
                        rate_cost(cfg->lr_fp, mvx, mvy, mvpx, mvpy);
            if (c < best) { best = c; bx = mvx; by = mvy; }
        }
    mv_out[0] = bx;
    mv_out[1] = by;
    return best;
}

static int64_t satd_dist(const orc_search_cfg *cfg, const uint8_t *cur, const uint8_t *planes,
                         int lx, int ly, int mvx, int mvy)
{
    uint8_t pred[256];
    orc_pred_block(planes, cfg->cw, cfg->ch, cfg->pad, lx, ly, cfg->bw, cfg->bh, mvx, mvy, pred);
    if (cfg->hadamard == 8)
        return orc_satd8_px(cur + (size_t)ly * cfg->cw + lx, cfg->cw, pred, cfg->bw, cfg->bw, cfg->bh);
    return orc_satd_px(cur + (size_t)ly * cfg->cw + lx, cfg->cw, pred, cfg->bw, cfg->bw, cfg->bh);
}

int64_t orc_satd_cost(const orc_search_cfg *cfg, const uint8_t *cur, const uint8_t *planes,
                      int lx, int ly, int mvx, int mvy, int mvpx, int mvpy)
{
    return satd_dist(cfg, cur, planes, lx, ly, mvx, mvy) +
           rate_cost(cfg->lr, mvx, mvy, mvpx, mvpy);
}

/* Quarter-pel window (build-defined, subpel_pat 4): the 24 positions around the
 * start (*bx, *by) within +-2 quarter pels, in raster order (dy, then dx), each
 * replacing the best on a strictly lower cost; *best holds the start's cost on
 * entry.  Returns the number of positions priced (24). */
static int qpel_window(const orc_search_cfg *cfg, const uint8_t *cur, const uint8_t *planes, int lx, int ly,
                       int mvpx, int mvpy, int *bx, int *by, int64_t *best)
{
    const int sx = *bx, sy = *by;
    int n = 0;
    for (int dy = -2; dy <= 2; dy++)
        for (int dx = -2; dx <= 2; dx++) {
            if (dx == 0 && dy == 0) continue;
            int64_t cc = orc_satd_cost(cfg, cur, planes, lx, ly, sx + dx, sy + dy, mvpx, mvpy);
            n++;
            if (cc < *best) { *best = cc; *bx = sx + dx; *by = sy + dy; }
        }
    return n;
}

/* Sub-pel ring walk, mb16.rs:3046-3046 ring8 order (s,0),(-s,0),(0,s),(0,-s),(s,s),(-s,-s),
 * (s,-s),(-s,s); cascade with a moving centre (mb16.rs:3230-3275); steps [2,1];
 * pattern bit0 = ring4, bit1 = single pass (mb16.rs:2976-2998). */
int64_t orc_subpel_refine_block(const orc_search_cfg *cfg, const uint8_t *cur,
                                const uint8_t *planes, int lx, int ly, int mvx, int mvy,
                                int mvpx, int mvpy, int32_t *mv_out, int32_t *evals)
{
    int bx = mvx, by = mvy;
    int64_t best = orc_satd_cost(cfg, cur, planes, lx, ly, bx, by, mvpx, mvpy);
    int ne = 1;
    int pat = cfg->subpel_pat;
    if (pat == 4) {
        ne += qpel_window(cfg, cur, planes, lx, ly, mvpx, mvpy, &bx, &by, &best);
    } else if (pat >= 0) {
        for (int step = 2; step >= 1; step--) {
            const int ring[8][2] = {{step, 0}, {-step, 0}, {0, step}, {0, -step},
                                    {step, step}, {-step, -step}, {step, -step}, {-step, step}};
            int nring = (pat & 1) ? 4 : 8;
            for (;;) {
                int improved = 0;
                for (int i = 0; i < nring; i++) {
                    int cx = bx + ring[i][0], cy = by + ring[i][1];
                    int64_t cc = orc_satd_cost(cfg, cur, planes, lx, ly, cx, cy, mvpx, mvpy);
                    ne++;
                    if (cc < best) { best = cc; bx = cx; by = cy; improved = 1; }
                }
                if (!improved || (pat & 2)) break;
            }
        }
    }
    mv_out[0] = bx;
    mv_out[1] = by;
    if (evals) *evals = ne;
    return best;
}

void orc_plain_search_frame(const orc_search_cfg *cfg, const uint8_t *cur, const uint8_t *planes,
                            const int32_t *centres, const int32_t *mvps, int32_t *int_mv,
                            int64_t *int_cost, int32_t *mv, int64_t *cost, int nthreads)
{
    int nbx = cfg->cw / cfg->bw, nby = cfg->ch / cfg->bh;
    int nb = nbx * nby;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (int b = 0; b < nb; b++) {
        int lx = (b % nbx) * cfg->bw, ly = (b / nbx) * cfg->bh;
        int cx = centres ? centres[2 * b] : 0, cy = centres ? centres[2 * b + 1] : 0;
        int px = mvps ? mvps[2 * b] : 0, py = mvps ? mvps[2 * b + 1] : 0;
        int32_t imv[2], fmv[2];
        int64_t ic = orc_full_search_block(cfg, cur, planes, lx, ly, cx, cy, px, py, imv);
        int64_t fc = ic;
        fmv[0] = imv[0];
        fmv[1] = imv[1];
        fc = orc_subpel_refine_block(cfg, cur, planes, lx, ly, imv[0], imv[1], px, py, fmv, NULL);
        if (int_mv) { int_mv[2 * b] = imv[0]; int_mv[2 * b + 1] = imv[1]; }
        if (int_cost) int_cost[b] = ic;
        if (mv) { mv[2 * b] = fmv[0]; mv[2 * b + 1] = fmv[1]; }
        if (cost) cost[b] = fc;
    }
}

/* SAD of any candidate through the prediction read (mc_sad_hp, mb16.rs:2018-2079) */
static int32_t sad_any(const orc_search_cfg *cfg, const uint8_t *cur, const uint8_t *planes,
                     int lx, int ly, int mvx, int mvy)
{
    if ((mvx & 3) == 0 && (mvy & 3) == 0) return sad_fullpel(cfg, cur, planes, lx, ly, mvx, mvy);
    uint8_t pred[256];
    orc_pred_block(planes, cfg->cw, cfg->ch, cfg->pad, lx, ly, cfg->bw, cfg->bh, mvx, mvy, pred);
    return orc_sad(cur + (size_t)ly * cfg->cw + lx, cfg->cw, pred, cfg->bw, cfg->bw, cfg->bh);
}

/* ---- reference-mode motion_search (mb16.rs:2433-3302) ------------------------
 *
 * The search restated step by step, every knob explicit (orc_diamond_cfg):
 *   rate  mvbits(d) = Exp-Golomb se(d) length (mb16.rs:2474-2475), or with a table
 *         (mv_cost_kind 1/2, mb16.rs:2459-2471) tab[min(|d|, 4095)];
 *   cost  the closure of mb16.rs:2540-2565: dist + (int64)(lam_r * rate), lam_r =
 *         lambda * 0.25 with a table, else lambda; dist = the SATD of the prediction
 *         (encoder domain) or, in the Fast preset (sad_only), its SAD;
 *   cost_fp  B2 (mb16.rs:2571-2578): SAD + (int64)(lambda_fp * rate), or cost;
 *   1. seeds (0,0) then the predictors, strict < (mb16.rs:2580-2594);
 *   2. snap a fractional winner to ((v+2) div_euclid 4)*4 (mb16.rs:2602-2607);
 *   3. diamond rungs, each walked until it stops improving (mb16.rs:2632-2727):
 *      with fc, on a full-pel centre whose +-step box lies inside the frame and a
 *      16x16/16x8/8x16/8x8 block, the fixed-centre argmin of the 4-point ring
 *      (first minimum in ring order, strict < against the centre, mb16.rs:2646-
 *      2706), 16-wide blocks priced as the x4 batch does, dist + (int64)(lam_fp *
 *      rate) with lam_fp = lambda_fp under B2 else lambda (mb16.rs:2680-2683);
 *      otherwise one cascade pass (first improver moves the centre, mb16.rs:2708-
 *      2725);
 *   4. B2 reprice of the winner and the seed in the cost domain (mb16.rs:2739-2742);
 *   5. the me_wide rescue (mb16.rs:2778-2889): when enabled, me_fast or the flat
 *      test passes, and (best - (int64)(lambda * rate(best))) / (w*h) exceeds the
 *      threshold: a step-2 grid over +-range around the snapped winner (strict <;
 *      a 16x16 block whose grid lies inside the frame prices dist + (int64)(lambda
 *      * rate), mb16.rs:2822-2836), then the 3x3 +-1 pel walk with a moving centre;
 *      events report that it ran and whether it cut the cost by >= 1/16;
 *   6. the sub-pel rings (steps 2, 1; ring8 or ring4; walked or single pass) or the
 *      build-defined quarter-pel window (subpel_pat 4), priced by cost;
 *   7. with the snap, the pre-snap seed is kept when cheaper (mb16.rs:3297-3300).
 * The sub-pel memo (mb16.rs:3006-3029) is value-transparent and not restated; the
 * sub-pel FC batches (RFF_SP_FC), the pattern dispatch and the iteration cap are off
 * by default (mb16.rs:663-676, 940-958, 740-750) and not restated.
 * Trace (optional): every priced candidate as {phase, mvx, mvy, cost}, int64 each;
 * phases 0 seed, 1 snap, 2 cascade, 3 fixed-centre, 4 reprice, 5 rescue grid,
 * 6 rescue walk, 7 sub-pel. */
static int div_euclid4(int v) { return v >= 0 ? v / 4 : -((-v + 3) / 4); }

typedef struct {
    const orc_search_cfg *cfg;
    const orc_diamond_cfg *dc;
    const uint8_t *cur, *planes;
    int lx, ly, cx0, cy0;
    int64_t *trace;
    int cap, n;
} ms_state;

uint32_t orc_mv_rate(const uint16_t *tab, int32_t d)
{
    if (!tab) return orc_mvbits(d);
    uint32_t a = d < 0 ? (uint32_t)(-(int64_t)d) : (uint32_t)d;
    return tab[a < 4095u ? a : 4095u];
}

static uint32_t ms_rate(const ms_state *S, int x, int y)
{
    return orc_mv_rate(S->dc->mv_tab, x - S->cx0) + orc_mv_rate(S->dc->mv_tab, y - S->cy0);
}

static int64_t lam_rate(double lam, uint32_t rate) { return (int64_t)(lam * (double)rate); }

static void ms_note(ms_state *S, int phase, int x, int y, int64_t c)
{
    if (S->trace && S->n < S->cap) {
        int64_t *t = S->trace + 4 * (size_t)S->n;
        t[0] = phase;
        t[1] = x;
        t[2] = y;
        t[3] = c;
    }
    S->n++;
}

static int64_t ms_dist(const ms_state *S, int x, int y, int sad)
{
    if (sad) return sad_any(S->cfg, S->cur, S->planes, S->lx, S->ly, x, y);
    return satd_dist(S->cfg, S->cur, S->planes, S->lx, S->ly, x, y);
}

/* the cost closure (mb16.rs:2540-2565) */
static int64_t ms_cost(ms_state *S, int phase, int x, int y)
{
    const orc_diamond_cfg *dc = S->dc;
    double lam_r = dc->mv_tab ? dc->lambda * 0.25 : dc->lambda;
    int64_t c = ms_dist(S, x, y, dc->sad_only) + lam_rate(lam_r, ms_rate(S, x, y));
    ms_note(S, phase, x, y, c);
    return c;
}

/* cost_fp (mb16.rs:2571-2578) */
static int64_t ms_cost_fp(ms_state *S, int phase, int x, int y)
{
    const orc_diamond_cfg *dc = S->dc;
    if (!dc->sadfp) return ms_cost(S, phase, x, y);
    int64_t c = ms_dist(S, x, y, 1) + lam_rate(dc->lambda_fp, ms_rate(S, x, y));
    ms_note(S, phase, x, y, c);
    return c;
}

/* a candidate priced with an explicit lambda (the x4 and grid batches) */
static int64_t ms_cost_lam(ms_state *S, int phase, int x, int y, int sad, double lam)
{
    int64_t c = ms_dist(S, x, y, sad) + lam_rate(lam, ms_rate(S, x, y));
    ms_note(S, phase, x, y, c);
    return c;
}

int64_t orc_motion_search_ex(const orc_search_cfg *cfg, const orc_diamond_cfg *dc, const uint8_t *cur,
                             const uint8_t *planes, int lx, int ly, const int32_t *preds, int npreds,
                             int32_t *mv_out, int32_t *events, int64_t *trace, int trace_cap, int *trace_n)
{
    ms_state S = {cfg, dc, cur, planes, lx, ly, npreds > 0 ? preds[0] : 0, npreds > 0 ? preds[1] : 0,
                  trace, trace_cap, 0};
    const int bw = cfg->bw, bh = cfg->bh;
    if (events) events[0] = events[1] = 0;
    /* 1. seeds */
    int bx = 0, by = 0;
    int64_t bc = ms_cost_fp(&S, 0, 0, 0);
    for (int i = 0; i < npreds; i++) {
        int64_t pc = ms_cost_fp(&S, 0, preds[2 * i], preds[2 * i + 1]);
        if (pc < bc) { bc = pc; bx = preds[2 * i]; by = preds[2 * i + 1]; }
    }
    const int seedx = bx, seedy = by;
    int64_t seed_c = bc;
    /* 2. snap */
    if (dc->snap && ((bx & 3) != 0 || (by & 3) != 0)) {
        bx = div_euclid4(bx + 2) * 4;
        by = div_euclid4(by + 2) * 4;
        bc = ms_cost_fp(&S, 1, bx, by);
    }
    /* 3. diamond */
    const int fc = dc->fc && !dc->sad_only && ((bw == 16 && (bh == 16 || bh == 8)) || (bw == 8 && (bh == 16 || bh == 8)));
    const double lam_fc = dc->sadfp ? dc->lambda_fp : dc->lambda;
    for (int si = 0; si < dc->nsteps; si++) {
        const int step = dc->steps[si];
        const int ring[4][2] = {{step, 0}, {-step, 0}, {0, step}, {0, -step}};
        for (;;) {
            if (fc && (bx & 3) == 0 && (by & 3) == 0) {
                const int st = step >> 2, ix = lx + (bx >> 2), iy = ly + (by >> 2);
                if (ix - st >= 0 && iy - st >= 0 && ix + st + bw <= cfg->cw && iy + st + bh <= cfg->ch) {
                    int bi = -1;
                    int64_t bcc = bc;
                    for (int i = 0; i < 4; i++) {
                        const int x = bx + ring[i][0], y = by + ring[i][1];
                        int64_t cc = bw == 16 ? ms_cost_lam(&S, 3, x, y, dc->sadfp, lam_fc) : ms_cost_fp(&S, 3, x, y);
                        if (cc < bcc) { bcc = cc; bi = i; }
                    }
                    if (bi < 0) break;
                    bc = bcc;
                    bx += ring[bi][0];
                    by += ring[bi][1];
                    continue;
                }
            }
            int improved = 0;
            for (int i = 0; i < 4; i++) {
                const int x = bx + ring[i][0], y = by + ring[i][1];
                int64_t cc = ms_cost_fp(&S, 2, x, y);
                if (cc < bc) { bc = cc; bx = x; by = y; improved = 1; }
            }
            if (!improved) break;
        }
    }
    /* 4. B2 reprice */
    if (dc->sadfp) {
        bc = ms_cost(&S, 4, bx, by);
        seed_c = ms_cost(&S, 4, seedx, seedy);
    }
    /* 5. me_wide rescue */
    int flat = 0;
    if (dc->rescue && !dc->sad_only && !dc->rescue_fast) {
        uint64_t s = 0, ss = 0, n = (uint64_t)bw * bh;
        for (int r = 0; r < bh; r++)
            for (int i = 0; i < bw; i++) {
                uint64_t v = cur[(size_t)(ly + r) * cfg->cw + lx + i];
                s += v;
                ss += v * v;
            }
        flat = (ss - s * s / n) / n < (uint64_t)dc->rescue_var;
    }
    if (dc->rescue && !dc->sad_only && (dc->rescue_fast || flat)) {
        int64_t dist = bc - lam_rate(dc->lambda, ms_rate(&S, bx, by));
        if (dist / (int64_t)(bw * bh) > dc->rescue_thresh) {
            const int64_t pre_c = bc;
            const int gx = div_euclid4(bx + 2) * 4, gy = div_euclid4(by + 2) * 4;
            const int r = dc->rescue_range;
            const int icx = lx + (gx >> 2), icy = ly + (gy >> 2);
            const int batched = bw == 16 && bh == 16 && icx >= r && icx + r + 16 <= cfg->cw && icy >= r &&
                                icy + r + 16 <= cfg->ch;
            int gbx = bx, gby = by;
            for (int dy = -r; dy <= r; dy += 2)
                for (int dx = -r; dx <= r; dx += 2) {
                    const int x = gx + 4 * dx, y = gy + 4 * dy;
                    int64_t cc = batched ? ms_cost_lam(&S, 5, x, y, 0, dc->lambda) : ms_cost(&S, 5, x, y);
                    if (cc < bc) { bc = cc; gbx = x; gby = y; }
                }
            bx = gbx;
            by = gby;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int x = bx + 4 * dx, y = by + 4 * dy;
                    int64_t cc = ms_cost(&S, 6, x, y);
                    if (cc < bc) { bc = cc; bx = x; by = y; }
                }
            if (events) {
                events[0] = 1;
                events[1] = bc * 16 <= pre_c * 15;
            }
        }
    }
    /* 6. sub-pel */
    const int pat = cfg->subpel_pat;
    if (pat == 4) {
        const int sx = bx, sy = by;
        for (int dy = -2; dy <= 2; dy++)
            for (int dx = -2; dx <= 2; dx++) {
                if (dx == 0 && dy == 0) continue;
                int64_t cc = ms_cost(&S, 7, sx + dx, sy + dy);
                if (cc < bc) { bc = cc; bx = sx + dx; by = sy + dy; }
            }
    } else if (pat >= 0) {
        for (int step = 2; step >= 1; step--) {
            const int ring[8][2] = {{step, 0}, {-step, 0}, {0, step}, {0, -step},
                                    {step, step}, {-step, -step}, {step, -step}, {-step, step}};
            const int nring = (pat & 1) ? 4 : 8;
            for (;;) {
                int improved = 0;
                for (int i = 0; i < nring; i++) {
                    const int x = bx + ring[i][0], y = by + ring[i][1];
                    int64_t cc = ms_cost(&S, 7, x, y);
                    if (cc < bc) { bc = cc; bx = x; by = y; improved = 1; }
                }
                if (!improved || (pat & 2)) break;
            }
        }
    }
    /* 7. the seed */
    if (dc->snap && seed_c < bc) { bx = seedx; by = seedy; bc = seed_c; }
    mv_out[0] = bx;
    mv_out[1] = by;
    if (trace_n) *trace_n = S.n;
    return bc;
}

int64_t orc_motion_search_ref(const orc_search_cfg *cfg, const orc_diamond_cfg *dc,
                              const uint8_t *cur, const uint8_t *planes, int lx, int ly,
                              const int32_t *preds, int npreds, int32_t *mv_out)
{
    return orc_motion_search_ex(cfg, dc, cur, planes, lx, ly, preds, npreds, mv_out, NULL, NULL, 0, NULL);
}

/* median3 (inter.rs:9-11) and predict_mv (inter.rs:36-73). */
static int median3(int a, int b, int c)
{
    int mn = a < b ? a : b, mx = a > b ? a : b;
    mn = c < mn ? c : mn;
    mx = c > mx ? c : mx;
    return a + b + c - mn - mx;
}

void orc_predict_mv(const int32_t nb[12], int cur_ref, int32_t out[2])
{
    int mv[3][2], ref[3];
    for (int k = 0; k < 3; k++) {
        const int32_t *n = nb + 4 * k;
        int usable = n[0] && n[3] >= 0; /* an intra or unavailable neighbour: (0,0), ref -1 */
        mv[k][0] = usable ? n[1] : 0;
        mv[k][1] = usable ? n[2] : 0;
        ref[k] = usable ? n[3] : -1;
    }
    if (!nb[4] && !nb[8] && nb[0]) { /* B and C unavailable, A available: both take A */
        for (int k = 1; k < 3; k++) {
            mv[k][0] = mv[0][0];
            mv[k][1] = mv[0][1];
            ref[k] = ref[0];
        }
    }
    int matches = (ref[0] == cur_ref) + (ref[1] == cur_ref) + (ref[2] == cur_ref);
    if (matches == 1) {
        int k = ref[0] == cur_ref ? 0 : ref[1] == cur_ref ? 1 : 2;
        out[0] = mv[k][0];
        out[1] = mv[k][1];
    } else {
        out[0] = median3(mv[0][0], mv[1][0], mv[2][0]);
        out[1] = median3(mv[0][1], mv[1][1], mv[2][1]);
    }
}

static int ref_bits_orc(int r, int nrefs) /* mb16.rs:5967-5980 */
{
    if (nrefs <= 1) return 0;
    if (nrefs == 2) return 1;
    unsigned n = (unsigned)r + 1;
    int len = 1;
    while (n > 1) {
        n >>= 1;
        len += 2;
    }
    return len;
}

void orc_motion_search_frame(const orc_search_cfg *cfg, const orc_diamond_cfg *dc, const uint8_t *cur,
                             const uint8_t *const *planes, int nrefs, double lambda_ref, int32_t *ref,
                             int32_t *mv, int64_t *cost)
{
    const int nbx = cfg->cw / 16, nby = cfg->ch / 16;
    /* the frame's rescue learning window (mb16.rs:2868-2886): searches whose grid ran,
     * those that paid, and the latch that turns the rescue off for the frame */
    int resc_n = 0, resc_big = 0, resc_off = 0;
    orc_diamond_cfg d = *dc;
    /* raster macroblock order, as the encoder's frame loop; each macroblock
     * takes its 16x16 best_part result (mb16.rs:4921-4959) */
    for (int y = 0; y < nby; y++)
        for (int x = 0; x < nbx; x++) {
            int32_t nb[12] = {0, 0, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1};
            int ia = y * nbx + x - 1, ib = (y - 1) * nbx + x, ic, cav;
            if (x > 0) { nb[0] = 1; nb[1] = mv[2 * ia]; nb[2] = mv[2 * ia + 1]; nb[3] = ref[ia]; }
            if (y > 0) { nb[4] = 1; nb[5] = mv[2 * ib]; nb[6] = mv[2 * ib + 1]; nb[7] = ref[ib]; }
            if (y > 0 && x + 1 < nbx) { ic = (y - 1) * nbx + x + 1; cav = 1; }  /* C = above-right */
            else { ic = (y - 1) * nbx + x - 1; cav = x > 0 && y > 0; }          /* else D = above-left */
            if (cav) { nb[8] = 1; nb[9] = mv[2 * ic]; nb[10] = mv[2 * ic + 1]; nb[11] = ref[ic]; }
            const int i = y * nbx + x;
            int64_t bc = INT64_MAX;
            int br = 0, bmx = 0, bmy = 0;
            for (int r = 0; r < nrefs; r++) {
                int32_t pr[2], m[2], ev[2] = {0, 0};
                orc_predict_mv(nb, r, pr);
                d.rescue = dc->rescue && !resc_off;
                int64_t c = orc_motion_search_ex(cfg, &d, cur, planes[r], 16 * x, 16 * y, pr, 1, m, ev, NULL, 0,
                                                 NULL);
                if (ev[0] && resc_n < dc->rescue_learn) {
                    resc_n++;
                    if (ev[1]) resc_big++;
                    if (resc_n == dc->rescue_learn && resc_big * 100 < dc->rescue_learn * dc->rescue_payoff)
                        resc_off = 1;
                }
                c += (int64_t)(lambda_ref * (double)ref_bits_orc(r, nrefs));
                if (c < bc) { bc = c; br = r; bmx = m[0]; bmy = m[1]; }
            }
            ref[i] = br;
            mv[2 * i] = bmx;
            mv[2 * i + 1] = bmy;
            cost[i] = bc;
        }
}

/* The P macroblock decision with partitions (mb16.rs:8849-9010, the SATD currency:
 * shape RD off, the all-8x8 P_8x8 arm, every macroblock inter), in raster order.
 * Per macroblock: the neighbours A/B/C(D) read on the 4x4 motion grid of the decided
 * macroblocks (mv_neighbors_block, mb16.rs:1871-1888, at the macroblock's top-left
 * 4x4 with width 4: A = the left macroblock's top-right quadrant, B and C = the
 * bottom-left quadrant of the above and above-right ones, D = the above-left one's
 * bottom-right quadrant); best_part (mb16.rs:4921-4959) of the 16x16 with no extra
 * seeds; when its cost exceeds split_gate, best_part of the two 16x8 halves, the two
 * 8x16 halves and (p8x8) the four 8x8 quadrants, each with the macroblock's
 * neighbours and the 16x16 vector as the extra seed (mb16.rs:8893-8948); the pick
 * is the first strict minimum of c16, ct + cb, cl + cr and (int64)(lme * 4) + the
 * quadrants' sum (mb16.rs:8897-8948).  Costs include (int64)(lme * ref_bits(r)).
 * rec per macroblock, 14 int32: mode (0 16x16, 1 16x8, 2 8x16, 3 8x8), cost, then
 * per quadrant (TL, TR, BL, BR) ref, mvx, mvy. */
static void pd_quad_nb(const int32_t *rec, int nbx, int x, int y, int q, int32_t out[4])
{
    const int32_t *r = rec + 14 * (size_t)(y * nbx + x) + 2 + 3 * q;
    out[0] = 1;
    out[1] = r[1];
    out[2] = r[2];
    out[3] = r[0];
}

/* the frame's rescue learning window (mb16.rs:2868-2886) over its searches in call order */
typedef struct {
    int n, big, off;
} resc_window;

static void window_tally(resc_window *w, const orc_diamond_cfg *dc, const int32_t ev[2])
{
    if (ev[0] && w->n < dc->rescue_learn) {
        w->n++;
        if (ev[1]) w->big++;
        if (w->n == dc->rescue_learn && w->big * 100 < dc->rescue_learn * dc->rescue_payoff) w->off = 1;
    }
}

static int64_t pd_best_part(const orc_search_cfg *cfg0, const orc_diamond_cfg *dc, const uint8_t *cur,
                            const uint8_t *const *planes, int nrefs, double lme, const int32_t nb[12], int lx, int ly,
                            int bw, int bh, const int32_t *extra, int nextra, int32_t *ref_out, int32_t *mv_out,
                            resc_window *w)
{
    orc_search_cfg cfg = *cfg0;
    cfg.bw = bw;
    cfg.bh = bh;
    orc_diamond_cfg d = *dc;
    int64_t bc = INT64_MAX;
    for (int r = 0; r < nrefs; r++) {
        int32_t seeds[8], m[2], ev[2] = {0, 0};
        orc_predict_mv(nb, r, seeds);
        for (int e = 0; e < nextra; e++) {
            seeds[2 + 2 * e] = extra[2 * e];
            seeds[3 + 2 * e] = extra[2 * e + 1];
        }
        d.rescue = dc->rescue && !w->off;
        int64_t c = orc_motion_search_ex(&cfg, &d, cur, planes[r], lx, ly, seeds, 1 + nextra, m, ev, NULL, 0, NULL);
        window_tally(w, dc, ev);
        c += (int64_t)(lme * (double)ref_bits_orc(r, nrefs));
        if (c < bc) { bc = c; *ref_out = r; mv_out[0] = m[0]; mv_out[1] = m[1]; }
    }
    return bc;
}

void orc_partition_decision_frame(const orc_search_cfg *cfg, const orc_diamond_cfg *dc, const uint8_t *cur,
                                  const uint8_t *const *planes, int nrefs, double lme, int64_t split_gate, int p8x8,
                                  int32_t *rec)
{
    const int nbx = cfg->cw / 16, nby = cfg->ch / 16;
    resc_window w = {0, 0, 0};
    for (int y = 0; y < nby; y++)
        for (int x = 0; x < nbx; x++) {
            int32_t nb[12] = {0, 0, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1};
            if (x > 0) pd_quad_nb(rec, nbx, x - 1, y, 1, nb);
            if (y > 0) pd_quad_nb(rec, nbx, x, y - 1, 2, nb + 4);
            if (y > 0 && x + 1 < nbx) pd_quad_nb(rec, nbx, x + 1, y - 1, 2, nb + 8);
            else if (x > 0 && y > 0) pd_quad_nb(rec, nbx, x - 1, y - 1, 3, nb + 8);
            const int lx = 16 * x, ly = 16 * y;
            int32_t r16 = 0, mv16[2] = {0, 0};
            const int64_t c16 =
                pd_best_part(cfg, dc, cur, planes, nrefs, lme, nb, lx, ly, 16, 16, NULL, 0, &r16, mv16, &w);
            int mode = 0;
            int64_t best = c16;
            int32_t qr[4] = {r16, r16, r16, r16}, qm[4][2];
            for (int q = 0; q < 4; q++) { qm[q][0] = mv16[0]; qm[q][1] = mv16[1]; }
            if (c16 > split_gate) {
                int32_t pr[8], pm[8][2];
                int64_t pc[8];
                /* 16x8 top, bottom; 8x16 left, right; 8x8 TL, TR, BL, BR */
                static const int geo[8][4] = {{0, 0, 16, 8}, {0, 8, 16, 8}, {0, 0, 8, 16}, {8, 0, 8, 16},
                                              {0, 0, 8, 8},  {8, 0, 8, 8},  {0, 8, 8, 8},  {8, 8, 8, 8}};
                const int nparts = p8x8 ? 8 : 4;
                for (int k = 0; k < nparts; k++)
                    pc[k] = pd_best_part(cfg, dc, cur, planes, nrefs, lme, nb, lx + geo[k][0], ly + geo[k][1],
                                         geo[k][2], geo[k][3], mv16, 1, &pr[k], pm[k], &w);
                if (pc[0] + pc[1] < best) {
                    best = pc[0] + pc[1];
                    mode = 1;
                    const int map[4] = {0, 0, 1, 1};
                    for (int q = 0; q < 4; q++) { qr[q] = pr[map[q]]; qm[q][0] = pm[map[q]][0]; qm[q][1] = pm[map[q]][1]; }
                }
                if (pc[2] + pc[3] < best) {
                    best = pc[2] + pc[3];
                    mode = 2;
                    const int map[4] = {2, 3, 2, 3};
                    for (int q = 0; q < 4; q++) { qr[q] = pr[map[q]]; qm[q][0] = pm[map[q]][0]; qm[q][1] = pm[map[q]][1]; }
                }
                if (p8x8) {
                    const int64_t c8 = (int64_t)(lme * 4.0) + pc[4] + pc[5] + pc[6] + pc[7];
                    if (c8 < best) {
                        best = c8;
                        mode = 3;
                        for (int q = 0; q < 4; q++) { qr[q] = pr[4 + q]; qm[q][0] = pm[4 + q][0]; qm[q][1] = pm[4 + q][1]; }
                    }
                }
            }
            int32_t *o = rec + 14 * (size_t)(y * nbx + x);
            o[0] = mode;
            o[1] = (int32_t)best;
            for (int q = 0; q < 4; q++) {
                o[2 + 3 * q] = qr[q];
                o[3 + 3 * q] = qm[q][0];
                o[4 + 3 * q] = qm[q][1];
            }
        }
}

/* The split gate of mb16.rs:8866-8868 for a QP: ((30 * (qstep16 + 160)) >> 3) * 2,
 * qstep16 = {10, 11, 13, 14, 16, 18}[qp % 6] << (qp / 6). */
int64_t orc_split_gate(int qp)
{
    static const int64_t qs[6] = {10, 11, 13, 14, 16, 18};
    const int64_t q = qs[qp % 6] << (qp / 6);
    return ((30 * (q + 160)) >> 3) * 2;
}

/* skip_mv (mb16.rs:1841-1853) from the neighbours in predict_mv's layout. */
void orc_skip_mv(const int32_t nb[12], int32_t out[2])
{
    const int32_t *a = nb, *b = nb + 4;
    if (!a[0] || !b[0] || (a[3] == 0 && a[1] == 0 && a[2] == 0) || (b[3] == 0 && b[1] == 0 && b[2] == 0)) {
        out[0] = out[1] = 0;
        return;
    }
    orc_predict_mv(nb, 0, out);
}

void orc_decision_predict(int cw, int ch, int nrefs, const uint8_t *const *refs, const int32_t *ref,
                          const int32_t *mv, int32_t *skip_mv, uint8_t *pred, uint8_t *skip_pred)
{
    const int nbx = cw / 16, nby = ch / 16;
    uint8_t blk[256];
    for (int y = 0; y < nby; y++)
        for (int x = 0; x < nbx; x++) {
            const int i = y * nbx + x;
            int32_t nb[12] = {0, 0, 0, -1, 0, 0, 0, -1, 0, 0, 0, -1};
            int ia = i - 1, ib = i - nbx, ic, cav;
            if (x > 0) { nb[0] = 1; nb[1] = mv[2 * ia]; nb[2] = mv[2 * ia + 1]; nb[3] = ref[ia]; }
            if (y > 0) { nb[4] = 1; nb[5] = mv[2 * ib]; nb[6] = mv[2 * ib + 1]; nb[7] = ref[ib]; }
            if (y > 0 && x + 1 < nbx) { ic = ib + 1; cav = 1; }
            else { ic = ib - 1; cav = x > 0 && y > 0; }
            if (cav) { nb[8] = 1; nb[9] = mv[2 * ic]; nb[10] = mv[2 * ic + 1]; nb[11] = ref[ic]; }
            orc_skip_mv(nb, skip_mv + 2 * i);
            int r = ref[i] < 0 ? 0 : ref[i] >= nrefs ? nrefs - 1 : ref[i];
            orc_mc_luma(refs[r], cw, ch, 16 * x, 16 * y, 16, 16, mv[2 * i], mv[2 * i + 1], blk);
            for (int k = 0; k < 16; k++) memcpy(pred + (size_t)(16 * y + k) * cw + 16 * x, blk + 16 * k, 16);
            orc_mc_luma(refs[0], cw, ch, 16 * x, 16 * y, 16, 16, skip_mv[2 * i], skip_mv[2 * i + 1], blk);
            for (int k = 0; k < 16; k++) memcpy(skip_pred + (size_t)(16 * y + k) * cw + 16 * x, blk + 16 * k, 16);
        }
}

/* ---------------------------------------------------------------------------
 * mb-tree lookahead
 * ------------------------------------------------------------------------- */

/* downsample2x, mbtree.rs:61-74: (a+b+c+d+2)/4 */
void orc_downsample2x(const uint8_t *y, int cw, int ch, uint8_t *out)
{
    int hw = cw / 2, hh = ch / 2;
    for (int j = 0; j < hh; j++)
        for (int i = 0; i < hw; i++) {
            unsigned s = (unsigned)y[(size_t)2 * j * cw + 2 * i] + y[(size_t)2 * j * cw + 2 * i + 1] +
                         y[(size_t)(2 * j + 1) * cw + 2 * i] + y[(size_t)(2 * j + 1) * cw + 2 * i + 1];
            out[(size_t)j * hw + i] = (uint8_t)((s + 2) / 4);
        }
}

/* intra_cost, mbtree.rs:78-93: AC SATD (coefficient 0 excluded), min i32::MAX, max 1 */
int32_t orc_intra_cost(const uint8_t *sy, int cw, int bx0, int by0, int bs)
{
    int64_t s = 0;
    for (int by = 0; by < bs / 4; by++)
        for (int bx = 0; bx < bs / 4; bx++) {
            int32_t blk[16], h[16];
            for (int dy = 0; dy < 4; dy++)
                for (int dx = 0; dx < 4; dx++)
                    blk[dy * 4 + dx] = sy[(size_t)(by0 + by * 4 + dy) * cw + bx0 + bx * 4 + dx];
            orc_hadamard_4x4(blk, h);
            for (int k = 1; k < 16; k++) s += iabs(h[k]);
        }
    if (s > INT32_MAX) s = INT32_MAX;
    return s < 1 ? 1 : (int32_t)s;
}

/* mc_satd, mbtree.rs:102-148: interior full-pel satd_px, else mc_luma + 4x4 SATD */
int64_t orc_mbtree_mc_satd(const uint8_t *sy, int cw, int ch, const uint8_t *ref, int bx0,
                           int by0, int bs, int mvx, int mvy)
{
    uint8_t pred[256];
    orc_mc_luma(ref, cw, ch, bx0, by0, bs, bs, mvx, mvy, pred);
    return orc_satd_px(sy + (size_t)by0 * cw + bx0, cw, pred, bs, bs, bs);
}

static int64_t mcs_noted(const uint8_t *sy, int cw, int ch, const uint8_t *ref, int bx0, int by0, int bs, int x,
                         int y, int64_t *trace, int cap, int *ne)
{
    int64_t v = orc_mbtree_mc_satd(sy, cw, ch, ref, bx0, by0, bs, x, y);
    if (trace && *ne < cap) {
        trace[3 * (size_t)*ne] = x;
        trace[3 * (size_t)*ne + 1] = y;
        trace[3 * (size_t)*ne + 2] = v;
    }
    (*ne)++;
    return v;
}

/* inter_cost, mbtree.rs:154-197, with an evaluation trace: up to cap entries {mvx, mvy, satd} (int64),
 * *n = the number of mc_satd evaluations. */
int32_t orc_inter_cost_ex(const uint8_t *sy, int cw, int ch, const uint8_t *ref, int bx0, int by0, int bs,
                          int seedx, int seedy, int max_step, int32_t *mv_out, int64_t *trace, int cap, int *n)
{
    int ne = 0;
#define MCS(X, Y) mcs_noted(sy, cw, ch, ref, bx0, by0, bs, (X), (Y), trace, cap, &ne)
    int bmx = 0, bmy = 0;
    int64_t best = MCS(0, 0);
    if (best != 0) {
        if (seedx != 0 || seedy != 0) {
            int64_t s = MCS(seedx, seedy);
            if (s < best) { best = s; bmx = seedx; bmy = seedy; }
        }
        for (int step = max_step; step >= 1; step >>= 1) {
            for (;;) {
                int moved = 0;
                const int ring[4][2] = {{step, 0}, {-step, 0}, {0, step}, {0, -step}};
                for (int i = 0; i < 4; i++) {
                    int x = bmx + ring[i][0] * 4, y = bmy + ring[i][1] * 4;
                    int64_t s = MCS(x, y);
                    if (s < best) { best = s; bmx = x; bmy = y; moved = 1; }
                }
                if (!moved) break;
            }
        }
    }
#undef MCS
    mv_out[0] = bmx;
    mv_out[1] = bmy;
    if (n) *n = ne;
    return best > INT32_MAX ? INT32_MAX : (int32_t)best;
}

int32_t orc_inter_cost(const uint8_t *sy, int cw, int ch, const uint8_t *ref, int bx0, int by0,
                       int bs, int seedx, int seedy, int max_step, int32_t *mv_out)
{
    return orc_inter_cost_ex(sy, cw, ch, ref, bx0, by0, bs, seedx, seedy, max_step, mv_out, NULL, 0, NULL);
}

/* lookahead::complexity, lookahead.rs:38-101: the rate-control score of a coded-size
 * frame.  IDR (ref NULL): sum over macroblocks of intra_activity (55-72), the per-4x4
 * AC SATD (coefficient 0 excluded).  P: sum over macroblocks of inter_activity
 * (76-101), the minimum over nine full-pel candidates of the residual's 4x4 SATD
 * against mc_luma of the reference.  Returned as max(total, 1) in f64. */
double orc_rc_complexity(const uint8_t *sy, const uint8_t *ref, int cw, int ch)
{
    static const int cands[9][2] = {{0, 0}, {4, 0}, {-4, 0}, {0, 4}, {0, -4}, {8, 0}, {-8, 0}, {0, 8}, {0, -8}};
    int64_t total = 0;
    for (int my = 0; my < ch / 16; my++)
        for (int mx = 0; mx < cw / 16; mx++) {
            if (!ref) {
                int64_t s = 0;
                for (int by = 0; by < 4; by++)
                    for (int bx = 0; bx < 4; bx++) {
                        int32_t blk[16], h[16];
                        for (int dy = 0; dy < 4; dy++)
                            for (int dx = 0; dx < 4; dx++)
                                blk[dy * 4 + dx] = sy[(size_t)(my * 16 + by * 4 + dy) * cw + mx * 16 + bx * 4 + dx];
                        orc_hadamard_4x4(blk, h);
                        for (int k = 1; k < 16; k++) s += iabs(h[k]);
                    }
                total += s;
            } else {
                int64_t best = INT64_MAX;
                for (int c = 0; c < 9; c++) {
                    uint8_t pred[256];
                    orc_mc_luma(ref, cw, ch, mx * 16, my * 16, 16, 16, cands[c][0], cands[c][1], pred);
                    int64_t s = orc_satd_px(sy + (size_t)my * 16 * cw + mx * 16, cw, pred, 16, 16, 16);
                    if (s < best) best = s;
                }
                total += best;
            }
        }
    return total < 1 ? 1.0 : (double)total;
}

/* Build-defined multi-reference lookahead (the 4K configuration's 3 refs, +-64 form;
 * no reference counterpart, whose inter_cost, mbtree.rs:154-197, walks a seeded
 * diamond against one reference).  One 8x8 block of frame f at (bx0, by0):
 * for k = 1..min(nrefs, f): exhaustive full-pel SAD over +-range against frame f-k
 * (reads clamped into the frame; raster first minimum, dy outer), then the 8x8 SATD
 * at the winner; the cheapest reference wins, ties to the smaller k. */
static int32_t la_intra8(const uint8_t *y, int cw, int bx0, int by0)
{
    int32_t d[64], co[64];
    for (int j = 0; j < 8; j++)
        for (int i = 0; i < 8; i++) d[j * 8 + i] = y[(size_t)(by0 + j) * cw + bx0 + i];
    orc_hadamard_8x8(d, co);
    int64_t s = 0;
    for (int k = 1; k < 64; k++) s += iabs(co[k]);
    s >>= 1;
    return s < 1 ? 1 : (int32_t)s;
}

/* SAD of the 8x8 block at (bx0, by0) of cur against ref displaced by (dx, dy), reads
 * clamped into the frame */
static int64_t la_sad8(const uint8_t *cur, const uint8_t *ref, int cw, int ch, int bx0, int by0, int dx, int dy)
{
    int64_t sad = 0;
    for (int j = 0; j < 8; j++)
        for (int i = 0; i < 8; i++)
            sad += iabs((int)cur[(size_t)(by0 + j) * cw + bx0 + i] - at(ref, cw, ch, bx0 + i + dx, by0 + j + dy));
    return sad;
}

/* Exhaustive full-pel search of +-r around (cx, cy), raster first minimum */
static int64_t la_window8(const uint8_t *cur, const uint8_t *ref, int cw, int ch, int bx0, int by0, int cx, int cy,
                          int r, int *wx, int *wy)
{
    int64_t bs = INT64_MAX;
    for (int dy = -r; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++) {
            int64_t sad = la_sad8(cur, ref, cw, ch, bx0, by0, cx + dx, cy + dy);
            if (sad < bs) { bs = sad; *wx = cx + dx; *wy = cy + dy; }
        }
    return bs;
}

/* The learned-seeded form (build-defined): per reference k the 8 centres
 * cen[(k-1)*16 + 2c .. +1] (half-res full pels, the network's proposals in rank order)
 * are each searched over +-local_range; the first minimum in candidate order (then in
 * raster order) wins; a best SAD above fb_sad falls back to the exhaustive +-range
 * search (bit k-1 of *fb_mask).  Then as orc_la_block_multi. */
int64_t orc_la_block_learned(const uint8_t *half, int cw, int ch, int f, int nrefs, int range, int local_range,
                             int fb_sad, const int32_t *cen, int bx0, int by0, int32_t *mv_out, int32_t *ref_out,
                             int32_t *fb_mask)
{
    const size_t fs = (size_t)cw * ch;
    const uint8_t *cur = half + (size_t)f * fs;
    int64_t best = INT64_MAX;
    int bmx = 0, bmy = 0, bk = 0;
    *fb_mask = 0;
    for (int k = 1; k <= nrefs && k <= f; k++) {
        const uint8_t *ref = half + (size_t)(f - k) * fs;
        int64_t bs = INT64_MAX;
        int wx = 0, wy = 0;
        for (int c = 0; c < 8; c++) {
            int cx2 = 0, cy2 = 0;
            int64_t s2 = la_window8(cur, ref, cw, ch, bx0, by0, cen[(k - 1) * 16 + 2 * c], cen[(k - 1) * 16 + 2 * c + 1],
                                    local_range, &cx2, &cy2);
            if (s2 < bs) { bs = s2; wx = cx2; wy = cy2; }
        }
        if (bs > fb_sad) {
            la_window8(cur, ref, cw, ch, bx0, by0, 0, 0, range, &wx, &wy);
            *fb_mask |= 1 << (k - 1);
        }
        uint8_t pred[64];
        for (int j = 0; j < 8; j++)
            for (int i = 0; i < 8; i++) pred[j * 8 + i] = (uint8_t)at(ref, cw, ch, bx0 + i + wx, by0 + j + wy);
        int64_t c = orc_satd8_px(cur + (size_t)by0 * cw + bx0, cw, pred, 8, 8, 8);
        if (c < best) { best = c; bmx = wx; bmy = wy; bk = k; }
    }
    mv_out[0] = 8 * bmx;
    mv_out[1] = 8 * bmy;
    *ref_out = bk;
    return best;
}

int64_t orc_la_block_multi(const uint8_t *half, int cw, int ch, int f, int nrefs, int range, int bx0, int by0,
                           int32_t *mv_out, int32_t *ref_out)
{
    const size_t fs = (size_t)cw * ch;
    const uint8_t *cur = half + (size_t)f * fs;
    int64_t best = INT64_MAX;
    int bmx = 0, bmy = 0, bk = 0;
    for (int k = 1; k <= nrefs && k <= f; k++) {
        const uint8_t *ref = half + (size_t)(f - k) * fs;
        int wx = 0, wy = 0;
        la_window8(cur, ref, cw, ch, bx0, by0, 0, 0, range, &wx, &wy);
        uint8_t pred[64];
        for (int j = 0; j < 8; j++)
            for (int i = 0; i < 8; i++) pred[j * 8 + i] = (uint8_t)at(ref, cw, ch, bx0 + i + wx, by0 + j + wy);
        int64_t c = orc_satd8_px(cur + (size_t)by0 * cw + bx0, cw, pred, 8, 8, 8);
        if (c < best) { best = c; bmx = wx; bmy = wy; bk = k; }
    }
    mv_out[0] = 8 * bmx;
    mv_out[1] = 8 * bmy;
    *ref_out = bk;
    return best;
}

void orc_mbtree_costs_multi(const uint8_t *half, int nframes, int mb_w, int mb_h, int nrefs, int range,
                            int32_t *intra, int32_t *inter, int32_t *mv, int32_t *ref, int nthreads)
{
    const int cw = mb_w * 8, ch = mb_h * 8, nm = mb_w * mb_h;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) collapse(2)
#else
    (void)nthreads;
#endif
    for (int f = 0; f < nframes; f++)
        for (int m = 0; m < nm; m++) {
            const size_t g = (size_t)f * nm + m;
            const int bx0 = (m % mb_w) * 8, by0 = (m / mb_w) * 8;
            const int32_t ic = la_intra8(half + (size_t)f * cw * ch, cw, bx0, by0);
            intra[g] = ic;
            if (f == 0) {
                inter[g] = ic;
                mv[2 * g] = mv[2 * g + 1] = 0;
                ref[g] = 0;
                continue;
            }
            int32_t m2[2], k;
            int64_t c = orc_la_block_multi(half, cw, ch, f, nrefs, range, bx0, by0, m2, &k);
            inter[g] = c < ic ? (int32_t)c : ic;
            mv[2 * g] = m2[0];
            mv[2 * g + 1] = m2[1];
            ref[g] = k;
        }
}

/* Multi-reference int32 propagation (build-defined): as orc_propagate_i32, with
 * frame f's amount going to frame f - ref[f][mb]; a frame's accumulator is final
 * once every later frame has been processed. */
void orc_propagate_multi_i32(int n, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                             const int32_t *mv, const int32_t *ref, int32_t *prop)
{
    int nm = mb_w * mb_h;
    int64_t *acc = (int64_t *)calloc((size_t)n * nm, sizeof(int64_t));
    for (int f = n - 1; f >= 0; f--) {
        for (int m = 0; m < nm; m++) {
            int64_t a = acc[(size_t)f * nm + m];
            prop[(size_t)f * nm + m] = a > INT32_MAX ? INT32_MAX : (int32_t)a;
        }
        if (f == 0) break;
        for (int y = 0; y < mb_h; y++)
            for (int x = 0; x < mb_w; x++) {
                int m = y * mb_w + x;
                size_t g = (size_t)f * nm + m;
                int64_t total = (int64_t)intra[g] + prop[g];
                int64_t amount = total * (int64_t)(intra[g] - inter[g]) / intra[g];
                if (amount <= 0) continue;
                int t = f - (ref ? ref[g] : 1);
                int rx = clampi(x * 16 + asr2(mv[2 * g]), 0, (mb_w - 1) * 16);
                int ry = clampi(y * 16 + asr2(mv[2 * g + 1]), 0, (mb_h - 1) * 16);
                int cx0 = rx / 16, cy0 = ry / 16, fx = rx % 16, fy = ry % 16;
                const int wys[2] = {16 - fy, fy}, wxs[2] = {16 - fx, fx};
                for (int dy = 0; dy < 2; dy++) {
                    if (wys[dy] <= 0) continue;
                    int cy = cy0 + dy < mb_h - 1 ? cy0 + dy : mb_h - 1;
                    for (int dx = 0; dx < 2; dx++) {
                        if (wxs[dx] <= 0) continue;
                        int cx = cx0 + dx < mb_w - 1 ? cx0 + dx : mb_w - 1;
                        acc[(size_t)t * nm + cy * mb_w + cx] += (amount * wxs[dx] * wys[dy]) >> 8;
                    }
                }
            }
    }
    free(acc);
}

/* Rust i32 `/` truncates toward zero (C99 does too). */
void orc_frame_costs_half(const uint8_t *half, const uint8_t *ref_half, int mb_w, int mb_h,
                          int32_t *intra, int32_t *inter, int32_t *mv)
{
    int cwh = mb_w * 8, chh = mb_h * 8;
    for (int y = 0; y < mb_h; y++)
        for (int x = 0; x < mb_w; x++) {
            int m = y * mb_w + x;
            int sfx = 0, sfy = 0;
            if (x > 0) { sfx = mv[2 * (m - 1)]; sfy = mv[2 * (m - 1) + 1]; }
            else if (y > 0) { sfx = mv[2 * (m - mb_w)]; sfy = mv[2 * (m - mb_w) + 1]; }
            int32_t ic = orc_intra_cost(half, cwh, x * 8, y * 8, 8);
            intra[m] = ic;
            if (ref_half) {
                int32_t mvp[2];
                int32_t c = orc_inter_cost(half, cwh, chh, ref_half, x * 8, y * 8, 8, sfx / 2,
                                           sfy / 2, 8, mvp);
                inter[m] = c < ic ? c : ic;
                mv[2 * m] = mvp[0] * 2;
                mv[2 * m + 1] = mvp[1] * 2;
            } else {
                inter[m] = ic;
                mv[2 * m] = 0;
                mv[2 * m + 1] = 0;
            }
        }
}

/* propagate_to, mbtree.rs:268-297 */
static void propagate_to(double *prev, int mb_w, int mb_h, int mb_x, int mb_y, int mvx, int mvy,
                         double amount)
{
    if (amount <= 0.0) return;
    int rx = clampi(mb_x * 16 + asr2(mvx), 0, (mb_w - 1) * 16);
    int ry = clampi(mb_y * 16 + asr2(mvy), 0, (mb_h - 1) * 16);
    int cx0 = rx / 16, cy0 = ry / 16;
    double fx = (double)(rx % 16), fy = (double)(ry % 16);
    double wl = 16.0 - fx, wt = 16.0 - fy;
    const double wys[2] = {wt, fy}, wxs[2] = {wl, fx};
    for (int dy = 0; dy < 2; dy++) {
        if (wys[dy] <= 0.0) continue;
        int cy = cy0 + dy < mb_h - 1 ? cy0 + dy : mb_h - 1;
        for (int dx = 0; dx < 2; dx++) {
            if (wxs[dx] <= 0.0) continue;
            int cx = cx0 + dx < mb_w - 1 ? cx0 + dx : mb_w - 1;
            prev[cy * mb_w + cx] += amount * (wxs[dx] * wys[dy]) / 256.0;
        }
    }
}

/* backward propagation, mbtree.rs:440-455 */
void orc_propagate_f64(int n, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                       const int32_t *mv, double *prop)
{
    int nm = mb_w * mb_h;
    memset(prop, 0, sizeof(double) * (size_t)n * nm);
    for (int f = n - 1; f >= 1; f--) {
        const double *cur = prop + (size_t)f * nm;
        double *prev = prop + (size_t)(f - 1) * nm;
        for (int y = 0; y < mb_h; y++)
            for (int x = 0; x < mb_w; x++) {
                int m = y * mb_w + x;
                size_t g = (size_t)f * nm + m;
                double total = (double)intra[g] + cur[m];
                double frac = (double)(intra[g] - inter[g]) / (double)intra[g];
                propagate_to(prev, mb_w, mb_h, x, y, mv[2 * g], mv[2 * g + 1], total * frac);
            }
    }
}

/* ---- the grain signature (signals.rs:528-533) of frame cur against ref -------
 * median_var: frame_median_mb_var (signals.rs:247-266), per macroblock the variance
 *   of its 64 samples at even rows and columns, (sq - sum*sum/64)/64 in int64, the
 *   median (sorted, element n/2);
 * floor: grain_floor (signals.rs:311-338), the zero-motion SAD per pixel of ~48
 *   interior macroblocks (the interior walk of mbw-4 x mbh-4 from (2,2), stride
 *   max(inner/48, 1)), the element len/4 of the sorted values; 0 below 6x6 MBs;
 * mgain: b2_mgain (signals.rs:61-129), over ~24 interior macroblocks (stride
 *   max(inner/24, 1)) the mean of (SAD0 - min SAD over the +-8 step-4 full-pel grid
 *   inside the frame) / (SAD0 + 1); 0 below 6x6 MBs;
 * grain = median_var < 200 && floor > 5 && mgain < 0.1.  stats: the three values. */
static uint32_t sad16_at(const uint8_t *cur, const uint8_t *ref, int cw, int bx, int by, int rx, int ry)
{
    uint32_t s = 0;
    for (int dy = 0; dy < 16; dy++)
        for (int dx = 0; dx < 16; dx++)
            s += (uint32_t)iabs((int)cur[(size_t)(by + dy) * cw + bx + dx] - (int)ref[(size_t)(ry + dy) * cw + rx + dx]);
    return s;
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

static int cmp_f64(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int orc_grain_signature(const uint8_t *cur, const uint8_t *ref, int cw, int ch, double *stats)
{
    const int mbw = cw / 16, mbh = ch / 16;
    int64_t *vs = (int64_t *)malloc(sizeof(int64_t) * (size_t)(mbw * mbh > 0 ? mbw * mbh : 1));
    int nv = 0;
    for (int my = 0; my < mbh; my++)
        for (int mx = 0; mx < mbw; mx++) {
            int64_t sum = 0, sq = 0;
            for (int r = 0; r < 16; r += 2)
                for (int c = 0; c < 16; c += 2) {
                    int64_t v = cur[(size_t)(my * 16 + r) * cw + mx * 16 + c];
                    sum += v;
                    sq += v * v;
                }
            vs[nv++] = (sq - sum * sum / 64) / 64;
        }
    qsort(vs, (size_t)nv, sizeof(int64_t), cmp_i64);
    const int64_t median_var = nv ? vs[nv / 2] : 0;
    free(vs);
    double floor_v = 0.0, mgain = 0.0;
    if (mbw >= 6 && mbh >= 6) {
        const int inner = (mbw - 4) * (mbh - 4);
        /* grain_floor */
        int stride = inner / 48 > 1 ? inner / 48 : 1;
        double *fl = (double *)malloc(sizeof(double) * (size_t)(inner / stride + 2));
        int nf = 0;
        for (int i = 0; i < inner; i += stride) {
            const int bx = (2 + i % (mbw - 4)) * 16, by = (2 + i / (mbw - 4)) * 16;
            fl[nf++] = (double)sad16_at(cur, ref, cw, bx, by, bx, by) / 256.0;
        }
        qsort(fl, (size_t)nf, sizeof(double), cmp_f64);
        floor_v = nf ? fl[nf / 4] : 0.0;
        free(fl);
        /* b2_mgain */
        stride = inner / 24 > 1 ? inner / 24 : 1;
        double acc = 0.0;
        unsigned n = 0;
        for (int i = 0; i < inner; i += stride) {
            const int bx = (2 + i % (mbw - 4)) * 16, by = (2 + i / (mbw - 4)) * 16;
            const uint32_t s0 = sad16_at(cur, ref, cw, bx, by, bx, by);
            uint32_t best = s0;
            for (int dy = -8; dy <= 8; dy += 4)
                for (int dx = -8; dx <= 8; dx += 4) {
                    const int rx = bx + dx, ry = by + dy;
                    if (rx < 0 || ry < 0 || rx + 16 > cw || ry + 16 > ch) continue;
                    const uint32_t sv = sad16_at(cur, ref, cw, bx, by, rx, ry);
                    if (sv < best) best = sv;
                }
            acc += (double)(s0 - best) / (double)(s0 + 1);
            n++;
        }
        mgain = n ? acc / (double)n : 0.0;
    }
    if (stats) {
        stats[0] = (double)median_var;
        stats[1] = floor_v;
        stats[2] = mgain;
    }
    return median_var < 200 && floor_v > 5.0 && mgain < 0.1;
}

/* gop_qp_offsets' policy (mbtree.rs:369-593) over given frame costs: the grain latch
 * (mbtree.rs:401-420: with grain set and n >= 2 every offset is 0), f64 propagation,
 * back-off latch, centred -s*log2 offsets, spread latch, round/clamp. */
void orc_mbtree_qp_offsets(int n, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                           const int32_t *mv, double strength, double res_min, double sd_min, int grain, int32_t *off,
                           double *stats)
{
    size_t nm = (size_t)mb_w * mb_h, N = (size_t)n * nm;
    double resid = 0.0, sd = 0.0, eff = 0.0;
    memset(off, 0, sizeof(int32_t) * N);
    if (strength > 0.0 && N > 0 && !(grain && n >= 2)) {
        double *prop = (double *)malloc(sizeof(double) * N), *o = (double *)malloc(sizeof(double) * N);
        orc_propagate_f64(n, mb_w, mb_h, intra, inter, mv, prop);
        double fsum = 0.0, fc = 0.0;
        for (int f = 1; f < n; f++)
            for (size_t m = 0; m < nm; m++) {
                size_t g = (size_t)f * nm + m;
                fsum += (double)(intra[g] - inter[g]) / (double)intra[g];
                fc += 1.0;
            }
        resid = 1.0 - (fc > 0.0 ? fsum / fc : 0.0);
        eff = (res_min > 0.0 && resid < res_min) ? 0.0 : strength;
        if (eff != 0.0) {
            double sum = 0.0, sq = 0.0;
            for (size_t g = 0; g < N; g++) {
                o[g] = -eff * log2(((double)intra[g] + prop[g]) / (double)intra[g]);
                sum += o[g];
            }
            double mean = sum / (double)N;
            for (size_t g = 0; g < N; g++) o[g] -= mean;
            for (size_t g = 0; g < N; g++) sq += o[g] * o[g];
            sd = sqrt(sq / (double)N) / (eff > 1e-9 ? eff : 1e-9);
            if (!(sd < sd_min))
                for (size_t g = 0; g < N; g++) {
                    double r = round(o[g]);
                    off[g] = (int32_t)(r < -6.0 ? -6.0 : r > 6.0 ? 6.0 : r);
                }
        }
        free(prop);
        free(o);
    }
    if (stats) { stats[0] = resid; stats[1] = sd; stats[2] = eff; }
}

/* Build-defined int32 propagation: same geometry as propagate_to, integer amounts.
 * amount = ((intra + prop) * (intra - inter)) / intra (int64, floor; all terms >= 0),
 * share = (amount * wx * wy) >> 8, accumulated in int64 and saturated to INT32_MAX. */
void orc_propagate_i32(int n, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                       const int32_t *mv, int32_t *prop)
{
    int nm = mb_w * mb_h;
    int64_t *acc = (int64_t *)calloc((size_t)nm, sizeof(int64_t));
    memset(prop, 0, sizeof(int32_t) * (size_t)n * nm);
    for (int f = n - 1; f >= 1; f--) {
        memset(acc, 0, sizeof(int64_t) * (size_t)nm);
        for (int y = 0; y < mb_h; y++)
            for (int x = 0; x < mb_w; x++) {
                int m = y * mb_w + x;
                size_t g = (size_t)f * nm + m;
                int64_t total = (int64_t)intra[g] + prop[g];
                int64_t amount = total * (int64_t)(intra[g] - inter[g]) / intra[g];
                if (amount <= 0) continue;
                int rx = clampi(x * 16 + asr2(mv[2 * g]), 0, (mb_w - 1) * 16);
                int ry = clampi(y * 16 + asr2(mv[2 * g + 1]), 0, (mb_h - 1) * 16);
                int cx0 = rx / 16, cy0 = ry / 16, fx = rx % 16, fy = ry % 16;
                const int wys[2] = {16 - fy, fy}, wxs[2] = {16 - fx, fx};
                for (int dy = 0; dy < 2; dy++) {
                    if (wys[dy] <= 0) continue;
                    int cy = cy0 + dy < mb_h - 1 ? cy0 + dy : mb_h - 1;
                    for (int dx = 0; dx < 2; dx++) {
                        if (wxs[dx] <= 0) continue;
                        int cx = cx0 + dx < mb_w - 1 ? cx0 + dx : mb_w - 1;
                        acc[cy * mb_w + cx] += (amount * wxs[dx] * wys[dy]) >> 8;
                    }
                }
            }
        for (int m = 0; m < nm; m++)
            prop[(size_t)(f - 1) * nm + m] = acc[m] > INT32_MAX ? INT32_MAX : (int32_t)acc[m];
    }
    free(acc);
}
