// rh264_accel.hip -- the per-block family of the drop-in boundary: the functions of
// crates/rusty_h264-accel with their names, argument order and value domains, served
// by the HIP kernels behind the batched entry points of rh264_me.hip.
//
// Host code only.  Every call stages its host blocks into device memory owned by the
// calling thread's default context, runs one batched launch (rh264_block_metric or
// rh264_build_planes) and copies the values back; the call is synchronous.  The
// reference calls these from GOP-parallel threads (crates/rusty_h264-encoder/src/
// lib.rs:746-800), so each host thread gets its own context on the device chosen by
// rh264_accel_set_device (default 0); rh264_accel_release frees the calling thread's.
//
// Status: 0 = done, 1 = declined (where the reference returns None / false: an
// uncovered shape or geometry), < 0 = error (rh264_last_error).

#include <stdint.h>
#include <string.h>

#include <vector>

#include "rh264_internal.h"
#include "rh264_me.h"

namespace {

int g_accel_device = 0;

// The calling thread's context and its growing device staging buffer.
struct AccelTls {
    rh264_ctx *ctx = nullptr;
    void *dev = nullptr;
    size_t cap = 0;
    std::vector<uint8_t> host;
};
thread_local AccelTls t_acc;

int tls_ctx(rh264_ctx **out)
{
    if (!t_acc.ctx) {
        int rc = rh264_init(g_accel_device, &t_acc.ctx);
        if (rc) return rc;
    }
    *out = t_acc.ctx;
    return 0;
}

int tls_stage(size_t bytes, void **dptr)
{
    if (bytes > t_acc.cap) {
        if (t_acc.dev) rh264_free(t_acc.ctx, t_acc.dev);
        t_acc.dev = nullptr;
        t_acc.cap = 0;
        const size_t cap = (bytes + 4095) & ~(size_t)4095;
        int rc = rh264_alloc(t_acc.ctx, cap, &t_acc.dev);
        if (rc) return rc;
        t_acc.cap = cap;
    }
    *dptr = t_acc.dev;
    return 0;
}

bool x4_shape(size_t w, size_t h) // satd_avg.rs:342-344
{
    return (w == 16 && (h == 16 || h == 8)) || (w == 8 && (h == 16 || h == 8));
}

// Packed host staging of n compare jobs: the source block, then `nref` reference
// blocks (each w x h at stride w), then the int64 offset tables.  One upload, one
// launch, one download.
struct Jobs {
    size_t w, h, nblk; // nblk = 1 source + references
    std::vector<uint8_t> &buf;
    explicit Jobs(size_t w_, size_t h_, size_t nref) : w(w_), h(h_), nblk(1 + nref), buf(t_acc.host)
    {
        buf.assign(nblk * w * h, 0);
    }
    void put(size_t slot, const uint8_t *p, size_t stride)
    {
        for (size_t r = 0; r < h; r++) memcpy(buf.data() + slot * w * h + r * w, p + r * stride, w);
    }
};

// One kind of metric over n pairs (source slot 0 against reference slots ra[i], and
// for SATD_AVG the average with rb[i]); out = n int32 values.
int run_jobs(Jobs &J, int kind, int n, const int *ra, const int *rb, int32_t *out)
{
    rh264_ctx *ctx;
    int rc = tls_ctx(&ctx);
    if (rc) return rc;
    const size_t pix = J.nblk * J.w * J.h, pixa = (pix + 7) & ~(size_t)7;
    const size_t tab = (size_t)n * 8;
    const size_t bytes = pixa + 3 * tab + (size_t)n * 4;
    void *dv;
    rc = tls_stage(bytes, &dv);
    if (rc) return rc;
    std::vector<uint8_t> &b = J.buf;
    b.resize(pixa + 3 * tab, 0);
    int64_t *so = (int64_t *)(b.data() + pixa), *oa = so + n, *ob = oa + n;
    for (int i = 0; i < n; i++) {
        so[i] = 0;
        oa[i] = (int64_t)ra[i] * (int64_t)(J.w * J.h);
        ob[i] = rb ? (int64_t)rb[i] * (int64_t)(J.w * J.h) : 0;
    }
    uint8_t *d = (uint8_t *)dv;
    rc = rh264_h2d(ctx, d, b.data(), pixa + 3 * tab);
    if (rc) return rc;
    const int64_t *dso = (const int64_t *)(d + pixa), *doa = dso + n, *dob = doa + n;
    int32_t *dout = (int32_t *)(d + pixa + 3 * tab);
    rc = rh264_block_metric(ctx, kind, (int)J.w, (int)J.h, n, d, (int)J.w, dso, d, (int)J.w, doa,
                            rb ? dob : nullptr, dout);
    if (rc) return rc;
    return rh264_d2h(ctx, out, dout, (size_t)n * 4);
}

int one_metric(int kind, size_t w, size_t h, const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out)
{
    if (!a || !b || !out) return rh264_fail("accel: null pointer");
    Jobs J(w, h, 1);
    J.put(0, a, sa);
    J.put(1, b, sb);
    const int ra = 1;
    return run_jobs(J, kind, 1, &ra, nullptr, out);
}

} // namespace

extern "C" {

int rh264_accel_set_device(int device)
{
    if (device < 0) return rh264_fail("accel: device %d", device);
    g_accel_device = device;
    return 0;
}

int rh264_accel_release(void)
{
    if (t_acc.dev) rh264_free(t_acc.ctx, t_acc.dev);
    if (t_acc.ctx) rh264_destroy(t_acc.ctx);
    t_acc.ctx = nullptr;
    t_acc.dev = nullptr;
    t_acc.cap = 0;
    std::vector<uint8_t>().swap(t_acc.host);
    return 0;
}

// satd_sad.rs:309-363: Wels SATD sum_4x4 ((sum|H d| + 1) >> 1) and SAD
#define RH264_ACCEL_PAIR(NAME, KIND, W, H)                                                                     \
    int NAME(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out)                          \
    {                                                                                                          \
        return one_metric(KIND, W, H, a, sa, b, sb, out);                                                      \
    }
RH264_ACCEL_PAIR(rh264_satd_16x16, RH264_METRIC_SATD_WELS, 16, 16)
RH264_ACCEL_PAIR(rh264_satd_16x8, RH264_METRIC_SATD_WELS, 16, 8)
RH264_ACCEL_PAIR(rh264_satd_8x16, RH264_METRIC_SATD_WELS, 8, 16)
RH264_ACCEL_PAIR(rh264_satd_8x8, RH264_METRIC_SATD_WELS, 8, 8)
RH264_ACCEL_PAIR(rh264_satd_4x4, RH264_METRIC_SATD_WELS, 4, 4)
RH264_ACCEL_PAIR(rh264_sad_16x16, RH264_METRIC_SAD, 16, 16)
RH264_ACCEL_PAIR(rh264_sad_16x8, RH264_METRIC_SAD, 16, 8)
RH264_ACCEL_PAIR(rh264_sad_8x16, RH264_METRIC_SAD, 8, 16)
#undef RH264_ACCEL_PAIR

// satd_avg.rs:462-492 (sad_x4) and 388-414 (satd_x4): one source against four
// offsets into one base
static int x4_base(int kind, const uint8_t *src, size_t ss, const uint8_t *base, const size_t o[4], size_t rs, size_t w,
                   size_t h, uint32_t out[4])
{
    if (!x4_shape(w, h)) return RH264_DECLINED;
    if (!src || !base || !o || !out) return rh264_fail("accel: null pointer");
    Jobs J(w, h, 4);
    J.put(0, src, ss);
    for (int i = 0; i < 4; i++) J.put(1 + i, base + o[i], rs);
    const int ra[4] = {1, 2, 3, 4};
    return run_jobs(J, kind, 4, ra, nullptr, (int32_t *)out);
}

int rh264_sad_x4(const uint8_t *src, size_t ss, const uint8_t *base, const size_t o[4], size_t rs, size_t w, size_t h,
                 uint32_t out[4])
{
    return x4_base(RH264_METRIC_SAD, src, ss, base, o, rs, w, h, out);
}

int rh264_satd_x4(const uint8_t *src, size_t ss, const uint8_t *base, const size_t o[4], size_t rs, size_t w,
                  size_t h, uint32_t out[4])
{
    return x4_base(RH264_METRIC_SATD, src, ss, base, o, rs, w, h, out);
}

// satd_avg.rs:350-381: four independent (plane, offset) operands, shared stride
int rh264_satd_x4p(const uint8_t *src, size_t ss, const uint8_t *const r[4], const size_t o[4], size_t rs, size_t w,
                   size_t h, uint32_t out[4])
{
    if (!x4_shape(w, h)) return RH264_DECLINED;
    if (!src || !r || !o || !out) return rh264_fail("accel: null pointer");
    Jobs J(w, h, 4);
    J.put(0, src, ss);
    for (int i = 0; i < 4; i++) J.put(1 + i, r[i] + o[i], rs);
    const int ra[4] = {1, 2, 3, 4};
    return run_jobs(J, RH264_METRIC_SATD, 4, ra, nullptr, (int32_t *)out);
}

// satd_avg.rs:536-571: four fused averages (a[i] + oa[i], b[i] + ob[i])
int rh264_satd_avg_x4(const uint8_t *src, size_t ss, const uint8_t *const a[4], const size_t oa[4],
                      const uint8_t *const b[4], const size_t ob[4], size_t rs, size_t w, size_t h, uint32_t out[4])
{
    if (!x4_shape(w, h)) return RH264_DECLINED;
    if (!src || !a || !oa || !b || !ob || !out) return rh264_fail("accel: null pointer");
    Jobs J(w, h, 8);
    J.put(0, src, ss);
    for (int i = 0; i < 4; i++) {
        J.put(1 + i, a[i] + oa[i], rs);
        J.put(5 + i, b[i] + ob[i], rs);
    }
    const int ra[4] = {1, 2, 3, 4}, rb[4] = {5, 6, 7, 8};
    return run_jobs(J, RH264_METRIC_SATD_AVG, 4, ra, rb, (int32_t *)out);
}

// satd_avg.rs:579-602: sum|H (src - (a+b+1)>>1)|
int rh264_satd_avg(const uint8_t *src, size_t src_stride, const uint8_t *a, const uint8_t *b, size_t stride, size_t w,
                   size_t h, uint32_t *out)
{
    if (!x4_shape(w, h)) return RH264_DECLINED;
    if (!src || !a || !b || !out) return rh264_fail("accel: null pointer");
    Jobs J(w, h, 2);
    J.put(0, src, src_stride);
    J.put(1, a, stride);
    J.put(2, b, stride);
    const int ra = 1, rb = 2;
    return run_jobs(J, RH264_METRIC_SATD_AVG, 1, &ra, &rb, (int32_t *)out);
}

// hpel.rs:34-43: H, V and C of the edge-padded plane f (pw x ph, stride pw), every
// tap clamped into f.  They are the interior of the plane sets that
// rh264_build_planes makes from f with a pad of 4: taps reach 3 samples past a
// sample, inside that replicated border.  Declines where that kernel does not apply
// (pw, ph multiples of 16).
int rh264_hpel_fused(const uint8_t *f, size_t pw, size_t ph, uint8_t *hp, uint8_t *vp, uint8_t *cp)
{
    if (pw < 24 || ph == 0 || pw % 16 || ph % 16 || pw > 16384 || ph > 16384) return RH264_DECLINED;
    if (!f || !hp || !vp || !cp) return rh264_fail("accel: null pointer");
    rh264_ctx *ctx;
    int rc = tls_ctx(&ctx);
    if (rc) return rc;
    const int pad = 4;
    const size_t fr = pw * ph, fra = (fr + 255) & ~(size_t)255;
    const size_t qw = pw + 2 * pad, qh = ph + 2 * pad, qn = qw * qh;
    const size_t pb = rh264_planes_bytes((int)pw, (int)ph, pad);
    void *dv;
    rc = tls_stage(fra + pb, &dv);
    if (rc) return rc;
    uint8_t *d = (uint8_t *)dv;
    rc = rh264_h2d(ctx, d, f, fr);
    if (rc) return rc;
    rc = rh264_build_planes(ctx, d, fr, 1, (int)pw, (int)ph, pad, d + fra);
    if (rc) return rc;
    std::vector<uint8_t> &b = t_acc.host;
    b.resize(3 * qn);
    rc = rh264_d2h(ctx, b.data(), d + fra + qn, 3 * qn);
    if (rc) return rc;
    uint8_t *dst[3] = {hp, vp, cp};
    for (int k = 0; k < 3; k++)
        for (size_t y = 0; y < ph; y++) memcpy(dst[k] + y * pw, b.data() + k * qn + (y + pad) * qw + pad, pw);
    return 0;
}

// ---- MeCtx (mectx.rs:27-166) --------------------------------------------------
struct rh264_mectx {
    rh264_ctx *ctx;
    uint8_t *dev;   // [F | H | V | C | source block], planes pw*ph each, the block w*h
    int64_t pw, ph, pad, lx, ly, w, h, px_max, py_max;
};

int rh264_mectx_new(const uint8_t *src, size_t src_len, size_t cw, const uint8_t *f, const uint8_t *hp,
                    const uint8_t *vp, const uint8_t *cp, size_t plane_len, size_t stride, size_t pad, size_t pw,
                    size_t ph, size_t lx, size_t ly, size_t w, size_t h, rh264_mectx **out)
{
    if (!out) return rh264_fail("mectx: null output");
    *out = nullptr;
    if (stride != pw || !x4_shape(w, h)) return RH264_DECLINED;          // mectx.rs:66-78
    if (!src || !f || !hp || !vp || !cp) return RH264_DECLINED;
    if (src_len < (h - 1) * cw + w) return RH264_DECLINED;                // mectx.rs:79-81
    if (ph != 0 && pw > SIZE_MAX / ph) return RH264_DECLINED;
    const size_t need = pw * ph;                                          // mectx.rs:82-85
    if (plane_len < need || need == 0) return RH264_DECLINED;
    if ((int64_t)pw - (int64_t)w - 1 < 0 || (int64_t)ph - (int64_t)h - 1 < 0) return RH264_DECLINED;
    rh264_ctx *ctx;
    int rc = tls_ctx(&ctx);
    if (rc) return rc;
    rh264_mectx *m = new rh264_mectx();
    m->ctx = ctx;
    m->pw = (int64_t)pw;
    m->ph = (int64_t)ph;
    m->pad = (int64_t)pad;
    m->lx = (int64_t)lx;
    m->ly = (int64_t)ly;
    m->w = (int64_t)w;
    m->h = (int64_t)h;
    m->px_max = (int64_t)pw - (int64_t)w - 1;
    m->py_max = (int64_t)ph - (int64_t)h - 1;
    void *dv = nullptr;
    rc = rh264_alloc(ctx, 4 * need + w * h + 256, &dv);
    if (rc) {
        delete m;
        return rc;
    }
    m->dev = (uint8_t *)dv;
    const uint8_t *pl[4] = {f, hp, vp, cp};
    for (int k = 0; k < 4 && !rc; k++) rc = rh264_h2d(ctx, m->dev + k * need, pl[k], need);
    std::vector<uint8_t> blk(w * h);
    for (size_t r = 0; r < h; r++) memcpy(blk.data() + r * w, src + r * cw, w);
    if (!rc) rc = rh264_h2d(ctx, m->dev + 4 * need, blk.data(), w * h);
    if (rc) {
        rh264_free(ctx, m->dev);
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

int rh264_mectx_free(rh264_mectx *m)
{
    if (!m) return 0;
    int rc = rh264_free(m->ctx, m->dev);
    delete m;
    return rc;
}

// The candidates of n vectors: ok[i] = 0 where the reference's eval returns None
// (outside the validated window, mectx.rs:114-118), else out[i] = the encoder-domain
// SATD (full and half phases: one plane; quarter phases: the (a+b+1)>>1 average of
// the operand pair of mectx.rs:143-156).
int rh264_mectx_eval_n(rh264_mectx *m, int n, const int32_t *mvs, uint32_t *out, uint8_t *ok)
{
    if (!m || (n > 0 && (!mvs || !out || !ok))) return rh264_fail("mectx: null pointer");
    if (n <= 0) return 0;
    const int64_t pn = m->pw * m->ph;
    // operand table (plane, dx, dy) per phase fx*4 + fy; one-plane phases name one
    static const int8_t qa[16][3] = {{0, 0, 0}, {0, 0, 0}, {2, 0, 0}, {0, 0, 1}, {0, 0, 0}, {1, 0, 0},
                                     {2, 0, 0}, {1, 0, 1}, {1, 0, 0}, {1, 0, 0}, {3, 0, 0}, {1, 0, 1},
                                     {0, 1, 0}, {1, 0, 0}, {2, 1, 0}, {1, 0, 1}};
    static const int8_t qb[16][3] = {{0, 0, 0}, {2, 0, 0}, {2, 0, 0}, {2, 0, 0}, {1, 0, 0}, {2, 0, 0},
                                     {3, 0, 0}, {2, 0, 0}, {1, 0, 0}, {3, 0, 0}, {3, 0, 0}, {3, 0, 0},
                                     {1, 0, 0}, {2, 1, 0}, {3, 0, 0}, {2, 1, 0}};
    std::vector<int64_t> one_off, avg_a, avg_b;
    std::vector<int> one_i, avg_i;
    for (int i = 0; i < n; i++) {
        const int32_t mvx = mvs[2 * i], mvy = mvs[2 * i + 1];
        const int64_t px = m->lx + (mvx >> 2) + m->pad, py = m->ly + (mvy >> 2) + m->pad;
        ok[i] = !(px < 0 || py < 0 || px > m->px_max || py > m->py_max);
        if (!ok[i]) continue;
        const int64_t base = py * m->pw + px;
        const int ph = (mvx & 3) * 4 + (mvy & 3);
        const int8_t *A = qa[ph], *Bq = qb[ph];
        const int64_t oa = A[0] * pn + base + A[1] + A[2] * m->pw;
        if (((mvx | mvy) & 1) == 0) {
            one_off.push_back(oa);
            one_i.push_back(i);
        } else {
            avg_a.push_back(oa);
            avg_b.push_back(Bq[0] * pn + base + Bq[1] + Bq[2] * m->pw);
            avg_i.push_back(i);
        }
    }
    const int kinds[2] = {RH264_METRIC_SATD, RH264_METRIC_SATD_AVG};
    for (int pass = 0; pass < 2; pass++) {
        const std::vector<int64_t> &oa = pass ? avg_a : one_off;
        const std::vector<int> &idx = pass ? avg_i : one_i;
        const int k = (int)idx.size();
        if (!k) continue;
        const size_t tab = (size_t)k * 8;
        void *dv;
        int rc = tls_stage(3 * tab + (size_t)k * 4, &dv);
        if (rc) return rc;
        std::vector<int64_t> host((size_t)3 * k);
        for (int j = 0; j < k; j++) {
            host[j] = 0;
            host[k + j] = oa[j];
            host[2 * k + j] = pass ? avg_b[j] : 0;
        }
        int64_t *d = (int64_t *)dv;
        rc = rh264_h2d(m->ctx, d, host.data(), 3 * tab);
        if (rc) return rc;
        int32_t *dout = (int32_t *)(d + 3 * k);
        rc = rh264_block_metric(m->ctx, kinds[pass], (int)m->w, (int)m->h, k, m->dev + 4 * pn, (int)m->w, d, m->dev,
                                (int)m->pw, d + k, pass ? d + 2 * k : nullptr, dout);
        if (rc) return rc;
        std::vector<int32_t> res(k);
        rc = rh264_d2h(m->ctx, res.data(), dout, (size_t)k * 4);
        if (rc) return rc;
        for (int j = 0; j < k; j++) out[idx[j]] = (uint32_t)res[j];
    }
    return 0;
}

// mectx.rs:112-166: 0 with *out = the SATD, or 1 (None) outside the window
int rh264_mectx_eval(rh264_mectx *m, int32_t mvx, int32_t mvy, uint32_t *out)
{
    const int32_t mv[2] = {mvx, mvy};
    uint8_t ok = 0;
    int rc = rh264_mectx_eval_n(m, 1, mv, out, &ok);
    if (rc) return rc;
    return ok ? 0 : RH264_DECLINED;
}

} // extern "C"
