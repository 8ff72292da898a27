// rh264_me.hip -- MI355X (gfx950) kernels and C ABI for rusty_h264's encoder
// motion-search hot path.  Declarations and the contract: include/rh264_me.h.
//
// Kernels (one launch per stage per batch of frames):
//   k_build_planes   F/H/V/C plane sets (inter.rs:786-987, hpel.rs:34-203)
//   k_fill_u64       clears the per-block argmin keys (a kernel, not a memset,
//                    so a captured graph replays it correctly)
//   k_full_search    exhaustive integer SAD, quad-SAD (v_qsad_pk_u16_u8) over a
//                    vertical ring of accumulators, segmented wave argmin +
//                    64-bit atomicMin on (cost << 32 | raster index)
//   k_subpel_refine  sub-pel ring walk (mb16.rs:2956-3291), one lane group per
//                    block, 4x4 Hadamards across quads with DPP
//   k_block_metric   the accel per-block family, batched
//   k_predict        qpel prediction of blocks from a plane set
//
// All arithmetic is integer and exact.

#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <cmath>
#include <utility>
#include <string>
#include <vector>

#include "rh264_me.h"
#include "rh264_internal.h"

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------

static thread_local std::string g_err;

int rh264_fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

// The argmin keys of rh264_plain_search_noalloc: a static device array, one per device
// the library runs on (each holds it until the process ends).  Host code never passes
// its name to a kernel: rh264_init looks up its device address on the context's
// device and keeps it in ctx->noalloc_keys.
static __device__ __align__(256) uint64_t scratch[RH264_NOALLOC_KEYS];

static int ensure_keys(rh264_ctx *ctx, size_t n)
{
    if (ctx->keys_cap >= n) return 0;
    HIP_TRY(hipStreamSynchronize(ctx->stream)); // earlier launches may still use the old buffer
    if (ctx->keys) HIP_TRY(hipFree(ctx->keys));
    ctx->keys = nullptr;
    ctx->keys_cap = 0;
    HIP_TRY(hipMalloc(&ctx->keys, n * sizeof(uint64_t)));
    ctx->keys_cap = n;
    return 0;
}

static int ensure_tmp_mv(rh264_ctx *ctx, size_t n)
{
    if (ctx->tmp_mv_cap >= n) return 0;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->tmp_mv) HIP_TRY(hipFree(ctx->tmp_mv));
    ctx->tmp_mv = nullptr;
    ctx->tmp_mv_cap = 0;
    HIP_TRY(hipMalloc(&ctx->tmp_mv, n * sizeof(int32_t)));
    ctx->tmp_mv_cap = n;
    return 0;
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }
__device__ __forceinline__ int clip_u8(int v) { return min(max(v, 0), 255); }
// clip_u8((v + round) >> shift) computed as a clamp followed by a logical shift:
// for v + round < 0 both give 0, above (256 << shift) - 1 both give 255.  Written
// this way so the shift/saturate/pack idiom is not formed: when it was, the
// compiler emitted gfx950's packed shift-and-saturate instruction for the plane
// kernel and the H/V/C bytes it produced differed from this scalar rule.
template <int SHIFT, int ROUND> __device__ __forceinline__ uint32_t round_clip_u8(int v)
{
    return (uint32_t)min(max(v + ROUND, 0), (256 << SHIFT) - 1) >> SHIFT;
}

// Exp-Golomb se(d) length, mb16.rs:2474-2475: 1 + 2*floor(log2(codenum+1)).
// |d| is clamped to 2^24 quarter pels (far beyond any coded MV) so 2d-1 cannot wrap.
__device__ __forceinline__ int mvbits(int d)
{
    d = clampi(d, -(1 << 24), 1 << 24);
    unsigned codenum = d > 0 ? (unsigned)(2 * d - 1) : (unsigned)(-2 * d);
    return 1 + 2 * (31 - __clz((int)(codenum + 1u)));
}

// Quarter-phase operand table, inter.rs:1079-1092 / mectx.rs:143-156: for phase
// fx*4+fy, the plane (0=F, 1=H, 2=V, 3=C) and integer offset (x, y) of operand a
// and of operand b; one-plane phases name the same operand twice.
struct QPhase {
    int pa, ax, ay, pb, bx, by;
};
constexpr QPhase k_qphase[16] = {
    /* fx=0: fy=0..3 */ {0, 0, 0, 0, 0, 0}, {0, 0, 0, 2, 0, 0}, {2, 0, 0, 2, 0, 0}, {0, 0, 1, 2, 0, 0},
    /* fx=1 */          {0, 0, 0, 1, 0, 0}, {1, 0, 0, 2, 0, 0}, {2, 0, 0, 3, 0, 0}, {1, 0, 1, 2, 0, 0},
    /* fx=2 */          {1, 0, 0, 1, 0, 0}, {1, 0, 0, 3, 0, 0}, {3, 0, 0, 3, 0, 0}, {1, 0, 1, 3, 0, 0},
    /* fx=3 */          {0, 1, 0, 1, 0, 0}, {1, 0, 0, 2, 1, 0}, {2, 1, 0, 3, 0, 0}, {1, 0, 1, 2, 1, 0},
};
// the same, packed 4 bits a field (plane_a, ax, ay, plane_b, bx, by from bit 0)
constexpr uint32_t qent(const QPhase &q)
{
    return (uint32_t)q.pa | ((uint32_t)q.ax << 4) | ((uint32_t)q.ay << 8) | ((uint32_t)q.pb << 12) |
           ((uint32_t)q.bx << 16) | ((uint32_t)q.by << 20);
}
__constant__ uint32_t c_qtab[16] = {
    qent(k_qphase[0]),  qent(k_qphase[1]),  qent(k_qphase[2]),  qent(k_qphase[3]),
    qent(k_qphase[4]),  qent(k_qphase[5]),  qent(k_qphase[6]),  qent(k_qphase[7]),
    qent(k_qphase[8]),  qent(k_qphase[9]),  qent(k_qphase[10]), qent(k_qphase[11]),
    qent(k_qphase[12]), qent(k_qphase[13]), qent(k_qphase[14]), qent(k_qphase[15]),
};
__device__ __forceinline__ uint32_t qtab(int idx) { return c_qtab[idx]; }

// Byte-wise differences of two packed rows of four pixels.
__device__ __forceinline__ void diff4(uint32_t s, uint32_t p, int &d0, int &d1, int &d2, int &d3)
{
    d0 = (int)(s & 255) - (int)(p & 255);
    d1 = (int)((s >> 8) & 255) - (int)((p >> 8) & 255);
    d2 = (int)((s >> 16) & 255) - (int)((p >> 16) & 255);
    d3 = (int)(s >> 24) - (int)(p >> 24);
}

// Per-byte rounded-up average (a+b+1)>>1 of four packed bytes.
__device__ __forceinline__ uint32_t avg4(uint32_t a, uint32_t b)
{
    return (a | b) - (((a ^ b) >> 1) & 0x7f7f7f7fu);
}

// 4 consecutive bytes at (x, y) of a pw-stride plane, x aligned or not, all
// bytes inside the buffer (caller guarantees): two aligned dwords + alignbyte.
__device__ __forceinline__ uint32_t load4_fast(const uint8_t *plane, int pw, int x, int y)
{
    size_t a = (size_t)y * pw + x;
    const uint32_t *p = (const uint32_t *)(plane + (a & ~(size_t)3));
    return __builtin_amdgcn_alignbyte(p[1], p[0], (uint32_t)(a & 3));
}

// Same with every coordinate clamped into the plane (infinite edge replication).
__device__ __forceinline__ uint32_t load4_clamped(const uint8_t *plane, int pw, int ph, int x, int y)
{
    const uint8_t *row = plane + (size_t)clampi(y, 0, ph - 1) * pw;
    uint32_t v = 0;
#pragma unroll
    for (int m = 0; m < 4; m++) v |= (uint32_t)row[clampi(x + m, 0, pw - 1)] << (8 * m);
    return v;
}

// Workgroups are dispatched to the 8 XCDs round-robin by linear workgroup id, and
// each XCD has its own L2.  The search kernels renumber their workgroups so that
// XCD x runs one contiguous share of the logical ones: neighbouring blocks, whose
// reference windows overlap, then meet in one L2 instead of being fetched by all
// eight.
#ifndef RH264_XCD_MAP
#define RH264_XCD_MAP 1
#endif
constexpr unsigned RH264_XCDS = 8;
__device__ __forceinline__ unsigned xcd_linear(unsigned id, unsigned n)
{
    if (!RH264_XCD_MAP) return id;
    const unsigned x = id % RH264_XCDS, i = id / RH264_XCDS, q = n / RH264_XCDS, r = n % RH264_XCDS;
    return x * q + min(x, r) + i;
}
// logical (x, y) of this workgroup in a 2-D grid, and (x, y, z) in a 3-D one
__device__ __forceinline__ void xcd_grid2(unsigned &bx, unsigned &by)
{
    const unsigned l = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    bx = l % gridDim.x;
    by = l / gridDim.x;
}
__device__ __forceinline__ void xcd_grid3(unsigned &bx, unsigned &by, unsigned &bz)
{
    const unsigned gxy = gridDim.x * gridDim.y;
    const unsigned l = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x + blockIdx.z * gxy, gxy * gridDim.z);
    bz = l / gxy;
    by = (l - bz * gxy) / gridDim.x;
    bx = l - bz * gxy - by * gridDim.x;
}

// ---------------------------------------------------------------------------
// plane build: F, H, V, C over the padded plane (inter.rs:816-843 + hpel.rs /
// build_hpel_fused inter.rs:908-987).  Every tap index is clamped, which
// equals edge replication of the source frame (pad_plane, inter.rs:786-814).
// Tile: 256 columns x 16 rows per 256-thread block; each thread makes a 4x4
// patch and stores whole dwords.
// ---------------------------------------------------------------------------

constexpr int PT_W = 256;
constexpr int PT_H = 16;
constexpr int PT_LW = PT_W + 8; // LDS row pitch (>= PT_W + 5)

__global__ __launch_bounds__(256) void k_build_planes(const uint8_t *__restrict__ frames,
                                                      size_t frame_stride, int cw, int ch, int pad,
                                                      uint8_t *__restrict__ planes,
                                                      size_t planes_bytes)
{
    __shared__ uint8_t s_in[PT_H + 5][PT_LW];
    __shared__ int16_t s_vt[PT_H][PT_LW];
    const int pw = cw + 2 * pad, ph = ch + 2 * pad;
    const size_t n = (size_t)pw * ph;
    const uint8_t *y = frames + (size_t)blockIdx.z * frame_stride;
    uint8_t *out = planes + (size_t)blockIdx.z * planes_bytes;
    const int x0 = blockIdx.x * PT_W, y0 = blockIdx.y * PT_H;
    const int tid = threadIdx.x;

    // source tile: rows y0-2 .. y0+PT_H+2, columns x0-2 .. x0+PT_W+3 (padded coords)
    for (int i = tid; i < (PT_H + 5) * (PT_W + 6); i += 256) {
        int r = i / (PT_W + 6), c = i - r * (PT_W + 6);
        int sy = clampi(y0 - 2 + r - pad, 0, ch - 1);
        int sx = clampi(x0 - 2 + c - pad, 0, cw - 1);
        s_in[r][c] = y[(size_t)sy * cw + sx];
    }
    __syncthreads();
    // vertical 6-tap intermediates (exact in int16: -2550..10710)
    for (int i = tid; i < PT_H * (PT_W + 6); i += 256) {
        int r = i / (PT_W + 6), c = i - r * (PT_W + 6);
        int v = (int)s_in[r][c] - 5 * (int)s_in[r + 1][c] + 20 * (int)s_in[r + 2][c] +
                20 * (int)s_in[r + 3][c] - 5 * (int)s_in[r + 4][c] + (int)s_in[r + 5][c];
        s_vt[r][c] = (int16_t)v;
    }
    __syncthreads();
    const int cx = (tid & 63) * 4;  // tile column of the 4-wide patch
    const int ry0 = (tid >> 6) * 4; // tile row of the 4-tall patch
    const int gx = x0 + cx;
    if (gx >= pw) return;
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        const int r = ry0 + rr, gy = y0 + r;
        if (gy >= ph) break;
        uint32_t f4 = 0, h4 = 0, v4 = 0, c4 = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int c = cx + k; // output column gx+k <-> tile column c+2
            const uint8_t *in = &s_in[r + 2][c];
            int hs = (int)in[0] - 5 * (int)in[1] + 20 * (int)in[2] + 20 * (int)in[3] -
                     5 * (int)in[4] + (int)in[5];
            const int16_t *vt = &s_vt[r][c];
            int cs = (int)vt[0] - 5 * (int)vt[1] + 20 * (int)vt[2] + 20 * (int)vt[3] -
                     5 * (int)vt[4] + (int)vt[5];
            f4 |= (uint32_t)in[2] << (8 * k);
            h4 |= round_clip_u8<5, 16>(hs) << (8 * k);
            v4 |= round_clip_u8<5, 16>((int)vt[2]) << (8 * k);
            c4 |= round_clip_u8<10, 512>(cs) << (8 * k);
        }
        const size_t o = (size_t)gy * pw + gx;
        *(uint32_t *)(out + o) = f4;
        *(uint32_t *)(out + n + o) = h4;
        *(uint32_t *)(out + 2 * n + o) = v4;
        *(uint32_t *)(out + 3 * n + o) = c4;
    }
    // zero the tail slack once per plane set
    if (blockIdx.x == 0 && blockIdx.y == 0 && tid < 64)
        *(uint32_t *)(out + 4 * n + 4 * tid) = 0u;
}

// ---------------------------------------------------------------------------
// key fill
// ---------------------------------------------------------------------------

__global__ void k_fill_u64(uint64_t *p, size_t n, uint64_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Plane build, register-window form: a 256 x PB_H output tile per 256-thread block.
// The clamped source rows (y0-2 .. y0+PB_H+2, columns x0-4 .. x0+263) are staged in
// LDS as dwords; thread (c, g) makes columns x0+4c .. +3 of PB_H/4 rows, sliding
// a window of six source rows held as 16-bit pairs down its strip.  Per row:
//   vt (vertical 6-tap, exact in int16: -2550..10710) at columns x-2 .. x+7,
//   V = clip((vt + 16) >> 5) at x .. x+3,
//   C = clip((6-tap of vt + 512) >> 10) with v_dot2_i32_i16 (32-bit sums),
//   H = clip((horizontal 6-tap + 16) >> 5) on the centre row in int16 pairs,
//   F = the centre row's source dword.
// Every staged dword lies wholly inside the frame or wholly in the replicated
// border (x0, pad and cw are multiples of 4), so interior dwords are plain loads.
#ifndef RH264_PB_H
#define RH264_PB_H 32
#endif
constexpr int PB_W = 256, PB_H = RH264_PB_H, PB_RPT = PB_H / 4, PB_DW = PB_W / 4 + 4, PB_ROWS = PB_H + 5;

typedef short pks2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ pks2 pk_of(uint32_t v) { return __builtin_bit_cast(pks2, v); }
__device__ __forceinline__ uint32_t u_of(pks2 v) { return __builtin_bit_cast(uint32_t, v); }
// (b[k], b[k+1]) of a little-endian dword as two 16-bit lanes, k = 0 or 2
__device__ __forceinline__ pks2 bytes_lo(uint32_t v) { return pk_of(__builtin_amdgcn_perm(0u, v, 0x0c010c00u)); }
__device__ __forceinline__ pks2 bytes_hi(uint32_t v) { return pk_of(__builtin_amdgcn_perm(0u, v, 0x0c030c02u)); }
// pair (a.hi, b.lo): the 16-bit pair that starts one lane to the right of a
__device__ __forceinline__ pks2 pk_shift(pks2 a, pks2 b) { return pk_of(__builtin_amdgcn_alignbit(u_of(b), u_of(a), 16)); }
// 6-tap (1, -5, 20, 20, -5, 1) over six pairs, lane-wise, exact in int16 for byte inputs
__device__ __forceinline__ pks2 tap6(pks2 a, pks2 b, pks2 c, pks2 d, pks2 e, pks2 f)
{
    const pks2 k20 = {20, 20}, k5 = {5, 5};
    return (a + f) + k20 * (c + d) - k5 * (b + e);
}
// clip((v + 16) >> 5) of two int16 lanes, as bytes 0 and 1 of the result
__device__ __forceinline__ uint32_t round5_pair(pks2 v)
{
    const pks2 r = {16, 16}, z = {0, 0}, m = {255, 255};
    pks2 t = (v + r) >> 5;
    t = __builtin_elementwise_min(__builtin_elementwise_max(t, z), m);
    return u_of(t);
}

__global__ __launch_bounds__(256) void k_build_planes_rw(const uint8_t *__restrict__ frames, size_t frame_stride,
                                                         int cw, int ch, int pad, uint8_t *__restrict__ planes,
                                                         size_t planes_bytes)
{
    __shared__ uint32_t s_src[PB_ROWS * PB_DW];
    const int pw = cw + 2 * pad, ph = ch + 2 * pad;
    const size_t n = (size_t)pw * ph;
    // tiles in row-major order, one contiguous share per XCD: the tiles that share
    // halo rows and boundary cache lines read them through one L2
    unsigned tx, ty, tz;
    xcd_grid3(tx, ty, tz);
    const uint8_t *fr = frames + (size_t)tz * frame_stride;
    uint8_t *out = planes + (size_t)tz * planes_bytes;
    const int x0 = (int)tx * PB_W, y0 = (int)ty * PB_H;
    const int tid = threadIdx.x;
    {
        // every load of a thread is issued before its first store: a dword left of the
        // frame is the first column's byte replicated (byte 0 of the dword at column 0),
        // right of it the last column's (byte 3 of the dword at cw - 4)
        constexpr int NS = (PB_ROWS * PB_DW + 255) / 256;
        uint32_t v[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const int i = min(tid + 256 * k, PB_ROWS * PB_DW - 1), r = i / PB_DW, d = i - r * PB_DW;
            const uint8_t *row = fr + (size_t)clampi(y0 - 2 + r - pad, 0, ch - 1) * cw;
            v[k] = *(const uint32_t *)(row + clampi(x0 - 4 + 4 * d - pad, 0, cw - 4));
        }
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const int i = tid + 256 * k, r = i / PB_DW, d = i - r * PB_DW, sx = x0 - 4 + 4 * d - pad;
            const uint32_t w = sx < 0 ? 0x01010101u * (v[k] & 0xffu) : sx >= cw ? 0x01010101u * (v[k] >> 24) : v[k];
            if (i < PB_ROWS * PB_DW) s_src[i] = w;
        }
    }
    __syncthreads();
    const int c = tid & 63, g = tid >> 6;
    const int x = x0 + 4 * c;
    if (x >= pw) return;
    // window: source rows (LDS rows 16g + k .. +5), each as 3 raw dwords and 6 pairs
    pks2 P[6][6];
    uint32_t raw[6];
    auto load_row = [&](int w, int lr) {
        const uint32_t *rp = s_src + lr * PB_DW + c;
        const uint32_t d0 = rp[0], d1 = rp[1], d2 = rp[2];
        raw[w] = d1;
        P[w][0] = bytes_lo(d0); P[w][1] = bytes_hi(d0);
        P[w][2] = bytes_lo(d1); P[w][3] = bytes_hi(d1);
        P[w][4] = bytes_lo(d2); P[w][5] = bytes_hi(d2);
    };
#pragma unroll
    for (int w = 0; w < 5; w++) load_row(w, PB_RPT * g + w);
#pragma unroll
    for (int k = 0; k < PB_RPT; k++) {
        const int w5 = (k + 5) % 6; // ring slot of the new row
        load_row(w5, PB_RPT * g + k + 5);
        const int oy = y0 + PB_RPT * g + k;
        if (oy >= ph) break;
        // ring slots of window rows 0..5 (source rows oy-2 .. oy+3)
        int sl[6];
#pragma unroll
        for (int j = 0; j < 6; j++) sl[j] = (k + j) % 6;
        // vertical 6-tap at column pairs 1..5 (columns x-2 .. x+7)
        pks2 vt[6];
#pragma unroll
        for (int m = 1; m < 6; m++)
            vt[m] = tap6(P[sl[0]][m], P[sl[1]][m], P[sl[2]][m], P[sl[3]][m], P[sl[4]][m], P[sl[5]][m]);
        // V at columns x .. x+3
        const uint32_t v01 = round5_pair(vt[2]), v23 = round5_pair(vt[3]);
        const uint32_t vq = __builtin_amdgcn_perm(v23, v01, 0x06040200u);
        // C: 6-tap of vt in 32 bits, outputs x+0..x+3
        const pks2 w15 = {1, -5}, w20 = {20, 20}, w51 = {-5, 1};
        const pks2 s1 = pk_shift(vt[1], vt[2]), s2 = pk_shift(vt[2], vt[3]), s3 = pk_shift(vt[3], vt[4]),
                   s4 = pk_shift(vt[4], vt[5]);
        int cs[4];
        cs[0] = __builtin_amdgcn_sdot2(vt[1], w15, __builtin_amdgcn_sdot2(vt[2], w20, __builtin_amdgcn_sdot2(vt[3], w51, 512, false), false), false);
        cs[1] = __builtin_amdgcn_sdot2(s1, w15, __builtin_amdgcn_sdot2(s2, w20, __builtin_amdgcn_sdot2(s3, w51, 512, false), false), false);
        cs[2] = __builtin_amdgcn_sdot2(vt[2], w15, __builtin_amdgcn_sdot2(vt[3], w20, __builtin_amdgcn_sdot2(vt[4], w51, 512, false), false), false);
        cs[3] = __builtin_amdgcn_sdot2(s2, w15, __builtin_amdgcn_sdot2(s3, w20, __builtin_amdgcn_sdot2(s4, w51, 512, false), false), false);
        // clip((c + 512) >> 10) as clamp-then-logical-shift (the +512 is in the sums):
        // this spelling keeps the compiler from forming gfx950's packed
        // shift-and-saturate instruction, whose bytes differ from the rule
        uint32_t cq = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) cq |= ((uint32_t)min(max(cs[q], 0), (256 << 10) - 1) >> 10) << (8 * q);
        // H on the centre row (window row 2)
        const pks2 *R = P[sl[2]];
        const pks2 r1 = pk_shift(R[1], R[2]), r2 = pk_shift(R[2], R[3]), r3 = pk_shift(R[3], R[4]),
                   r4 = pk_shift(R[4], R[5]);
        const uint32_t h01 = round5_pair(tap6(R[1], r1, R[2], r2, R[3], r3));
        const uint32_t h23 = round5_pair(tap6(R[2], r2, R[3], r3, R[4], r4));
        const uint32_t hq = __builtin_amdgcn_perm(h23, h01, 0x06040200u);
        const size_t o = (size_t)oy * pw + x;
        *(uint32_t *)(out + o) = raw[sl[2]];
        *(uint32_t *)(out + n + o) = hq;
        *(uint32_t *)(out + 2 * n + o) = vq;
        *(uint32_t *)(out + 3 * n + o) = cq;
    }
    if (tx == 0 && ty == 0 && tid < 64) *(uint32_t *)(out + 4 * n + 4 * tid) = 0u;
}

// ---------------------------------------------------------------------------
// exhaustive integer search
// ---------------------------------------------------------------------------
//
// Lane <-> (block, column group g): the lane owns the four horizontally adjacent
// candidates dx = -R + 4g + k (k = 0..3) for every dy, walking the reference rows
// of its window top to bottom.  Row j feeds candidate row t = j - k for source
// row k, so a ring of BH accumulators (four packed u16 SADs each: a BWxBH block
// SAD is at most 255*256 = 65280 < 65536) holds the candidate rows in flight;
// after row j the candidate row j-(BH-1) is complete and is priced in raster
// order.  v_qsad_pk_u16_u8 computes the four alignments of one source dword
// against eight reference bytes in one instruction.

template <int BW, int BH> struct FsShape {
    static constexpr int SD = BW / 4;  // source dwords per row
    static constexpr int RD = SD + 1;  // reference dwords per row (before realign)
};

__device__ __forceinline__ uint64_t wave_seg_min(uint64_t key, int seg, int lane)
{
    // segmented min over lanes with equal `seg`, lanes sorted by seg: afterwards
    // the first lane of every segment holds that segment's minimum.
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint64_t ok = __shfl_down(key, off, 64);
        int os = __shfl_down(seg, off, 64);
        if (lane + off < 64 && os == seg && ok < key) key = ok;
    }
    return key;
}

// State of one lane of the integer search.
template <int BW, int BH> struct FsLane {
    uint32_t src[BH][FsShape<BW, BH>::SD];
    uint64_t acc[BH];
    uint32_t best, bidx;
    int rxb[4];
};

// Rows jb .. jb+BH-1 of the lane's window.  MODE 0: every (row, source row) pair;
// MODE 1: first block of a window whose height is a multiple of BH -- only pairs
// with candidate row t = j - k >= 0; MODE 2: last such block -- only t <= 2R.  The
// masks are compile-time after unrolling, so no quad-SAD is spent on candidate rows
// outside [-R, R].
template <int BW, int BH, int MODE>
__device__ __forceinline__ void fs_rows(FsLane<BW, BH> &L, const uint8_t *__restrict__ F, int pw, int ph,
                                        bool fast, int px0, int py0, int shift, int jb, int NJ, int R, int NC,
                                        int g, int cy, int mpy, const int32_t *s_lr)
{
    using S = FsShape<BW, BH>;
#pragma unroll
    for (int jj = 0; jj < BH; jj++) {
        const int j = jb + jj;
        if (MODE == 0 && j >= NJ) break;
        uint32_t w[S::RD];
        const int ry = py0 + j;
        if (fast) {
            const uint32_t *rp = (const uint32_t *)(F + (size_t)ry * pw + (px0 & ~3));
            uint32_t a[S::RD + 1];
#pragma unroll
            for (int i = 0; i < S::RD + 1; i++) a[i] = rp[i];
#pragma unroll
            for (int i = 0; i < S::RD; i++) w[i] = __builtin_amdgcn_alignbyte(a[i + 1], a[i], shift);
        } else {
#pragma unroll
            for (int i = 0; i < S::RD; i++) w[i] = load4_clamped(F, pw, ph, px0 + 4 * i, ry);
        }
#pragma unroll
        for (int k = 0; k < BH; k++) {
            if (MODE == 1 && k > jj) continue;
            if (MODE == 2 && k < jj) continue;
            const int slot = (jj - k + BH) % BH;
#pragma unroll
            for (int i = 0; i < S::SD; i++) {
                uint64_t ref8 = ((uint64_t)w[i + 1] << 32) | w[i];
                L.acc[slot] = __builtin_amdgcn_qsad_pk_u16_u8(ref8, L.src[k][i], L.acc[slot]);
            }
        }
        const int t = j - (BH - 1);
        const int sd = (jj + 1) % BH;
        if (t >= 0) {
            const int ry_bits = mvbits(cy + 4 * (t - R) - mpy);
            const uint64_t a4 = L.acc[sd];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int c = 4 * g + k;
                const uint32_t sad = (uint32_t)(a4 >> (16 * k)) & 0xffffu;
                const uint32_t cost = sad + (uint32_t)s_lr[L.rxb[k] + ry_bits];
                if (c < NC && cost < L.best) {
                    L.best = cost;
                    L.bidx = (uint32_t)(t * NC + c);
                }
            }
        }
        L.acc[sd] = 0;
    }
}

#ifndef RH264_FS_WAVES
#define RH264_FS_WAVES 2
#endif
// One lane's share of the integer search of one block: the block's source, its
// reference plane F and search geometry.
struct FsJob {
    const uint8_t *srcp; // top-left sample of the source block (row stride cw)
    const uint8_t *F;    // full-pel plane of the reference's plane set
    int cw, pw, ph, pad, R;
    int lx, ly, cx, cy, mpx, mpy, g;
};

// LSRC: the source block is staged in LDS at lsrc (BH rows of SD dwords) instead of
// read from J.srcp.
template <int BW, int BH, bool LSRC = false>
__device__ __forceinline__ uint64_t fs_qsad_key(const FsJob &J, const int32_t *s_lr, const uint32_t *lsrc = nullptr)
{
    using S = FsShape<BW, BH>;
    const int R = J.R, NC = 2 * R + 1, g = J.g;
    FsLane<BW, BH> L;
#pragma unroll
    for (int k = 0; k < BH; k++)
#pragma unroll
        for (int i = 0; i < S::SD; i++) {
            if constexpr (LSRC)
                L.src[k][i] = lsrc[k * S::SD + i];
            else
                L.src[k][i] = *(const uint32_t *)(J.srcp + (size_t)k * J.cw + 4 * i);
        }
    const int dx0 = -R + 4 * g;
    const int px0 = J.lx + (J.cx >> 2) + J.pad + dx0; // padded column of candidate dx0
    const int py0 = J.ly + (J.cy >> 2) + J.pad - R;   // padded row of candidate dy = -R
    // whole window of the block inside the padded plane -> direct aligned reads
    const int wx_lo = J.lx + (J.cx >> 2) + J.pad - R, wx_hi = wx_lo + 2 * R + BW - 1;
    const bool fast = wx_lo >= 0 && wx_hi <= J.pw - 1 && py0 >= 0 && py0 + 2 * R + BH - 1 <= J.ph - 1;
    const int shift = px0 & 3;
#pragma unroll
    for (int k = 0; k < 4; k++) L.rxb[k] = mvbits(J.cx + 4 * (dx0 + k) - J.mpx);
#pragma unroll
    for (int sl = 0; sl < BH; sl++) L.acc[sl] = 0;
    L.best = 0xffffffffu;
    L.bidx = 0xffffffffu;
    const int NJ = 2 * R + BH;
    if ((2 * R) % BH == 0) {
        // window height NJ is a whole number of ring blocks: peel the first and last
        fs_rows<BW, BH, 1>(L, J.F, J.pw, J.ph, fast, px0, py0, shift, 0, NJ, R, NC, g, J.cy, J.mpy, s_lr);
        for (int jb = BH; jb < NJ - BH; jb += BH)
            fs_rows<BW, BH, 0>(L, J.F, J.pw, J.ph, fast, px0, py0, shift, jb, NJ, R, NC, g, J.cy, J.mpy, s_lr);
        fs_rows<BW, BH, 2>(L, J.F, J.pw, J.ph, fast, px0, py0, shift, NJ - BH, NJ, R, NC, g, J.cy, J.mpy, s_lr);
    } else {
        for (int jb = 0; jb < NJ; jb += BH)
            fs_rows<BW, BH, 0>(L, J.F, J.pw, J.ph, fast, px0, py0, shift, jb, NJ, R, NC, g, J.cy, J.mpy, s_lr);
    }
    return ((uint64_t)L.best << 32) | L.bidx;
}

// Segmented wave argmin of the lanes' keys (lanes of one block are contiguous),
// then one 64-bit atomicMin per block segment: the global first raster minimum.
__device__ __forceinline__ void fs_commit(uint64_t key, long seg, bool active, uint64_t *slot)
{
    const int lane = threadIdx.x & 63;
    key = active ? key : ~0ull;
    const int segi = active ? (int)(seg & 0x3fffffff) : 0x7fffffff;
    key = wave_seg_min(key, segi, lane);
    const int prev = __shfl_up(segi, 1, 64);
    if (active && (lane == 0 || prev != segi)) atomicMin((unsigned long long *)slot, (unsigned long long)key);
}

// The same search with v_sad_u8 (4 abs-diff-adds per lane-op at full issue rate):
// a lane owns NQ (1, 2 or 4) candidates at ABSOLUTE padded columns base..base+NQ-1,
// base = (left edge of the window & ~3) + NQ*g, so each candidate's reference
// dwords are the lane's aligned row dwords realigned by a byte shift in 0..3.
// Accumulators are u32, one per (candidate row in flight, candidate).
template <int BW, int BH, int NQ> struct FsLane8 {
    uint32_t src[BH][FsShape<BW, BH>::SD];
    uint32_t acc[BH][NQ];
    uint32_t pre[FsShape<BW, BH>::SD + 1]; // aligned dwords of the current reference row
    uint32_t best, bidx;
    int rxb[NQ];
    int c0;  // column index (0..2R) of the lane's first candidate, may be out of range
    int sh;  // byte shift of the first candidate inside its aligned dword (0 or 2 for NQ=2)
};

template <int BW, int BH, int NQ>
__device__ __forceinline__ void fs8_load_row(FsLane8<BW, BH, NQ> &L, const uint8_t *__restrict__ F, int pw, int ph,
                                             bool fast, int abase, int ry)
{
    using S = FsShape<BW, BH>;
    if (fast) {
        const uint32_t *rp = (const uint32_t *)(F + (size_t)ry * pw + abase);
#pragma unroll
        for (int i = 0; i < S::SD + 1; i++) L.pre[i] = rp[i];
    } else {
#pragma unroll
        for (int i = 0; i < S::SD + 1; i++) L.pre[i] = load4_clamped(F, pw, ph, abase + 4 * i, ry);
    }
}

template <int BW, int BH, int NQ, int MODE>
__device__ __forceinline__ void fs8_rows(FsLane8<BW, BH, NQ> &L, const uint8_t *__restrict__ F, int pw, int ph,
                                         bool fast, int abase, int py0, int jb, int NJ, int R, int NC, int cy,
                                         int mpy, const int32_t *s_lr)
{
    using S = FsShape<BW, BH>;
#pragma unroll
    for (int jj = 0; jj < BH; jj++) {
        const int j = jb + jj;
        if (MODE == 0 && j >= NJ) break;
        uint32_t a[S::SD + 1];
        fs8_load_row<BW, BH, NQ>(L, F, pw, ph, fast, abase, py0 + j);
#pragma unroll
        for (int i = 0; i < S::SD + 1; i++) a[i] = L.pre[i];
        uint32_t w[NQ][S::SD];
#pragma unroll
        for (int q = 0; q < NQ; q++)
#pragma unroll
            for (int i = 0; i < S::SD; i++)
                w[q][i] = NQ == 4 ? (q == 0 ? a[i] : __builtin_amdgcn_alignbyte(a[i + 1], a[i], q))
                                  : __builtin_amdgcn_alignbyte(a[i + 1], a[i], L.sh + q);
#pragma unroll
        for (int k = 0; k < BH; k++) {
            if (MODE == 1 && k > jj) continue;
            if (MODE == 2 && k < jj) continue;
            const int slot = (jj - k + BH) % BH;
#pragma unroll
            for (int q = 0; q < NQ; q++)
#pragma unroll
                for (int i = 0; i < S::SD; i++)
                    L.acc[slot][q] = __builtin_amdgcn_sad_u8(L.src[k][i], w[q][i], L.acc[slot][q]);
        }
        const int t = j - (BH - 1);
        const int sd = (jj + 1) % BH;
        if (t >= 0) {
            const int ry_bits = mvbits(cy + 4 * (t - R) - mpy);
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int c = L.c0 + q;
                const uint32_t cost = L.acc[sd][q] + (uint32_t)s_lr[L.rxb[q] + ry_bits];
                if ((unsigned)c < (unsigned)NC && cost < L.best) {
                    L.best = cost;
                    L.bidx = (uint32_t)(t * NC + c);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) L.acc[sd][q] = 0;
    }
}

template <int NQ> __host__ __device__ constexpr int fs8_groups(int NC) { return (NC + 3 + NQ - 1) / NQ; }

template <int BW, int BH, int NQ>
__device__ __forceinline__ uint64_t fs_sad8_key(const FsJob &J, const int32_t *s_lr)
{
    using S = FsShape<BW, BH>;
    const int R = J.R, NC = 2 * R + 1;
    FsLane8<BW, BH, NQ> L;
#pragma unroll
    for (int k = 0; k < BH; k++)
#pragma unroll
        for (int i = 0; i < S::SD; i++) L.src[k][i] = *(const uint32_t *)(J.srcp + (size_t)k * J.cw + 4 * i);
    const int wx_lo = J.lx + (J.cx >> 2) + J.pad - R, wx_hi = wx_lo + 2 * R + BW - 1;
    const int py0 = J.ly + (J.cy >> 2) + J.pad - R;
    const int base = (wx_lo & ~3) + NQ * J.g; // absolute padded column of the lane's first candidate
    const int abase = base & ~3;
    L.sh = base & 3;
    L.c0 = base - wx_lo;
    const bool fast = wx_lo >= 0 && wx_hi <= J.pw - 1 && py0 >= 0 && py0 + 2 * R + BH - 1 <= J.ph - 1;
#pragma unroll
    for (int q = 0; q < NQ; q++) L.rxb[q] = mvbits(J.cx + 4 * (L.c0 + q - R) - J.mpx);
#pragma unroll
    for (int sl = 0; sl < BH; sl++)
#pragma unroll
        for (int q = 0; q < NQ; q++) L.acc[sl][q] = 0;
    L.best = 0xffffffffu;
    L.bidx = 0xffffffffu;
    const int NJ = 2 * R + BH;
    if ((2 * R) % BH == 0) {
        fs8_rows<BW, BH, NQ, 1>(L, J.F, J.pw, J.ph, fast, abase, py0, 0, NJ, R, NC, J.cy, J.mpy, s_lr);
        for (int jb = BH; jb < NJ - BH; jb += BH)
            fs8_rows<BW, BH, NQ, 0>(L, J.F, J.pw, J.ph, fast, abase, py0, jb, NJ, R, NC, J.cy, J.mpy, s_lr);
        fs8_rows<BW, BH, NQ, 2>(L, J.F, J.pw, J.ph, fast, abase, py0, NJ - BH, NJ, R, NC, J.cy, J.mpy, s_lr);
    } else {
        for (int jb = 0; jb < NJ; jb += BH)
            fs8_rows<BW, BH, NQ, 0>(L, J.F, J.pw, J.ph, fast, abase, py0, jb, NJ, R, NC, J.cy, J.mpy, s_lr);
    }
    return ((uint64_t)L.best << 32) | L.bidx;
}

// Lanes per block of each integer-search variant (V: 0 quad-SAD, 1 v_sad_u8 x2, 2 x4,
// 3 v_sad_u8 with one candidate per lane: the most lanes, for short block lists).
template <int V> __host__ __device__ constexpr int fs_groups(int NC)
{
    return V == 0 ? (NC + 3) / 4 : (V == 1 ? fs8_groups<2>(NC) : (V == 2 ? fs8_groups<4>(NC) : fs8_groups<1>(NC)));
}

template <int BW, int BH, int V>
__device__ __forceinline__ uint64_t fs_key(const FsJob &J, const int32_t *s_lr)
{
    if (V == 0) return fs_qsad_key<BW, BH>(J, s_lr);
    if (V == 1) return fs_sad8_key<BW, BH, 2>(J, s_lr);
    if (V == 2) return fs_sad8_key<BW, BH, 4>(J, s_lr);
    return fs_sad8_key<BW, BH, 1>(J, s_lr);
}

// Single-stream integer search: nframes frames of one geometry (blockIdx.y = frame).
template <int BW, int BH, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RH264_FS_WAVES)))
void k_full_search(rh264_search_params p, const uint8_t *__restrict__ cur, size_t cur_stride,
                   const uint8_t *__restrict__ planes, size_t planes_stride, const int32_t *__restrict__ centres,
                   const int32_t *__restrict__ mvps, uint64_t *__restrict__ keys)
{
    __shared__ int32_t s_lr[RH264_LR_N];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr_fp[threadIdx.x];
    __syncthreads();
    const int NG = fs_groups<V>(2 * p.range + 1);
    const int nbx = p.cw / BW, nblk = nbx * (p.ch / BH);
    unsigned wgx, wgy;
    xcd_grid2(wgx, wgy);
    const int frame = (int)wgy;
    const long gi = (long)wgx * 256 + threadIdx.x;
    const int b = (int)(gi / NG);
    const bool active = b < nblk;
    const int bs = active ? b : nblk - 1;
    const size_t boff = (size_t)frame * nblk + bs;
    FsJob J;
    J.cw = p.cw;
    J.pw = p.cw + 2 * p.pad;
    J.ph = p.ch + 2 * p.pad;
    J.pad = p.pad;
    J.R = p.range;
    J.g = (int)(gi - (long)b * NG);
    J.lx = (bs % nbx) * BW;
    J.ly = (bs / nbx) * BH;
    J.cx = centres ? centres[2 * boff] : 0;
    J.cy = centres ? centres[2 * boff + 1] : 0;
    J.mpx = mvps ? mvps[2 * boff] : 0;
    J.mpy = mvps ? mvps[2 * boff + 1] : 0;
    J.F = planes + (size_t)frame * planes_stride;
    J.srcp = cur + (size_t)frame * cur_stride + (size_t)J.ly * p.cw + J.lx;
    const uint64_t key = fs_key<BW, BH, V>(J, s_lr);
    fs_commit(key, b, active, keys + boff);
}

// ---------------------------------------------------------------------------
// sub-pel refinement
// ---------------------------------------------------------------------------
//
// A group of LPB = BW*BH/4 lanes owns one block: lane (4x4 block bi, row r)
// holds one source row of four pixels.  A candidate costs one or two 4-byte
// plane reads per lane, the row butterfly in registers, the column butterfly
// across the quad (DPP quad_perm), |.| and a group reduction.

template <int LPB> __device__ __forceinline__ int group_sum(int v)
{
    v += __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true); // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true); // quad_perm [2,3,0,1]
    if (LPB >= 8) v += __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true); // row_half_mirror
    if (LPB >= 16) v += __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true); // row_mirror
    if (LPB == 32) v += __shfl_xor(v, 16, 64);
    if (LPB == 64) // the group is the wave: four row sums, read into scalars
        v = __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) +
            __builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48);
    return v;
}

// sum |H d| of the 4x4 block whose row r this lane holds (d = 4 diffs).
__device__ __forceinline__ int hadamard_row_quad(int d0, int d1, int d2, int d3, int r)
{
    // row transform (hadamard_1d, transform.rs:631-633)
    int e0 = d0 + d1 + d2 + d3, e1 = d0 + d1 - d2 - d3;
    int e2 = d0 - d1 - d2 + d3, e3 = d0 - d1 + d2 - d3;
    // column transform across the quad: stage 1 pairs rows r, r^2; stage 2 r, r^1.
    const bool hi2 = (r & 2) != 0, odd = (r & 1) != 0;
    int y0 = __builtin_amdgcn_mov_dpp(e0, 0x4E, 0xF, 0xF, true);
    int y1 = __builtin_amdgcn_mov_dpp(e1, 0x4E, 0xF, 0xF, true);
    int y2 = __builtin_amdgcn_mov_dpp(e2, 0x4E, 0xF, 0xF, true);
    int y3 = __builtin_amdgcn_mov_dpp(e3, 0x4E, 0xF, 0xF, true);
    int s0 = hi2 ? y0 - e0 : e0 + y0, s1 = hi2 ? y1 - e1 : e1 + y1;
    int s2 = hi2 ? y2 - e2 : e2 + y2, s3 = hi2 ? y3 - e3 : e3 + y3;
    int z0 = __builtin_amdgcn_mov_dpp(s0, 0xB1, 0xF, 0xF, true);
    int z1 = __builtin_amdgcn_mov_dpp(s1, 0xB1, 0xF, 0xF, true);
    int z2 = __builtin_amdgcn_mov_dpp(s2, 0xB1, 0xF, 0xF, true);
    int z3 = __builtin_amdgcn_mov_dpp(s3, 0xB1, 0xF, 0xF, true);
    int c0 = odd ? z0 - s0 : s0 + z0, c1 = odd ? z1 - s1 : s1 + z1;
    int c2 = odd ? z2 - s2 : s2 + z2, c3 = odd ? z3 - s3 : s3 + z3;
    return abs(c0) + abs(c1) + abs(c2) + abs(c3);
}

struct RefineGeom {
    const uint8_t *pl; // plane set base
    size_t n;          // bytes per plane
    int pw, ph, pad;
};

// The refine walk reads one block's neighbourhood many times.  Each group stages
// the four planes around its start vector in LDS once: columns [ix-4, ix+BW+4),
// rows [iy-3, iy+BH+3) of the padded planes (ix, iy = padded integer position of
// the start).  Candidates whose integer part lies within [-3, +2] pels of the start
// on both axes are then served from LDS; the rest read the planes in HBM.
template <int BW, int BH> struct RefineWin {
    static constexpr int PITCH = BW + 8;      // bytes per staged row (multiple of 4)
    static constexpr int ROWS = BH + 6;
    static constexpr int PLANE = PITCH * ROWS;
    static constexpr int BYTES = 4 * PLANE;   // F, H, V, C
};

template <int BW, int BH>
__device__ __forceinline__ void stage_window(uint8_t *win, const RefineGeom &G, int wx0, int wy0, int li, int lpb)
{
    using W = RefineWin<BW, BH>;
    constexpr int DPR = W::PITCH / 4;       // dwords per staged row
    constexpr int TOTAL = 4 * W::ROWS * DPR;
    const bool inside = wx0 >= 0 && wy0 >= 0 && wx0 + W::PITCH <= G.pw && wy0 + W::ROWS <= G.ph;
    for (int i = li; i < TOTAL; i += lpb) {
        const int k = i / (W::ROWS * DPR);
        const int rem = i - k * (W::ROWS * DPR);
        const int row = rem / DPR, d = rem - row * DPR;
        const uint8_t *pl = G.pl + G.n * k;
        const int x = wx0 + 4 * d, y = wy0 + row;
        const uint32_t v = inside ? load4_fast(pl, G.pw, x, y) : load4_clamped(pl, G.pw, G.ph, x, y);
        *(uint32_t *)(win + k * W::PLANE + row * W::PITCH + 4 * d) = v;
    }
}

__device__ __forceinline__ uint32_t lds_load4(const uint8_t *rowp, int x)
{
    const uint32_t *p = (const uint32_t *)(rowp + (x & ~3));
    return __builtin_amdgcn_alignbyte(p[1], p[0], (uint32_t)(x & 3));
}

// Refinement with one lane per 4x4 sub-block: the lane keeps its 16 source pixels
// in four dwords, reads four prediction rows per candidate (one or two planes) and
// does the whole 4x4 Hadamard in registers; a block's NSB = (BW/4)*(BH/4) lanes
// then add their sums with DPP (no cross-lane butterflies).  Blocks of at least
// 8 sub-blocks can stage their +-3 pel neighbourhood in LDS (USE_LDS).
#ifndef RH264_LDS_MIN_NSB
#define RH264_LDS_MIN_NSB 8 // the batch's refinements stage windows in LDS from this many lanes per block
#endif
template <int BW, int BH> struct RefShape {
    static constexpr int NSB = (BW / 4) * (BH / 4); // lanes per block
    static constexpr int GPW = 256 / NSB;           // blocks per 256-thread workgroup
};

template <int LPB> __device__ __forceinline__ int group_sum_n(int v)
{
    if (LPB >= 2) v += __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    if (LPB >= 4) v += __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    if (LPB >= 8) v += __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true); // row_half_mirror
    if (LPB >= 16) v += __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true); // row_mirror
    return v;
}

// sum |H d| of a 4x4 block, d = s - p, rows as packed dwords (transform.rs:637-655),
// in packed 16-bit arithmetic (every intermediate fits: |d| <= 255,
// after the column butterflies <= 1020, after the first row stage <= 2040).  The
// source row is held as bytes (0,2) and (1,3) in the two halves of a dword; the
// column transform runs on both pairs at once, the row transform's first stage on
// pairs too, and its last stage uses |a + b| + |a - b| = 2 max(|a|, |b|).
typedef short pk16 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ pk16 as_pk(uint32_t v) { return __builtin_bit_cast(pk16, v); }
__device__ __forceinline__ uint32_t as_u32(pk16 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t bytes02(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x0c020c00u); }
__device__ __forceinline__ uint32_t bytes13(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x0c030c01u); }

struct SrcPk {
    uint32_t a[4], b[4]; // per row: bytes (0,2) and (1,3) as 16-bit halves
};

__device__ __forceinline__ SrcPk src_pk(const uint32_t s[4])
{
    SrcPk S;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        S.a[r] = bytes02(s[r]);
        S.b[r] = bytes13(s[r]);
    }
    return S;
}

__device__ __forceinline__ void hadamard4_pk(const pk16 x[4], pk16 e[4])
{
    const pk16 a = x[0] + x[1], b = x[2] + x[3], c = x[0] - x[1], d = x[2] - x[3];
    e[0] = a + b;
    e[1] = a - b;
    e[2] = c - d;
    e[3] = c + d;
}

__device__ __forceinline__ int satd4x4_pk(const SrcPk &S, const uint32_t p[4])
{
    pk16 x02[4], x13[4], P[4], Q[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        x02[r] = as_pk(S.a[r]) - as_pk(bytes02(p[r]));
        x13[r] = as_pk(S.b[r]) - as_pk(bytes13(p[r]));
    }
    hadamard4_pk(x02, P); // columns 0 and 2
    hadamard4_pk(x13, Q); // columns 1 and 3
    pk16 acc = {0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const pk16 u = P[k] + Q[k], v = P[k] - Q[k]; // (c0+c1, c2+c3), (c0-c1, c2-c3)
        const pk16 lo = as_pk(__builtin_amdgcn_perm(as_u32(v), as_u32(u), 0x05040100u)); // (u.x, v.x)
        const pk16 hi = as_pk(__builtin_amdgcn_perm(as_u32(v), as_u32(u), 0x07060302u)); // (u.y, v.y)
        const pk16 mx = __builtin_elementwise_max(lo, hi), mn = __builtin_elementwise_min(lo, hi);
        acc += __builtin_elementwise_max(mx, -mn); // max(|lo|, |hi|)
    }
    return 2 * ((int)acc.x + (int)acc.y);
}

// Candidate read from the staged window; di, dj in [-3, 2] (integer offset from the
// start vector), (sbx, sby) = the lane's sub-block.
template <int BW, int BH>
__device__ __forceinline__ int eval_lane_lds(const uint8_t *win, const SrcPk &s, int sbx, int sby, int di,
                                             int dj, int fx, int fy)
{
    using W = RefineWin<BW, BH>;
    const uint32_t q = qtab(fx * 4 + fy);
    const int x = di + 4 + sbx, y = dj + 3 + sby;
    const uint8_t *pa = win + (q & 15) * W::PLANE + (y + ((q >> 8) & 15)) * W::PITCH;
    const int xa = x + ((q >> 4) & 15);
    uint32_t p[4];
#pragma unroll
    for (int r = 0; r < 4; r++) p[r] = lds_load4(pa + r * W::PITCH, xa);
    if ((fx | fy) & 1) {
        const uint8_t *pb = win + ((q >> 12) & 15) * W::PLANE + (y + ((q >> 20) & 15)) * W::PITCH;
        const int xb = x + ((q >> 16) & 15);
#pragma unroll
        for (int r = 0; r < 4; r++) p[r] = avg4(p[r], lds_load4(pb + r * W::PITCH, xb));
    }
    return satd4x4_pk(s, p);
}

// Candidate read from the plane set in HBM (clamped at the plane border).
template <int BW, int BH>
__device__ __forceinline__ int eval_lane_global(const RefineGeom &G, const SrcPk &s, int lx, int ly, int sbx,
                                                int sby, int mvx, int mvy)
{
    const int bpx = lx + (mvx >> 2) + G.pad, bpy = ly + (mvy >> 2) + G.pad;
    const uint32_t q = qtab((mvx & 3) * 4 + (mvy & 3));
    const bool two = ((mvx | mvy) & 1) != 0;
    // acceptance guard of hpel_block / MeCtx (inter.rs:1030, 1075): +1 slack
    const bool inb = bpx >= 0 && bpy >= 0 && bpx + BW + 1 <= G.pw && bpy + BH + 1 <= G.ph;
    const uint8_t *pa = G.pl + G.n * (q & 15), *pb = G.pl + G.n * ((q >> 12) & 15);
    const int xa = bpx + sbx + ((q >> 4) & 15), ya = bpy + sby + ((q >> 8) & 15);
    const int xb = bpx + sbx + ((q >> 16) & 15), yb = bpy + sby + ((q >> 20) & 15);
    uint32_t p[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        p[r] = inb ? load4_fast(pa, G.pw, xa, ya + r) : load4_clamped(pa, G.pw, G.ph, xa, ya + r);
        if (two) p[r] = avg4(p[r], inb ? load4_fast(pb, G.pw, xb, yb + r) : load4_clamped(pb, G.pw, G.ph, xb, yb + r));
    }
    return satd4x4_pk(s, p);
}

// The ring walk of one block by its lane group (mb16.rs:3230-3275).
template <int BW, int BH, bool USE_LDS>
__device__ __forceinline__ void refine_walk(const RefineGeom &G, const uint8_t *win, const uint32_t s[4], int lx,
                                            int ly, int sbx, int sby, int sxi, int syi, int &bx, int &by, int mpx,
                                            int mpy, int pat, const int32_t *s_lr, int &best, int &evals)
{
    constexpr int NSB = RefShape<BW, BH>::NSB;
    const SrcPk S = src_pk(s);
    auto cost_of = [&](int mx, int my) -> int {
        const int di = (mx >> 2) - sxi, dj = (my >> 2) - syi;
        int v;
        if (USE_LDS && di >= -3 && di <= 2 && dj >= -3 && dj <= 2)
            v = eval_lane_lds<BW, BH>(win, S, sbx, sby, di, dj, mx & 3, my & 3);
        else
            v = eval_lane_global<BW, BH>(G, S, lx, ly, sbx, sby, mx, my);
        return group_sum_n<NSB>(v) + s_lr[mvbits(mx - mpx) + mvbits(my - mpy)];
    };
    best = cost_of(bx, by);
    evals = 1;
    if (pat == 4) {
        // quarter-pel window: the start first, then raster order
        const int sx = bx, sy = by;
        for (int dy = -2; dy <= 2; dy++)
            for (int dx = -2; dx <= 2; dx++) {
                if (dx == 0 && dy == 0) continue;
                const int cc = cost_of(sx + dx, sy + dy);
                evals++;
                if (cc < best) { best = cc; bx = sx + dx; by = sy + dy; }
            }
    } else if (pat >= 0) {
        const int nring = (pat & 1) ? 4 : 8;
        for (int step = 2; step >= 1; step--) {
            const int rx[8] = {step, -step, 0, 0, step, -step, step, -step};
            const int ry[8] = {0, 0, step, -step, step, -step, -step, step};
            // terminates: every extra pass strictly lowers a cost bounded below by 0
            for (;;) {
                bool improved = false;
                for (int i = 0; i < nring; i++) {
                    const int mx = bx + rx[i], my = by + ry[i];
                    const int cc = cost_of(mx, my);
                    evals++;
                    if (cc < best) { best = cc; bx = mx; by = my; improved = true; }
                }
                if (!improved || (pat & 2)) break;
            }
        }
    }
}

template <int BW, int BH> __device__ __forceinline__ void load_src4(const uint8_t *p, int stride, uint32_t s[4])
{
#pragma unroll
    for (int r = 0; r < 4; r++) s[r] = *(const uint32_t *)(p + (size_t)r * stride);
}

template <int BW, int BH, bool USE_LDS>
__global__ __launch_bounds__(256) void k_subpel_refine(rh264_search_params p,
                                                       const uint8_t *__restrict__ cur,
                                                       size_t cur_stride,
                                                       const uint8_t *__restrict__ planes,
                                                       size_t planes_stride,
                                                       const uint64_t *__restrict__ keys,
                                                       const int32_t *__restrict__ centres,
                                                       const int32_t *__restrict__ start_mv,
                                                       const int32_t *__restrict__ mvps,
                                                       int32_t *__restrict__ int_mv,
                                                       int32_t *__restrict__ int_cost,
                                                       int32_t *__restrict__ out_mv,
                                                       int32_t *__restrict__ out_cost,
                                                       int32_t *__restrict__ out_evals)
{
    constexpr int NSB = RefShape<BW, BH>::NSB, GPW = RefShape<BW, BH>::GPW;
    using W = RefineWin<BW, BH>;
    __shared__ int32_t s_lr[RH264_LR_N];
    __shared__ __align__(16) uint8_t s_win[USE_LDS ? GPW * W::BYTES : 16];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr[threadIdx.x];

    const int nbx = p.cw / BW, nby = p.ch / BH, nblk = nbx * nby;
    // workgroups renumbered per XCD: consecutive block strips share one L2
    unsigned wgx, wgy;
    xcd_grid2(wgx, wgy);
    const int frame = (int)wgy;
    const int grp = threadIdx.x / NSB, li = threadIdx.x % NSB;
    const int b0 = (int)wgx * GPW + grp;
    const bool active = b0 < nblk;
    const int b = active ? b0 : nblk - 1;
    const size_t boff = (size_t)frame * nblk + b;
    const int lx = (b % nbx) * BW, ly = (b / nbx) * BH;
    const int sbx = 4 * (li % (BW / 4)), sby = 4 * (li / (BW / 4));

    RefineGeom G;
    G.pw = p.cw + 2 * p.pad;
    G.ph = p.ch + 2 * p.pad;
    G.pad = p.pad;
    G.n = (size_t)G.pw * G.ph;
    G.pl = planes + (size_t)frame * planes_stride;
    uint32_t s[4];
    load_src4<BW, BH>(cur + (size_t)frame * cur_stride + (size_t)(ly + sby) * p.cw + lx + sbx, p.cw, s);
    const int mpx = mvps ? mvps[2 * boff] : 0, mpy = mvps ? mvps[2 * boff + 1] : 0;

    int bx, by;
    if (keys) {
        const uint64_t key = keys[boff];
        const int NC = 2 * p.range + 1;
        const uint32_t idx = (uint32_t)key;
        const int t = (int)(idx / (uint32_t)NC), c = (int)(idx % (uint32_t)NC);
        const int cx = centres ? centres[2 * boff] : 0, cy = centres ? centres[2 * boff + 1] : 0;
        bx = cx + 4 * (c - p.range);
        by = cy + 4 * (t - p.range);
        if (li == 0 && active) {
            if (int_mv) { int_mv[2 * boff] = bx; int_mv[2 * boff + 1] = by; }
            if (int_cost) int_cost[boff] = (int32_t)(key >> 32);
        }
    } else {
        bx = start_mv[2 * boff];
        by = start_mv[2 * boff + 1];
    }

    uint8_t *win = s_win + (USE_LDS ? grp * W::BYTES : 0);
    const int sxi = bx >> 2, syi = by >> 2;
    if (USE_LDS) stage_window<BW, BH>(win, G, lx + sxi + G.pad - 4, ly + syi + G.pad - 3, li, NSB);
    __syncthreads();
    if (!active) return; // whole groups exit together

    int best, evals;
    refine_walk<BW, BH, USE_LDS>(G, win, s, lx, ly, sbx, sby, sxi, syi, bx, by, mpx, mpy, p.subpel_pat, s_lr, best,
                                 evals);
    if (li == 0) {
        out_mv[2 * boff] = bx;
        out_mv[2 * boff + 1] = by;
        out_cost[boff] = best;
        if (out_evals) out_evals[boff] = evals;
    }
}

// ---------------------------------------------------------------------------
// sub-pel refinement of 16x16 blocks on the matrix cores
// ---------------------------------------------------------------------------
//
// One wave per block; the ring walk's control flow is wave-uniform (scalar).  The
// sixteen 4x4 Hadamards of a candidate are two 16x16x16 f16 matrix products with
// FP32 accumulation:  X = P * Hb (row transforms), Y = Hb * X (column transforms),
// Hb = diag(H4, H4, H4, H4), H4 the symmetric Sylvester matrix (a row order of the
// reference's hadamard_1d, transform.rs:631-655; sum |.| does not depend on it).
// Every value is an integer that fp16 / fp32 hold exactly: P enters as p + 1024
// (fp16 spacing is 1 on [1024, 2048)), the accumulator starts at -4096 on the DC
// columns so X = P * Hb with |X| <= 1020 < 2048, and the second product starts at
// -Hb S Hb (the source block's transform, computed once), so Y = Hb (P - S) Hb with
// |Y| <= 8160.  sum |Y| is the encoder-domain SATD (mb16.rs:6027-6074).
//
// Lane l holds pixel row l%16, columns 4(l/16)..+3 of the block (the A operand of
// the first product); the first product's result is already in the layout of the
// second one's B operand.  Each wave stages the four planes around its start
// vector in LDS (integer offsets -4..+3 of the start on both axes, i.e. quarter-pel
// offsets -16..+15); farther candidates read the planes in HBM.  Per-candidate
// tables live in the lanes of a VGPR and are read with v_readlane: the operand
// offsets of the 16 quarter phases, the mv-bit counts of the offsets -32..+31
// from the start on each axis, and the 128-entry lambda-rate table.

typedef _Float16 mx_h4 __attribute__((ext_vector_type(4)));
typedef float mx_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t mx_u2 __attribute__((ext_vector_type(2)));

constexpr int MX_WL = 4;                 // window reach in full pels
constexpr int MX_ROWS = 2 * MX_WL + 16;  // staged rows per plane
constexpr int MX_DW = 7;                 // staged dwords per row (28 bytes)
constexpr int MX_PITCH = 40;             // row pitch in bytes (10 dwords: conflict-free reads)
constexpr int MX_PLANE = MX_ROWS * MX_PITCH;
constexpr int MX_WIN = 4 * MX_PLANE;     // bytes per wave

// The quarter-phase operand table (qtab) as byte offsets inside a staged window:
// plane * MX_PLANE + row offset * MX_PITCH + column offset, operand a in the low
// half, operand b in the high half.
constexpr uint32_t mxoff(const QPhase &q)
{
    return (uint32_t)(q.pa * MX_PLANE + q.ay * MX_PITCH + q.ax) | ((uint32_t)(q.pb * MX_PLANE + q.by * MX_PITCH + q.bx) << 16);
}
__constant__ uint32_t c_mxoff[16] = {
    mxoff(k_qphase[0]),  mxoff(k_qphase[1]),  mxoff(k_qphase[2]),  mxoff(k_qphase[3]),
    mxoff(k_qphase[4]),  mxoff(k_qphase[5]),  mxoff(k_qphase[6]),  mxoff(k_qphase[7]),
    mxoff(k_qphase[8]),  mxoff(k_qphase[9]),  mxoff(k_qphase[10]), mxoff(k_qphase[11]),
    mxoff(k_qphase[12]), mxoff(k_qphase[13]), mxoff(k_qphase[14]), mxoff(k_qphase[15]),
};

struct MxConst {
    mx_h4 hb; // Hb entries of this lane: B operand of the first product, A operand of the second
    mx_f4 c1; // accumulator start of the first product
};

// T8: Hb = diag(H8, H8) (the 8x8 transform, build-defined): |X| <= 2040 < 2048 and
// |Y| <= 16320, still exact; the DC columns of the first product start at -8192.
template <bool T8 = false> __device__ __forceinline__ MxConst mx_const(int lane)
{
    MxConst K;
    const int q = lane >> 4, n = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k = 4 * q + i; // row of Hb
        // Hb[k][n]: nonzero inside the diagonal blocks, H[k][n] = (-1)^popc(k & n) within a block
        const float h = T8 ? (((k >> 3) == (n >> 3)) ? ((__popc(k & n & 7) & 1) ? -1.0f : 1.0f) : 0.0f)
                           : ((q == (n >> 2)) ? ((__popc(i & n & 3) & 1) ? -1.0f : 1.0f) : 0.0f);
        K.hb[i] = (_Float16)h;
        K.c1[i] = T8 ? ((n & 7) == 0 ? -8192.0f : 0.0f) : ((n & 3) == 0 ? -4096.0f : 0.0f);
    }
    return K;
}

// Sum over the wave, returned as a scalar: row sums with DPP, then rows 0+1 and
// 2+3 with row_bcast:15, the total into lane 63 with row_bcast:31.
__device__ __forceinline__ int mx_wave_sum(int v)
{
    v += __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true); // row_half_mirror
    v += __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true); // row_mirror
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false); // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false); // row_bcast:31 into rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}

// Hb (P - S) Hb, P = the lane's 4 predicted pixels, c2 = -(Hb S Hb) in accumulator
// layout (zero to transform S itself).  Returns the lane's four coefficients.
__device__ __forceinline__ mx_f4 mx_transform(uint32_t p, const MxConst &K, mx_f4 c2)
{
    mx_u2 a2;
    a2[0] = __builtin_amdgcn_perm(0x64646464u, p, 0x04010400u); // (1024 + p0, 1024 + p1) as fp16
    a2[1] = __builtin_amdgcn_perm(0x64646464u, p, 0x04030402u); // (1024 + p2, 1024 + p3)
    const mx_h4 a = __builtin_bit_cast(mx_h4, a2);
    const mx_f4 x = __builtin_amdgcn_mfma_f32_16x16x16f16(a, K.hb, K.c1, 0, 0, 0);
    mx_u2 b2;
    b2[0] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(x[0], x[1]));
    b2[1] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(x[2], x[3]));
    const mx_h4 b = __builtin_bit_cast(mx_h4, b2);
    return __builtin_amdgcn_mfma_f32_16x16x16f16(K.hb, b, c2, 0, 0, 0);
}

// Minimum over the wave, returned as a scalar: the same DPP steps as mx_wave_sum,
// each lane keeping its own value where a step has no source.
__device__ __forceinline__ uint32_t mx_wave_min(uint32_t v)
{
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x141, 0xF, 0xF, false)); // row_half_mirror
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xF, 0xF, false)); // row_mirror
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x142, 0xA, 0xF, false)); // row_bcast:15
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x143, 0xC, 0xF, false)); // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// SATD of the candidate: the wave's sum |Y|, halved for the 8x8 transform (exact:
// the sum over each 8x8 block is even)
template <bool T8 = false> __device__ __forceinline__ int mx_satd(uint32_t p, const MxConst &K, mx_f4 c2)
{
    const mx_f4 y = mx_transform(p, K, c2);
    const float s = (fabsf(y[0]) + fabsf(y[1])) + (fabsf(y[2]) + fabsf(y[3]));
    const int t = mx_wave_sum((int)s);
    return T8 ? t >> 1 : t;
}

// 4 bytes at a dword-aligned LDS address plus a byte shift of 0..3.
__device__ __forceinline__ uint32_t mx_lds4(const uint8_t *p4, int shift)
{
    const uint32_t *p = (const uint32_t *)p4;
    return __builtin_amdgcn_alignbyte(p[1], p[0], (uint32_t)shift);
}

// Wave-uniform state of one block's refinement.
struct MxBlock {
    const uint8_t *pl;  // plane set
    int pw, ph;
    size_t pn;          // bytes per plane
    int ix, iy;         // padded integer position of the start vector
    int sx, sy;         // start vector (quarter pel, multiple of 4)
    int mpx, mpy;       // rate predictor
    int xoff;           // byte offset of column ix - MX_WL inside the staged rows
};

// The lane's four predicted pixels of the candidate at offset (rx, ry) from the
// start's integer position, read from the plane set in HBM, clamped at its border.
__device__ __forceinline__ uint32_t mx_operand_far(const MxBlock &B, int lane, int rx, int ry)
{
    const int r = lane & 15, qd = lane >> 4;
    const int fx = rx & 3, fy = ry & 3;
    const uint32_t q = qtab(fx * 4 + fy);
    const int bpx = B.ix + (rx >> 2), bpy = B.iy + (ry >> 2);
    const bool inb = bpx >= 0 && bpy >= 0 && bpx + 17 <= B.pw && bpy + 17 <= B.ph;
    const uint8_t *plA = B.pl + B.pn * (q & 15);
    const int xa = bpx + ((q >> 4) & 15) + 4 * qd, ya = bpy + ((q >> 8) & 15) + r;
    uint32_t v = inb ? load4_fast(plA, B.pw, xa, ya) : load4_clamped(plA, B.pw, B.ph, xa, ya);
    if ((fx | fy) & 1) {
        const uint8_t *plB = B.pl + B.pn * ((q >> 12) & 15);
        const int xb = bpx + ((q >> 16) & 15) + 4 * qd, yb = bpy + ((q >> 20) & 15) + r;
        const uint32_t w = inb ? load4_fast(plB, B.pw, xb, yb) : load4_clamped(plB, B.pw, B.ph, xb, yb);
        v = __builtin_amdgcn_lerp(v, w, 0x01010101u);
    }
    return v;
}

// Cost of a candidate outside the staged window: the plane set in HBM, clamped at
// its border, and the rate computed directly.  Kept out of line so that its table
// reads and address arithmetic stay off the common path.  Its arguments are plain
// values (the block's fields, the transform constants rebuilt from the lane), so the
// caller keeps nothing in scratch memory for it.
template <bool T8>
__device__ __attribute__((noinline)) int mx_cost_far(const uint8_t *pl, int pw, int ph, int ix, int iy, int sx, int sy,
                                                      int mpx, int mpy, mx_f4 c2, const int32_t *s_lr, int lane,
                                                      int rx, int ry)
{
    MxBlock B;
    B.pl = pl;
    B.pw = pw;
    B.ph = ph;
    B.pn = (size_t)pw * ph;
    B.ix = ix;
    B.iy = iy;
    B.sx = sx;
    B.sy = sy;
    B.mpx = mpx;
    B.mpy = mpy;
    B.xoff = 0;
    const MxConst K = mx_const<T8>(lane);
    const int s = mx_satd<T8>(mx_operand_far(B, lane, rx, ry), K, c2);
    return __builtin_amdgcn_readfirstlane(s + s_lr[mvbits(sx + rx - mpx) + mvbits(sy + ry - mpy)]);
}

template <int PAT, bool T8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_refine_mx16(rh264_search_params p, const uint8_t *__restrict__ cur,
                                                     size_t cur_stride, const uint8_t *__restrict__ planes,
                                                     size_t planes_stride, const uint64_t *__restrict__ keys,
                                                     const int32_t *__restrict__ centres,
                                                     const int32_t *__restrict__ start_mv,
                                                     const int32_t *__restrict__ mvps, int32_t *__restrict__ int_mv,
                                                     int32_t *__restrict__ int_cost, int32_t *__restrict__ out_mv,
                                                     int32_t *__restrict__ out_cost, int32_t *__restrict__ out_evals)
{
    __shared__ int32_t s_lr[RH264_LR_N];
    __shared__ __align__(16) uint8_t s_win[4 * MX_WIN];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr[threadIdx.x];
    __syncthreads();

    const int nbx = p.cw >> 4, nblk = nbx * (p.ch >> 4);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    unsigned wgx, wgy;
    xcd_grid2(wgx, wgy);
    const int b = (int)wgx * 4 + wave; // wave-uniform
    if (b >= nblk) return;
    const int frame = (int)wgy;
    const size_t boff = (size_t)frame * nblk + b;
    const int lx = (b % nbx) << 4, ly = (b / nbx) << 4;
    const int r = lane & 15, qd = lane >> 4; // pixel row, column quad of this lane

    MxBlock B;
    B.pw = p.cw + 2 * p.pad;
    B.ph = p.ch + 2 * p.pad;
    B.pn = (size_t)B.pw * B.ph;
    B.pl = planes + (size_t)frame * planes_stride;
    B.mpx = __builtin_amdgcn_readfirstlane(mvps ? mvps[2 * boff] : 0);
    B.mpy = __builtin_amdgcn_readfirstlane(mvps ? mvps[2 * boff + 1] : 0);
    const uint32_t src = *(const uint32_t *)(cur + (size_t)frame * cur_stride + (size_t)(ly + r) * p.cw + lx + 4 * qd);

    int bx, by;
    if (keys) {
        const uint64_t key = keys[boff];
        const int NC = 2 * p.range + 1;
        const uint32_t idx = __builtin_amdgcn_readfirstlane((uint32_t)key);
        const int t = (int)(idx / (uint32_t)NC), c = (int)(idx % (uint32_t)NC);
        const int cx = __builtin_amdgcn_readfirstlane(centres ? centres[2 * boff] : 0);
        const int cy = __builtin_amdgcn_readfirstlane(centres ? centres[2 * boff + 1] : 0);
        bx = cx + 4 * (c - p.range);
        by = cy + 4 * (t - p.range);
        if (lane == 0) {
            if (int_mv) { int_mv[2 * boff] = bx; int_mv[2 * boff + 1] = by; }
            if (int_cost) int_cost[boff] = (int32_t)(key >> 32);
        }
    } else {
        bx = __builtin_amdgcn_readfirstlane(start_mv[2 * boff]);
        by = __builtin_amdgcn_readfirstlane(start_mv[2 * boff + 1]);
    }
    // the walk runs on offsets (rx, ry) from S = the start's integer position, so
    // the phase of a candidate is (rx & 3, ry & 3) and its integer offset (rx >> 2, ry >> 2)
    const int fx0 = bx & 3, fy0 = by & 3;
    B.sx = bx - fx0;
    B.sy = by - fy0;
    B.ix = lx + (B.sx >> 2) + p.pad;
    B.iy = ly + (B.sy >> 2) + p.pad;

    // stage the window: planes F, H, V, C, rows iy-4 .. iy+19, columns from wxa
    // (ix-4 rounded down to a dword) for 28 bytes
    const int wx0 = B.ix - MX_WL, wy0 = B.iy - MX_WL;
    const int wxa = wx0 & ~3;
    B.xoff = wx0 - wxa;
    uint8_t *win = s_win + wave * MX_WIN;
    {
        // lane -> (row in a group of 9, dword of the row); lane 63 idles
        const int srow = lane / MX_DW, sd = lane - srow * MX_DW;
        const bool inside = wxa >= 0 && wy0 >= 0 && wxa + 4 * MX_DW <= B.pw && wy0 + MX_ROWS <= B.ph;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint8_t *plk = B.pl + B.pn * k;
#pragma unroll
            for (int r0 = 0; r0 < MX_ROWS; r0 += 64 / MX_DW) {
                const int row = r0 + srow;
                if (srow < 64 / MX_DW && row < MX_ROWS) {
                    const int x = wxa + 4 * sd, y = wy0 + row;
                    const uint32_t v = inside ? *(const uint32_t *)(plk + (size_t)y * B.pw + x)
                                              : load4_clamped(plk, B.pw, B.ph, x, y);
                    *(uint32_t *)(win + k * MX_PLANE + row * MX_PITCH + 4 * sd) = v;
                }
            }
        }
    }
    // lane tables: operand offsets per phase, mv bits per offset -32..31, lambda-rate
    const uint32_t t_off = c_mxoff[lane & 15];
    const int t_bx = mvbits(B.sx + lane - 32 - B.mpx), t_by = mvbits(B.sy + lane - 32 - B.mpy);
    const int t_lr = s_lr[lane];
    // the rate table covers every in-window candidate when the largest bit count
    // there (at an end of each axis' offset range, mvbits is monotone in |d|) is < 64
    const bool rate_fast = max(mvbits(B.sx - 16 - B.mpx), mvbits(B.sx + 15 - B.mpx)) +
                               max(mvbits(B.sy - 16 - B.mpy), mvbits(B.sy + 15 - B.mpy)) < 64;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const uint8_t *lwin = win + r * MX_PITCH + 4 * qd; // this lane's row and column quad
    const MxConst K = mx_const<T8>(lane);
    mx_f4 c2 = -mx_transform(src, K, mx_f4{0.0f, 0.0f, 0.0f, 0.0f});

    auto cost_of = [&](int rx, int ry) -> int {
        if (!rate_fast || (unsigned)(rx + 4 * MX_WL) >= (unsigned)(8 * MX_WL) ||
            (unsigned)(ry + 4 * MX_WL) >= (unsigned)(8 * MX_WL))
            return __builtin_amdgcn_readfirstlane(
                mx_cost_far<T8>(B.pl, B.pw, B.ph, B.ix, B.iy, B.sx, B.sy, B.mpx, B.mpy, c2, s_lr, lane, rx, ry));
        const int ph = ((rx & 3) << 2) | (ry & 3);
        const uint32_t qo = __builtin_amdgcn_readlane(t_off, ph);
        // byte offsets (wave-uniform) of the operands' top-left samples in the staged
        // window (one-plane phases list the same operand twice: lerp(a, a) = a); the
        // lane adds its own row and column quad
        const int o0 = ((ry >> 2) + MX_WL) * MX_PITCH + B.xoff + MX_WL + (rx >> 2);
        const int oa = o0 + (int)(qo & 0xffffu), ob = o0 + (int)(qo >> 16);
        const uint32_t v = __builtin_amdgcn_lerp(mx_lds4(lwin + (oa & ~3), oa & 3), mx_lds4(lwin + (ob & ~3), ob & 3),
                                                 0x01010101u);
        const int bits = __builtin_amdgcn_readlane(t_bx, rx + 32) + __builtin_amdgcn_readlane(t_by, ry + 32);
        return mx_satd<T8>(v, K, c2) + __builtin_amdgcn_readlane(t_lr, bits);
    };

    // Ring tables: lane i < 8 holds candidate i of the ring around the current centre
    // -- its two operands' dword offsets in the window, their byte shifts and its
    // lambda-rate cost, or -1 when the candidate lies outside the window -- built in
    // a few vector instructions per ring (and again after a move), so a candidate
    // costs four v_readlane instead of scalar address and rate arithmetic.
    const int li8 = lane & 7;
    const int lrx = (li8 < 2 || li8 >= 4) ? 1 - 2 * (li8 & 1) : 0;                      // RX[i]
    const int lry = li8 < 2 ? 0 : li8 < 4 ? 1 - 2 * (li8 & 1) : ((li8 & 1) == ((li8 >> 1) & 1) ? 1 : -1); // RY[i]
    struct Tab {
        uint32_t ta, tb, sh;
        int lr;
    };
    auto make_tab = [&](int ccx, int ccy, int step) -> Tab {
        const int rx = ccx + step * lrx, ry = ccy + step * lry;
        const bool inw = rate_fast && (unsigned)(rx + 4 * MX_WL) < (unsigned)(8 * MX_WL) &&
                         (unsigned)(ry + 4 * MX_WL) < (unsigned)(8 * MX_WL);
        const uint32_t qo = (uint32_t)__builtin_amdgcn_ds_bpermute((((rx & 3) << 2) | (ry & 3)) << 2, (int)t_off);
        const int o0 = ((ry >> 2) + MX_WL) * MX_PITCH + B.xoff + MX_WL + (rx >> 2);
        const int oa = o0 + (int)(qo & 0xffffu), ob = o0 + (int)(qo >> 16);
        const int bits = __builtin_amdgcn_ds_bpermute((rx + 32) << 2, t_bx) +
                         __builtin_amdgcn_ds_bpermute((ry + 32) << 2, t_by);
        const int lrc = __builtin_amdgcn_ds_bpermute(bits << 2, t_lr);
        Tab T;
        T.ta = (uint32_t)(oa & ~3);
        T.tb = (uint32_t)(ob & ~3);
        T.sh = (uint32_t)((oa & 3) | ((ob & 3) << 8));
        T.lr = inw ? lrc : -1;
        return T;
    };
    auto eval_tab = [&](const Tab &T, int i, int rx, int ry) -> int {
        const int lrc = __builtin_amdgcn_readlane(T.lr, i);
        if (lrc < 0)
            return __builtin_amdgcn_readfirstlane(
                mx_cost_far<T8>(B.pl, B.pw, B.ph, B.ix, B.iy, B.sx, B.sy, B.mpx, B.mpy, c2, s_lr, lane, rx, ry));
        const uint32_t ta = __builtin_amdgcn_readlane(T.ta, i), tb = __builtin_amdgcn_readlane(T.tb, i);
        const uint32_t sh = __builtin_amdgcn_readlane(T.sh, i);
        const uint32_t v = __builtin_amdgcn_lerp(mx_lds4(lwin + ta, sh & 3), mx_lds4(lwin + tb, (sh >> 8) & 3),
                                                 0x01010101u);
        return mx_satd<T8>(v, K, c2) + lrc;
    };

    int cx = fx0, cy = fy0; // current best, as an offset from S
    int best = cost_of(cx, cy);
    int evals = 1;
    if (PAT >= 0) {
        constexpr int NRING = (PAT & 1) ? 4 : 8;
        constexpr int RX[8] = {1, -1, 0, 0, 1, -1, 1, -1};
        constexpr int RY[8] = {0, 0, 1, -1, 1, -1, -1, 1};
#pragma unroll
        for (int step = 2; step >= 1; step--) {
            // terminates: every extra pass strictly lowers a cost bounded below by 0
            for (;;) {
                bool improved = false;
                Tab T = make_tab(cx, cy, step);
#pragma unroll
                for (int i = 0; i < NRING; i++) {
                    const int mx = cx + step * RX[i], my = cy + step * RY[i];
                    const int cc = eval_tab(T, i, mx, my);
                    evals++;
                    if (cc < best) {
                        best = cc;
                        cx = mx;
                        cy = my;
                        improved = true;
                        if (i + 1 < NRING) T = make_tab(cx, cy, step); // the rest of the pass moves with the centre
                    }
                }
                if (!improved || (PAT & 2)) break;
            }
        }
    }
    if (lane == 0) {
        out_mv[2 * boff] = B.sx + cx;
        out_mv[2 * boff + 1] = B.sy + cy;
        out_cost[boff] = best;
        if (out_evals) out_evals[boff] = evals;
    }
}

// ---------------------------------------------------------------------------
// Quarter-pel window refinement of 16x16 blocks (subpel_pat 4), one wave per
// block: the 25 positions start + (dx, dy), dx, dy in -2..2 quarter pels, each
// priced as the ring walk prices a candidate (SATD of the prediction + lambda
// rate); the cheapest wins, ties going to the start, then to the earlier position
// in raster order (dy, then dx).  The positions do not depend on each other, so:
//  * for an integer-pel start every position's phase and integer offset are
//    compile-time constants; its prediction is one of 16 operands, or the
//    rounded average of two (planes F, H, V, C at integer offsets -1 and 0 on
//    each axis), which each lane reads once with one 8-byte load per operand pair;
//  * each position's Hadamards are k_refine_mx16's two MFMAs, after which each
//    lane holds its partial sum of |coefficients| (exact in fp32);
//  * one butterfly over the 25 partials (two permlane swaps, four DPP stages)
//    leaves each lane with one position's whole SATD, so the rates and the
//    minimum are computed once, across the lanes.
// A fractional start (not produced by the integer search) prices the positions
// one by one from the plane set in HBM.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float wn_f(uint32_t v) { return __builtin_bit_cast(float, v); }
__device__ __forceinline__ uint32_t wn_u(float v) { return __builtin_bit_cast(uint32_t, v); }

// Butterfly stages.  Each takes a pair (a, b) of per-lane partials and returns a
// on the lanes where the stage's lane bit is clear, b where it is set, summed over
// the lane and its partner (the lane with that bit flipped; for the stage on bit
// 2, the partner l ^ 7 of row_half_mirror, which keeps bits 3..5).
__device__ __forceinline__ float wn_swap32(float a, float b) // lane bit 5
{
    const auto r = __builtin_amdgcn_permlane32_swap(wn_u(a), wn_u(b), false, false);
    return wn_f(r[0]) + wn_f(r[1]);
}
__device__ __forceinline__ float wn_swap16(float a, float b) // lane bit 4
{
    const auto r = __builtin_amdgcn_permlane16_swap(wn_u(a), wn_u(b), false, false);
    return wn_f(r[0]) + wn_f(r[1]);
}
template <int CTRL> __device__ __forceinline__ float wn_dpp(float a, float b, bool hi)
{
    const float keep = hi ? b : a, send = hi ? a : b;
    return keep + wn_f((uint32_t)__builtin_amdgcn_mov_dpp((int)wn_u(send), CTRL, 0xF, 0xF, true));
}

// Whole-wave sums of the partials t[0..24]: lane l ends with the sum of position
// wn_lane_pos(l) (bits 5, 4, 2, 0, 1 of l as bits 0..4 of the position; lanes
// whose position is >= 25 hold 0).  The partner sets of the six stages are the
// lane xors 32, 16, 7, 1, 2 and 8, which span all 64 lanes.
__device__ __forceinline__ float wn_sum25(const float (&t)[25], int lane)
{
    float u[13], w[7], x[4];
#pragma unroll
    for (int k = 0; k < 13; k++) u[k] = wn_swap32(t[2 * k], 2 * k + 1 < 25 ? t[2 * k + 1] : 0.0f);
#pragma unroll
    for (int k = 0; k < 7; k++) w[k] = wn_swap16(u[2 * k], 2 * k + 1 < 13 ? u[2 * k + 1] : 0.0f);
    const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4;
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = wn_dpp<0x141>(w[2 * k], 2 * k + 1 < 7 ? w[2 * k + 1] : 0.0f, b2); // row_half_mirror
    const float y0 = wn_dpp<0xB1>(x[0], x[1], b0), y1 = wn_dpp<0xB1>(x[2], x[3], b0); // quad_perm [1,0,3,2]
    const float z = wn_dpp<0x4E>(y0, y1, b1);                                          // quad_perm [2,3,0,1]
    return z + wn_f((uint32_t)__builtin_amdgcn_mov_dpp((int)wn_u(z), 0x128, 0xF, 0xF, true)); // row_ror:8
}
__device__ __forceinline__ int wn_lane_pos(int lane)
{
    return ((lane >> 5) & 1) | ((lane >> 3) & 2) | (lane & 4) | ((lane & 1) << 3) | ((lane & 2) << 3);
}

// Prediction of position C (dx = C % 5 - 2, dy = C / 5 - 2) around an integer
// start from the operands o[plane][y offset + 1][x offset + 1].
template <int C> __device__ __forceinline__ uint32_t wn_operand(const uint32_t (&o)[4][2][2])
{
    constexpr int dx = C % 5 - 2, dy = C / 5 - 2;
    constexpr QPhase q = k_qphase[(dx & 3) * 4 + (dy & 3)];
    constexpr int ox = dx < 0 ? 0 : 1, oy = dy < 0 ? 0 : 1; // integer offset -1 or 0, plus 1
    static_assert(ox + q.ax <= 1 && oy + q.ay <= 1 && ox + q.bx <= 1 && oy + q.by <= 1, "operand not loaded");
    const uint32_t a = o[q.pa][oy + q.ay][ox + q.ax];
    if constexpr (q.pa == q.pb && q.ax == q.bx && q.ay == q.by)
        return a;
    else
        return __builtin_amdgcn_lerp(a, o[q.pb][oy + q.by][ox + q.bx], 0x01010101u);
}
template <typename F, int... C>
__device__ __forceinline__ void wn_partials(float (&t)[25], const uint32_t (&o)[4][2][2], F &&partial,
                                            std::integer_sequence<int, C...>)
{
    ((t[C] = partial(wn_operand<C>(o))), ...);
}

// One block's +-2 quarter-pel window refinement by one wave: the 25 positions around
// the start (bx, by), SATD + rate against the predictor (mpx, mpy) priced with the
// table s_lr, the first minimum (the start first, then raster order); lane 0 writes
// out_mv / out_cost (/ out_evals) of block boff.  cur_f: the block's frame of the
// current picture (stride p.cw); pl: its reference plane set.
template <bool T8>
__device__ __forceinline__ void wn16_block(const rh264_search_params &p, const int32_t *s_lr,
                                           const uint8_t *__restrict__ cur_f, const uint8_t *__restrict__ pl, int lx,
                                           int ly, int bx, int by, int mpx, int mpy, int lane, size_t boff,
                                           int32_t *__restrict__ out_mv, int32_t *__restrict__ out_cost,
                                           int32_t *__restrict__ out_evals)
{
    const int r = lane & 15, qd = lane >> 4; // pixel row, column quad of this lane
    MxBlock B;
    B.pw = p.cw + 2 * p.pad;
    B.ph = p.ch + 2 * p.pad;
    B.pn = (size_t)B.pw * B.ph;
    B.pl = pl;
    B.mpx = mpx;
    B.mpy = mpy;
    const uint32_t src = *(const uint32_t *)(cur_f + (size_t)(ly + r) * p.cw + lx + 4 * qd);
    const int fx0 = bx & 3, fy0 = by & 3;
    B.sx = bx - fx0;
    B.sy = by - fy0;
    B.ix = lx + (B.sx >> 2) + p.pad;
    B.iy = ly + (B.sy >> 2) + p.pad;
    B.xoff = 0;

    const MxConst K = mx_const<T8>(lane);
    const mx_f4 c2 = -mx_transform(src, K, mx_f4{0.0f, 0.0f, 0.0f, 0.0f});
    int best, best_pos; // wave-uniform
    if ((fx0 | fy0) == 0) {
        // o[k][j][i]: plane k, row iy - 1 + j + r, columns ix - 1 + i + 4 qd .. +3
        uint32_t o[4][2][2];
        if (B.ix >= 1 && B.iy >= 1 && B.ix + 16 <= B.pw && B.iy + 16 <= B.ph) {
            // every used byte is inside the plane; the 8-byte load at the dword
            // below column ix - 1 + 4 qd ends at most 3 bytes past column ix + 15,
            // inside the buffer (the next row, the next plane or the tail)
            const int col = B.ix - 1 + 4 * qd;
            const uint32_t sh = (uint32_t)(B.ix - 1) & 3u; // the plane stride is a multiple of 4
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const uint8_t *row = B.pl + B.pn * k + (size_t)(B.iy - 1 + j + r) * B.pw;
                    const uint32_t *d = (const uint32_t *)(row + (col & ~3));
                    const uint32_t d0 = d[0], d1 = d[1];
                    o[k][j][0] = __builtin_amdgcn_alignbyte(d1, d0, sh);
                    o[k][j][1] = sh == 3u ? d1 : __builtin_amdgcn_alignbyte(d1, d0, sh + 1u);
                }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int i = 0; i < 2; i++)
                        o[k][j][i] = load4_clamped(B.pl + B.pn * k, B.pw, B.ph, B.ix - 1 + i + 4 * qd, B.iy - 1 + j + r);
        }
        float t[25];
        wn_partials(t, o,
                    [&](uint32_t v) -> float {
                        const mx_f4 y = mx_transform(v, K, c2);
                        return (fabsf(y[0]) + fabsf(y[1])) + (fabsf(y[2]) + fabsf(y[3]));
                    },
                    std::make_integer_sequence<int, 25>{});
        const int satd = T8 ? (int)wn_sum25(t, lane) >> 1 : (int)wn_sum25(t, lane);
        const int pos = wn_lane_pos(lane);
        const int dx = pos % 5 - 2, dy = pos / 5 - 2;
        const bool valid = pos < 25;
        const int rate = s_lr[valid ? mvbits(bx + dx - B.mpx) + mvbits(by + dy - B.mpy) : 0];
        // (cost, order): the start first, then raster order; cost < 2^22 + 2^26 (the
        // SATD bound and the lambda tables' bound, check_params) fits 27 bits
        const uint32_t key = valid ? ((uint32_t)(satd + rate) << 5) | (uint32_t)(pos == 12 ? 0 : pos + 1) : ~0u;
        const uint32_t kmin = mx_wave_min(key);
        best = (int)(kmin >> 5);
        const int ord = (int)(kmin & 31);
        best_pos = ord == 0 ? 12 : ord - 1;
    } else {
        auto cost_at = [&](int c) -> int {
            const int dx = c % 5 - 2, dy = c / 5 - 2;
            const int s = mx_satd<T8>(mx_operand_far(B, lane, fx0 + dx, fy0 + dy), K, c2);
            return __builtin_amdgcn_readfirstlane(s + s_lr[mvbits(bx + dx - B.mpx) + mvbits(by + dy - B.mpy)]);
        };
        best = cost_at(12);
        best_pos = 12;
        for (int c = 0; c < 25; c++) {
            if (c == 12) continue;
            const int cc = cost_at(c);
            if (cc < best) { best = cc; best_pos = c; }
        }
    }
    if (lane == 0) {
        out_mv[2 * boff] = bx + best_pos % 5 - 2;
        out_mv[2 * boff + 1] = by + best_pos / 5 - 2;
        out_cost[boff] = best;
        if (out_evals) out_evals[boff] = 25;
    }
}


// The +-2 quarter-pel window refinement of every 16x16 block (one wave each), from
// the integer search's keys or from given starts; with fflag, blocks flagged in it
// start from their fall-back search's key in fkeys (centre 0, radius p.range)
// instead, or (fskip) are left to k_refine_wn16_list.
template <bool T8>
__global__ __launch_bounds__(256) void k_refine_wn16(rh264_search_params p, const uint8_t *__restrict__ cur,
                                                     size_t cur_stride, const uint8_t *__restrict__ planes,
                                                     size_t planes_stride, const uint64_t *__restrict__ keys,
                                                     const int32_t *__restrict__ centres,
                                                     const int32_t *__restrict__ start_mv,
                                                     const int32_t *__restrict__ mvps, int32_t *__restrict__ int_mv,
                                                     int32_t *__restrict__ int_cost, int32_t *__restrict__ out_mv,
                                                     int32_t *__restrict__ out_cost, int32_t *__restrict__ out_evals,
                                                     const int32_t *__restrict__ fflag,
                                                     const uint64_t *__restrict__ fkeys, int fskip)
{
    __shared__ int32_t s_lr[RH264_LR_N];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr[threadIdx.x];
    __syncthreads();

    const int nbx = p.cw >> 4, nblk = nbx * (p.ch >> 4);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    unsigned wgx, wgy;
    xcd_grid2(wgx, wgy);
    const int b = (int)wgx * 4 + wave; // wave-uniform
    if (b >= nblk) return;
    const int frame = (int)wgy;
    const size_t boff = (size_t)frame * nblk + b;
    const int lx = (b % nbx) << 4, ly = (b / nbx) << 4;
    const int mpx = __builtin_amdgcn_readfirstlane(mvps ? mvps[2 * boff] : 0);
    const int mpy = __builtin_amdgcn_readfirstlane(mvps ? mvps[2 * boff + 1] : 0);
    int bx, by;
    if (keys) {
        const uint64_t key = keys[boff];
        const int NC = 2 * p.range + 1;
        const uint32_t idx = __builtin_amdgcn_readfirstlane((uint32_t)key);
        const int t = (int)(idx / (uint32_t)NC), c = (int)(idx % (uint32_t)NC);
        const int cx = __builtin_amdgcn_readfirstlane(centres ? centres[2 * boff] : 0);
        const int cy = __builtin_amdgcn_readfirstlane(centres ? centres[2 * boff + 1] : 0);
        bx = cx + 4 * (c - p.range);
        by = cy + 4 * (t - p.range);
        if (lane == 0) {
            if (int_mv) { int_mv[2 * boff] = bx; int_mv[2 * boff + 1] = by; }
            if (int_cost) int_cost[boff] = (int32_t)(key >> 32);
        }
    } else if (fflag && __builtin_amdgcn_readfirstlane(fflag[boff])) {
        if (fskip) return; // wave-uniform, no barrier follows
        // a fall-back block starts at its zero-centred full search's vector
        const uint32_t idx = __builtin_amdgcn_readfirstlane((uint32_t)fkeys[boff]);
        const int NC = 2 * p.range + 1;
        bx = 4 * ((int)(idx % (uint32_t)NC) - p.range);
        by = 4 * ((int)(idx / (uint32_t)NC) - p.range);
    } else {
        bx = __builtin_amdgcn_readfirstlane(start_mv[2 * boff]);
        by = __builtin_amdgcn_readfirstlane(start_mv[2 * boff + 1]);
    }
    wn16_block<T8>(p, s_lr, cur + (size_t)frame * cur_stride, planes + (size_t)frame * planes_stride, lx, ly, bx, by,
                   mpx, mpy, lane, boff, out_mv, out_cost, out_evals);
}

// The same refinement of the listed blocks only (frame-major global block indices, no
// rate predictor), each from its fall-back search's key in fkeys; the waves of a capped
// grid loop over the list, whose length is read here.
template <bool T8>
__global__ __launch_bounds__(256) void k_refine_wn16_list(rh264_search_params p, const uint8_t *__restrict__ cur,
                                                          size_t cur_stride, const uint8_t *__restrict__ planes,
                                                          size_t planes_stride, const int32_t *__restrict__ list,
                                                          const unsigned *__restrict__ count,
                                                          const uint64_t *__restrict__ fkeys,
                                                          int32_t *__restrict__ out_mv, int32_t *__restrict__ out_cost)
{
    __shared__ int32_t s_lr[RH264_LR_N];
    const long n = (long)*count;
    if ((long)blockIdx.x * 4 >= n) return;
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr[threadIdx.x];
    __syncthreads();
    const int nbx = p.cw >> 4, nblk = nbx * (p.ch >> 4), NC = 2 * p.range + 1;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    for (long e = (long)blockIdx.x * 4 + wave; e < n; e += (long)gridDim.x * 4) {
        const int gb = __builtin_amdgcn_readfirstlane(list[e]);
        const int frame = gb / nblk, b = gb - frame * nblk;
        const uint32_t idx = __builtin_amdgcn_readfirstlane((uint32_t)fkeys[gb]);
        const int bx = 4 * ((int)(idx % (uint32_t)NC) - p.range), by = 4 * ((int)(idx / (uint32_t)NC) - p.range);
        wn16_block<T8>(p, s_lr, cur + (size_t)frame * cur_stride, planes + (size_t)frame * planes_stride,
                       (b % nbx) << 4, (b / nbx) << 4, bx, by, 0, 0, lane, (size_t)gb, out_mv, out_cost, nullptr);
    }
}

// ---------------------------------------------------------------------------
// Reference-mode motion search (motion_search, mb16.rs:2433-3302, the default knobs
// explicit), one wave per 16x16 block.  Every candidate is read from the plane set in
// HBM (mx_operand_far: any vector, clamped at the plane's border) and priced with
// the matrix-core SATD of the refinement kernels, or with v_sad_u8 and a wave sum in
// the SAD phase of the B2 mode; the walk itself is wave-uniform scalar code.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int ms_div_euclid4(int v) { return v >= 0 ? v / 4 : -((-v + 3) / 4); }

// One block's reference-mode search (motion_search, mb16.rs:2433-3302, every knob
// explicit in rh264_diamond) by one wave.  The block (p.bw x p.bh, at (lx, ly) of the
// frame at frame_px) occupies the top-left of the wave's 16x16 transform tile; lanes
// outside it hold 0 for both source and prediction, so their 4x4 blocks transform to
// zero and add nothing to the SATD or the SAD.  Seeds (0, 0) then pr[0..npreds)
// (wave-uniform quarter-pel pairs), pr[0] pricing the rate.  Phases, in order:
//   seeds -> snap -> diamond rungs (fixed-centre argmin-of-4 passes where the
//   reference's x4 batch applies, cascade passes elsewhere) -> B2 reprice -> me_wide
//   rescue (grid, then the +-1 pel walk) -> sub-pel rings or the qpel window -> seed.
// Prices: the cost closure dist + lam_r * rate (lam_r = lambda / 4 under a rate
// table, mb16.rs:2553); the B2 phase SAD + lambda_fp * rate; the reference's
// batches (fixed-centre x4 of 16-wide blocks, the 16x16 rescue grid) price with
// lambda_fp under B2 else lambda (mb16.rs:2680-2683, 2834).  oev: bit 0 the rescue
// grid ran, bit 1 it cut the cost to <= 15/16 (mb16.rs:2878-2880).
template <bool T8>
// rescue: this search's me_wide switch (dc.rescue, or the decision's per-search
// switch under the frame's learning window); passed apart from dc so that dc stays
// the read-only kernel argument.
__device__ __forceinline__ void ms_search(const rh264_search_params &p, const rh264_diamond &dc, int rescue,
                                          const int32_t *s_lr, const int32_t *s_lrfp, const uint8_t *frame_px,
                                          const uint8_t *pl, int lx, int ly, int bw, int bh, const int32_t *pr,
                                          int npreds, int lane, int &obx, int &oby, int &obc, int &oev)
{
    MxBlock B;
    B.pw = p.cw + 2 * p.pad;
    B.ph = p.ch + 2 * p.pad;
    B.pn = (size_t)B.pw * B.ph;
    B.pl = pl;
    B.sx = B.sy = 0; // candidates are addressed by their vectors
    B.ix = lx + p.pad;
    B.iy = ly + p.pad;
    B.xoff = 0;
    B.mpx = npreds > 0 ? __builtin_amdgcn_readfirstlane(pr[0]) : 0; // the rate predictor
    B.mpy = npreds > 0 ? __builtin_amdgcn_readfirstlane(pr[1]) : 0;
    const int r = lane & 15, qd = lane >> 4;
    const bool act = r < bh && 4 * qd < bw; // the lane's 4 pixels lie inside the block
    const uint32_t src = act ? *(const uint32_t *)(frame_px + (size_t)(ly + r) * p.cw + lx + 4 * qd) : 0u;
    const MxConst K = mx_const<T8>(lane);
    const mx_f4 c2 = -mx_transform(src, K, mx_f4{0.0f, 0.0f, 0.0f, 0.0f});
    // rate of a candidate; with a table the components are tab[min(|d|, 4095)]
    auto bits = [&](int mx, int my) -> int {
        if (!dc.mv_cost) return mvbits(mx - B.mpx) + mvbits(my - B.mpy);
        // the index min(|d|, 4095) from a 64-bit difference: any predictor stays in the table
        const int64_t dx = (int64_t)mx - B.mpx, dy = (int64_t)my - B.mpy;
        const int ax = (int)min(dx < 0 ? -dx : dx, (int64_t)4095), ay = (int)min(dy < 0 ? -dy : dy, (int64_t)4095);
        return (int)dc.mv_tab[ax] + (int)dc.mv_tab[ay];
    };
    // (int64)(lambda * rate) of mb16.rs:2563; kind 0 = the cost closure's lambda
    // (lr), 1 = the B2 lambda (lr_fp), 2 = lambda itself
    auto price = [&](int kind, int rate) -> int {
        if (!dc.mv_cost) return (kind == 1 ? s_lrfp : s_lr)[rate];
        const double l = kind == 0 ? dc.lambda * 0.25 : kind == 1 ? dc.lambda_fp : dc.lambda;
        return (int)(l * (double)rate);
    };
    auto operand = [&](int mx, int my) -> uint32_t { return act ? mx_operand_far(B, lane, mx, my) : 0u; };
    auto dist = [&](int mx, int my, bool sad) -> int {
        const uint32_t o = operand(mx, my);
        if (sad) return mx_wave_sum((int)__builtin_amdgcn_sad_u8(src, o, 0u));
        return mx_satd<T8>(o, K, c2);
    };
    auto cost = [&](int mx, int my) -> int { // the cost closure (mb16.rs:2540-2565)
        return __builtin_amdgcn_readfirstlane(dist(mx, my, dc.sad_only) + price(0, bits(mx, my)));
    };
    auto cost_fp = [&](int mx, int my) -> int { // seed and diamond phases
        if (!dc.sadfp) return cost(mx, my);
        return __builtin_amdgcn_readfirstlane(dist(mx, my, true) + price(1, bits(mx, my)));
    };
    auto cost_lam = [&](int mx, int my, bool sad, int kind) -> int { // the reference's batches
        return __builtin_amdgcn_readfirstlane(dist(mx, my, sad) + price(kind, bits(mx, my)));
    };
    oev = 0;
    // seeds: (0, 0), then the predictors (strict <)
    int bx = 0, by = 0;
    int bc = cost_fp(0, 0);
    for (int i = 0; i < npreds; i++) {
        const int px = __builtin_amdgcn_readfirstlane(pr[2 * i]), py = __builtin_amdgcn_readfirstlane(pr[2 * i + 1]);
        const int pc = cost_fp(px, py);
        if (pc < bc) {
            bc = pc;
            bx = px;
            by = py;
        }
    }
    const int seedx = bx, seedy = by;
    int seed_c = bc;
    if (dc.snap && ((bx & 3) | (by & 3))) {
        bx = ms_div_euclid4(bx + 2) * 4;
        by = ms_div_euclid4(by + 2) * 4;
        bc = cost_fp(bx, by);
    }
    // diamond rungs, each walked until it stops improving
    const bool fc = dc.fc && !dc.sad_only && (bw == 16 || bw == 8) && (bh == 16 || bh == 8);
    const int fc_kind = dc.sadfp ? 1 : 2;
    for (int si = 0; si < dc.nsteps; si++) {
        const int step = dc.steps[si];
        for (;;) {
            if (fc && ((bx | by) & 3) == 0) {
                const int st = step >> 2, ix = lx + (bx >> 2), iy = ly + (by >> 2);
                if (ix - st >= 0 && iy - st >= 0 && ix + st + bw <= p.cw && iy + st + bh <= p.ch) {
                    // fixed-centre pass: the first minimum of the 4-point ring
                    int bi = -1, bcc = bc;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int x = bx + (i == 0 ? step : i == 1 ? -step : 0);
                        const int y = by + (i == 2 ? step : i == 3 ? -step : 0);
                        const int cc = bw == 16 ? cost_lam(x, y, dc.sadfp, fc_kind) : cost_fp(x, y);
                        if (cc < bcc) {
                            bcc = cc;
                            bi = i;
                        }
                    }
                    if (bi < 0) break;
                    bc = bcc;
                    bx += bi == 0 ? step : bi == 1 ? -step : 0;
                    by += bi == 2 ? step : bi == 3 ? -step : 0;
                    continue;
                }
            }
            bool improved = false;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int x = bx + (i == 0 ? step : i == 1 ? -step : 0), y = by + (i == 2 ? step : i == 3 ? -step : 0);
                const int cc = cost_fp(x, y);
                if (cc < bc) {
                    bc = cc;
                    bx = x;
                    by = y;
                    improved = true;
                }
            }
            if (!improved) break;
        }
    }
    if (dc.sadfp) {
        bc = cost(bx, by);
        seed_c = cost(seedx, seedy);
    }
    // me_wide rescue
    bool gate = rescue && !dc.sad_only;
    if (gate && !dc.rescue_fast) { // the flat test: the block's variance below rescue_var
        const int v0 = act ? (int)(src & 255) : 0, v1 = act ? (int)((src >> 8) & 255) : 0;
        const int v2 = act ? (int)((src >> 16) & 255) : 0, v3 = act ? (int)(src >> 24) : 0;
        const uint64_t sm = (uint64_t)mx_wave_sum(v0 + v1 + v2 + v3);
        const uint64_t sq = (uint64_t)mx_wave_sum(v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3);
        const uint64_t n = (uint64_t)(bw * bh);
        gate = (sq - sm * sm / n) / n < (uint64_t)dc.rescue_var;
    }
    if (gate) {
        const int dist0 = bc - price(2, bits(bx, by));
        if (dist0 / (bw * bh) > dc.rescue_thresh) {
            const int pre_c = bc;
            const int gx = ms_div_euclid4(bx + 2) * 4, gy = ms_div_euclid4(by + 2) * 4;
            const int rr = dc.rescue_range, icx = lx + (gx >> 2), icy = ly + (gy >> 2);
            const bool batched = bw == 16 && bh == 16 && icx >= rr && icx + rr + 16 <= p.cw && icy >= rr &&
                                 icy + rr + 16 <= p.ch;
            int gbx = bx, gby = by;
            for (int dy = -rr; dy <= rr; dy += 2)
                for (int dx = -rr; dx <= rr; dx += 2) {
                    const int x = gx + 4 * dx, y = gy + 4 * dy;
                    const int cc = batched ? cost_lam(x, y, false, 2) : cost(x, y);
                    if (cc < bc) {
                        bc = cc;
                        gbx = x;
                        gby = y;
                    }
                }
            bx = gbx;
            by = gby;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int x = bx + 4 * dx, y = by + 4 * dy;
                    const int cc = cost(x, y);
                    if (cc < bc) {
                        bc = cc;
                        bx = x;
                        by = y;
                    }
                }
            oev = 1 | ((int64_t)bc * 16 <= (int64_t)pre_c * 15 ? 2 : 0);
        }
    }
    // the sub-pel stage
    const int pat = p.subpel_pat;
    if (pat == 4) {
        const int sx = bx, sy = by;
        for (int dy = -2; dy <= 2; dy++)
            for (int dx = -2; dx <= 2; dx++) {
                if (dx == 0 && dy == 0) continue;
                const int cc = cost(sx + dx, sy + dy);
                if (cc < bc) {
                    bc = cc;
                    bx = sx + dx;
                    by = sy + dy;
                }
            }
    } else if (pat >= 0) {
        const int nring = (pat & 1) ? 4 : 8;
        for (int step = 2; step >= 1; step--) {
            for (;;) {
                bool improved = false;
                for (int i = 0; i < nring; i++) {
                    // ring8 order (s,0),(-s,0),(0,s),(0,-s),(s,s),(-s,-s),(s,-s),(-s,s)
                    const int rx = (i < 2 || i >= 4) ? ((i & 1) ? -step : step) : 0;
                    const int ry = i < 2 ? 0 : i < 4 ? ((i & 1) ? -step : step) : ((i == 4 || i == 7) ? step : -step);
                    const int cc = cost(bx + rx, by + ry);
                    if (cc < bc) {
                        bc = cc;
                        bx += rx;
                        by += ry;
                        improved = true;
                    }
                }
                if (!improved || (pat & 2)) break;
            }
        }
    }
    if (dc.snap && seed_c < bc) {
        bx = seedx;
        by = seedy;
        bc = seed_c;
    }
    obx = bx;
    oby = by;
    obc = bc;
}

template <bool T8>
__global__ __launch_bounds__(256) void k_motion_search(rh264_search_params p, rh264_diamond dc,
                                                       const uint8_t *__restrict__ cur, size_t cur_stride,
                                                       const uint8_t *__restrict__ planes, size_t planes_stride,
                                                       const int32_t *__restrict__ preds, int npreds,
                                                       int32_t *__restrict__ out_mv, int32_t *__restrict__ out_cost,
                                                       int32_t *__restrict__ out_ev)
{
    __shared__ int32_t s_lr[RH264_LR_N], s_lrfp[RH264_LR_N];
    if (threadIdx.x < RH264_LR_N) {
        s_lr[threadIdx.x] = p.lr[threadIdx.x];
        s_lrfp[threadIdx.x] = p.lr_fp[threadIdx.x];
    }
    __syncthreads();
    const int nbx = p.cw / p.bw, nblk = nbx * (p.ch / p.bh);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    unsigned wgx, wgy;
    xcd_grid2(wgx, wgy);
    const int b = (int)wgx * 4 + wave; // wave-uniform
    if (b >= nblk) return;
    const int frame = (int)wgy;
    const size_t boff = (size_t)frame * nblk + b;
    int bx, by, bc, ev;
    ms_search<T8>(p, dc, dc.rescue, s_lr, s_lrfp, cur + (size_t)frame * cur_stride, planes + (size_t)frame * planes_stride,
                  (b % nbx) * p.bw, (b / nbx) * p.bh, p.bw, p.bh, preds + boff * (size_t)(npreds > 0 ? npreds : 1) * 2,
                  npreds, lane, bx, by, bc, ev);
    if (lane == 0) {
        out_mv[2 * boff] = bx;
        out_mv[2 * boff + 1] = by;
        out_cost[boff] = bc;
        if (out_ev) out_ev[boff] = ev;
    }
}

// median3 (inter.rs:9-11)
__device__ __forceinline__ int ms_median3(int a, int b, int c)
{
    return a + b + c - min(min(a, b), c) - max(max(a, b), c);
}

// predict_mv (inter.rs:36-73) against reference cur_ref for macroblock (x, y), from
// the decided vectors and references of its neighbours A (left), B (above) and C
// (above-right, else D = above-left; mv_neighbors, mb16.rs:1813-1836).  Every
// decided macroblock is an inter block, so an available neighbour is usable.
__device__ __forceinline__ void ms_predict_mv(const int32_t *mv, const int32_t *ref, int nbx, int x, int y,
                                              int cur_ref, int &px, int &py)
{
    const bool av_a = x > 0, av_b = y > 0;
    const bool c_is_c = y > 0 && x + 1 < nbx;
    const bool av_c = c_is_c || (x > 0 && y > 0);
    const int ia = y * nbx + x - 1, ib = (y - 1) * nbx + x, ic = c_is_c ? (y - 1) * nbx + x + 1 : (y - 1) * nbx + x - 1;
    int ax = 0, ay = 0, ra = -1, bx = 0, by = 0, rb = -1, cx = 0, cy = 0, rc = -1;
    if (av_a) { ax = mv[2 * ia]; ay = mv[2 * ia + 1]; ra = ref[ia]; }
    if (av_b) { bx = mv[2 * ib]; by = mv[2 * ib + 1]; rb = ref[ib]; }
    if (av_c) { cx = mv[2 * ic]; cy = mv[2 * ic + 1]; rc = ref[ic]; }
    if (!av_b && !av_c && av_a) { // B and C take A
        bx = cx = ax;
        by = cy = ay;
        rb = rc = ra;
    }
    const int matches = (ra == cur_ref) + (rb == cur_ref) + (rc == cur_ref);
    if (matches == 1) {
        px = ra == cur_ref ? ax : rb == cur_ref ? bx : cx;
        py = ra == cur_ref ? ay : rb == cur_ref ? by : cy;
    } else {
        px = ms_median3(ax, bx, cx);
        py = ms_median3(ay, by, cy);
    }
}

// The encoder's 16x16 decision (best_part, mb16.rs:4921-4959, no extra seeds) as a
// wavefront: launch wf_t covers the macroblocks (x, y) with x + 2y = wf_t of every
// frame, whose left, above and above-right (or above-left) neighbours were decided
// by earlier launches.  RW waves share a macroblock, wave r searching reference r
// (plane set frame + nrefs - 1 - r) from its single seed predict_mv(.., r); the
// cheapest cost + refc[r] wins, the first in reference order on ties.
struct MsRefCost {
    int32_t c[4];
};
template <bool T8, int RW>
__global__ __launch_bounds__(256) void k_ms_wavefront(rh264_search_params p, rh264_diamond dc,
                                                      const uint8_t *__restrict__ cur, size_t cur_stride,
                                                      const uint8_t *__restrict__ planes, size_t planes_stride,
                                                      int nrefs, MsRefCost refc, int wf_t, int32_t *__restrict__ out_ref,
                                                      int32_t *__restrict__ out_mv, int32_t *__restrict__ out_cost)
{
    __shared__ int32_t s_lr[RH264_LR_N], s_lrfp[RH264_LR_N];
    __shared__ int32_t s_res[4][3];
    if (threadIdx.x < RH264_LR_N) {
        s_lr[threadIdx.x] = p.lr[threadIdx.x];
        s_lrfp[threadIdx.x] = p.lr_fp[threadIdx.x];
    }
    __syncthreads();
    const int nbx = p.cw >> 4, nby = p.ch >> 4, nblk = nbx * nby;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // the anti-diagonal's macroblocks: y from y_lo while x = wf_t - 2y stays >= 0
    const int y_lo = max(0, (wf_t - (nbx - 1) + 1) / 2), y_hi = min(nby - 1, wf_t / 2);
    const int n_t = y_hi - y_lo + 1;
    const int j = (int)blockIdx.x * (4 / RW) + wave / RW, rf = wave % RW; // wave-uniform
    const int frame = (int)blockIdx.y;
    const bool mb_ok = j < n_t;
    const int y = y_lo + j, x = wf_t - 2 * y;
    const size_t boff = (size_t)frame * nblk + (size_t)y * nbx + x;
    if (mb_ok && rf < nrefs) {
        int32_t pr[2];
        int px, py;
        ms_predict_mv(out_mv + (size_t)frame * nblk * 2, out_ref + (size_t)frame * nblk, nbx, x, y, rf, px, py);
        pr[0] = __builtin_amdgcn_readfirstlane(px);
        pr[1] = __builtin_amdgcn_readfirstlane(py);
        int bx, by, bc, ev;
        ms_search<T8>(p, dc, dc.rescue, s_lr, s_lrfp, cur + (size_t)frame * cur_stride,
                      planes + (size_t)(frame + nrefs - 1 - rf) * planes_stride, x << 4, y << 4, 16, 16, pr, 1, lane,
                      bx, by, bc, ev);
        bc += refc.c[rf];
        if constexpr (RW == 1) {
            if (lane == 0) {
                out_ref[boff] = 0;
                out_mv[2 * boff] = bx;
                out_mv[2 * boff + 1] = by;
                out_cost[boff] = bc;
            }
        } else if (lane == 0) {
            s_res[wave][0] = bx;
            s_res[wave][1] = by;
            s_res[wave][2] = bc;
        }
    }
    if constexpr (RW > 1) {
        __syncthreads();
        if (mb_ok && rf == 0 && lane == 0) {
            int br = 0;
            for (int q = 1; q < nrefs; q++)
                if (s_res[wave + q][2] < s_res[wave + br][2]) br = q;
            out_ref[boff] = br;
            out_mv[2 * boff] = s_res[wave + br][0];
            out_mv[2 * boff + 1] = s_res[wave + br][1];
            out_cost[boff] = s_res[wave + br][2];
        }
    }
}

// The same decision as a fixed point.  S = the (ref, mv) of every macroblock.  A
// pass re-searches exactly the macroblocks whose predictors, computed from S,
// differ from the predictors their current result was searched with.  The
// encoder-order decision is the only S with no such macroblock (by induction over
// the anti-diagonals), and after pass k every macroblock on the first k
// anti-diagonals holds its final result, so at most (number of anti-diagonals + 1)
// passes reach it; vectors that settle early end it much sooner.
//
// Check: one thread per macroblock of every frame, predictors from S for each
// reference; a changed macroblock stores its new predictors in used[i][r] and
// joins the work list (one atomic per wave).
// With the rescue (usedf non-null) a search's switch joins its inputs: on unless the
// frame's learning window closed before it (fstate[f] = {order index of the search
// that filled the window, or INT32_MAX; closed}); usedf[i][r] keeps the switch used.
__global__ __launch_bounds__(256) void k_ms_check(int nbx, int nblk, int ntot, int nrefs,
                                                  const int32_t *__restrict__ ref, const int32_t *__restrict__ mv,
                                                  int32_t *__restrict__ used, int32_t *__restrict__ list,
                                                  uint32_t *__restrict__ count, int32_t *__restrict__ usedf,
                                                  const int32_t *__restrict__ fstate)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    bool dirty = false;
    if (i < ntot) {
        const int frame = i / nblk, b = i - frame * nblk;
        for (int r = 0; r < nrefs; r++) {
            int px, py;
            ms_predict_mv(mv + (size_t)frame * nblk * 2, ref + (size_t)frame * nblk, nbx, b % nbx, b / nbx, r, px, py);
            int32_t *u = used + (size_t)i * 8 + 2 * r;
            if (u[0] != px || u[1] != py) {
                u[0] = px;
                u[1] = py;
                dirty = true;
            }
            if (usedf) {
                const int on = !(fstate[2 * frame + 1] && b * nrefs + r > fstate[2 * frame]);
                if (usedf[(size_t)i * 4 + r] != on) {
                    usedf[(size_t)i * 4 + r] = on;
                    dirty = true;
                }
            }
        }
    }
    const uint64_t m = __ballot(dirty);
    if (m == 0) return;
    const int lane = threadIdx.x & 63, leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (dirty) list[base + below] = i;
}

// Search of the listed macroblocks with their stored predictors: RW waves per
// macroblock, wave r searching reference r, the workgroups striding over the list
// (count entries, written by the check just before).  Block 0 clears the other
// counter for the next check.
template <bool T8, int RW>
__global__ __launch_bounds__(256) void k_ms_listed(rh264_search_params p, rh264_diamond dc,
                                                   const uint8_t *__restrict__ cur, size_t cur_stride,
                                                   const uint8_t *__restrict__ planes, size_t planes_stride,
                                                   int nrefs, MsRefCost refc, const int32_t *__restrict__ used,
                                                   const int32_t *__restrict__ list, const uint32_t *__restrict__ count,
                                                   uint32_t *__restrict__ next_count, int32_t *__restrict__ out_ref,
                                                   int32_t *__restrict__ out_mv, int32_t *__restrict__ out_cost,
                                                   const int32_t *__restrict__ usedf, int32_t *__restrict__ events)
{
    __shared__ int32_t s_lr[RH264_LR_N], s_lrfp[RH264_LR_N];
    __shared__ int32_t s_res[4][3];
    if (threadIdx.x < RH264_LR_N) {
        s_lr[threadIdx.x] = p.lr[threadIdx.x];
        s_lrfp[threadIdx.x] = p.lr_fp[threadIdx.x];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
    __syncthreads();
    const int nbx = p.cw >> 4, nblk = nbx * (p.ch >> 4);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int per_wg = 4 / RW, slot = wave / RW, rf = wave % RW;
    const int n = (int)__builtin_amdgcn_readfirstlane(*count);
    for (int e0 = (int)blockIdx.x * per_wg; e0 < n; e0 += (int)gridDim.x * per_wg) { // workgroup-uniform
        const int e = e0 + slot;
        const bool ok = e < n;
        const int i = ok ? __builtin_amdgcn_readfirstlane(list[e]) : 0;
        const int frame = i / nblk, b = i - frame * nblk;
        if (ok && rf < nrefs) {
            int bx, by, bc, ev;
            const int resc = usedf ? __builtin_amdgcn_readfirstlane(usedf[(size_t)i * 4 + rf]) : 0;
            ms_search<T8>(p, dc, resc, s_lr, s_lrfp, cur + (size_t)frame * cur_stride,
                          planes + (size_t)(frame + nrefs - 1 - rf) * planes_stride, (b % nbx) << 4, (b / nbx) << 4,
                          16, 16, used + (size_t)i * 8 + 2 * rf, 1, lane, bx, by, bc, ev);
            if (events && lane == 0) events[(size_t)i * 4 + rf] = ev;
            bc += refc.c[rf];
            if constexpr (RW == 1) {
                if (lane == 0) {
                    out_ref[i] = 0;
                    out_mv[2 * i] = bx;
                    out_mv[2 * i + 1] = by;
                    out_cost[i] = bc;
                }
            } else if (lane == 0) {
                s_res[wave][0] = bx;
                s_res[wave][1] = by;
                s_res[wave][2] = bc;
            }
        }
        if constexpr (RW > 1) {
            __syncthreads();
            if (ok && rf == 0 && lane == 0) {
                int br = 0;
                for (int q = 1; q < nrefs; q++)
                    if (s_res[wave + q][2] < s_res[wave + br][2]) br = q;
                out_ref[i] = br;
                out_mv[2 * i] = s_res[wave + br][0];
                out_mv[2 * i + 1] = s_res[wave + br][1];
                out_cost[i] = s_res[wave + br][2];
            }
            __syncthreads();
        }
    }
}

// The frame's rescue learning window from the searches' events (mb16.rs:2868-2886),
// one wave per frame over the searches in order (macroblock, then reference): the
// search whose grid ran for the learn-th time fills the window, which closes when
// fewer than payoff % of those learn searches cut their cost to <= 15/16.  fstate[f]
// = {order index of that search, or INT32_MAX when the window does not fill; closed}.
__global__ void k_ms_latch_init(int nframes, int32_t *__restrict__ fstate)
{
    const int f = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (f < nframes) {
        fstate[2 * f] = INT32_MAX;
        fstate[2 * f + 1] = 0;
    }
}

// slots: searches per macroblock (their order inside it), stride: event words per
// macroblock.
__global__ __launch_bounds__(64) void k_ms_latch(int nblk, int slots, int stride, int learn, int payoff,
                                                 const int32_t *__restrict__ events, int32_t *__restrict__ fstate)
{
    const int f = blockIdx.x, lane = threadIdx.x;
    const int total = nblk * slots;
    int n = 0, big = 0, pstar = INT32_MAX, closed = 0;
    if (learn > 0)
        for (int base = 0; base < total; base += 64) { // wave-uniform
            const int idx = base + lane;
            const int e = idx < total ? events[((size_t)f * nblk + idx / slots) * stride + idx % slots] : 0;
            const bool ran = e & 1, paid = ran && (e & 2);
            const uint64_t mr = __ballot(ran), mp = __ballot(paid);
            const int cnt = __popcll(mr);
            if (n + cnt >= learn) {
                // the (learn - n)-th search of this chunk whose grid ran
                const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mr >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)mr, 0u));
                const uint64_t hit = __ballot(ran && below == learn - n - 1);
                const int p = __ffsll((long long)hit) - 1;
                const uint64_t upto = p == 63 ? ~0ull : ((1ull << (p + 1)) - 1);
                big += __popcll(mp & upto);
                pstar = base + p;
                closed = big * 100 < learn * payoff;
                break;
            }
            n += cnt;
            big += __popcll(mp);
        }
    if (lane == 0) {
        fstate[2 * f] = pstar;
        fstate[2 * f + 1] = closed;
    }
}

// ---- the P macroblock decision with partitions (mb16.rs:8849-9010, SATD currency) --
// rec per macroblock: {mode, cost, (ref, mvx, mvy) x quadrants TL, TR, BL, BR}.
constexpr int PD_REC = 14;

// predict_mv (inter.rs:36-73) against cur_ref from the neighbours read on the 4x4
// motion grid at the macroblock's top-left block (mv_neighbors_block, mb16.rs:
// 1871-1888): A = the left macroblock's top-right quadrant, B and C = the bottom-left
// quadrant of the above and above-right ones, D = the above-left bottom-right.
__device__ __forceinline__ void pd_predict_mv(const int32_t *rec, int nbx, int x, int y, int cur_ref, int &px,
                                              int &py)
{
    const bool av_a = x > 0, av_b = y > 0, c_is_c = y > 0 && x + 1 < nbx;
    const bool av_c = c_is_c || (x > 0 && y > 0);
    const int32_t *A = rec + (size_t)PD_REC * (y * nbx + x - 1) + 2 + 3 * 1;
    const int32_t *B = rec + (size_t)PD_REC * ((y - 1) * nbx + x) + 2 + 3 * 2;
    const int32_t *Cq = c_is_c ? rec + (size_t)PD_REC * ((y - 1) * nbx + x + 1) + 2 + 3 * 2
                               : rec + (size_t)PD_REC * ((y - 1) * nbx + x - 1) + 2 + 3 * 3;
    int ax = 0, ay = 0, ra = -1, bx = 0, by = 0, rb = -1, cx = 0, cy = 0, rc = -1;
    if (av_a) { ra = A[0]; ax = A[1]; ay = A[2]; }
    if (av_b) { rb = B[0]; bx = B[1]; by = B[2]; }
    if (av_c) { rc = Cq[0]; cx = Cq[1]; cy = Cq[2]; }
    if (!av_b && !av_c && av_a) {
        bx = cx = ax;
        by = cy = ay;
        rb = rc = ra;
    }
    const int matches = (ra == cur_ref) + (rb == cur_ref) + (rc == cur_ref);
    if (matches == 1) {
        px = ra == cur_ref ? ax : rb == cur_ref ? bx : cx;
        py = ra == cur_ref ? ay : rb == cur_ref ? by : cy;
    } else {
        px = ms_median3(ax, bx, cx);
        py = ms_median3(ay, by, cy);
    }
}

// Fixed-point check, as k_ms_check, with the predictors from the quadrant grid.
// With the rescue (usedf non-null) every search slot of the macroblock (PD_SLOTS words
// each: the 16x16 references, then each partition's references) keeps its switch, as in
// k_ms_check; slots = the slots in use.
constexpr int PD_SLOTS = 36; // (1 + 8 partitions) x 4 references
__global__ __launch_bounds__(256) void k_pd_check(int nbx, int nblk, int ntot, int nrefs, const int32_t *__restrict__ rec,
                                                  int32_t *__restrict__ used, int32_t *__restrict__ list,
                                                  uint32_t *__restrict__ count, int slots, int32_t *__restrict__ usedf,
                                                  const int32_t *__restrict__ fstate)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    bool dirty = false;
    if (i < ntot) {
        const int frame = i / nblk, b = i - frame * nblk;
        for (int r = 0; r < nrefs; r++) {
            int px, py;
            pd_predict_mv(rec + (size_t)frame * nblk * PD_REC, nbx, b % nbx, b / nbx, r, px, py);
            int32_t *u = used + (size_t)i * 8 + 2 * r;
            if (u[0] != px || u[1] != py) {
                u[0] = px;
                u[1] = py;
                dirty = true;
            }
        }
        if (usedf)
            for (int sl = 0; sl < slots; sl++) {
                const int on = !(fstate[2 * frame + 1] && b * slots + sl > fstate[2 * frame]);
                if (usedf[(size_t)i * PD_SLOTS + sl] != on) {
                    usedf[(size_t)i * PD_SLOTS + sl] = on;
                    dirty = true;
                }
            }
    }
    const uint64_t m = __ballot(dirty);
    if (m == 0) return;
    const int lane = threadIdx.x & 63, leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (dirty) list[base + below] = i;
}

// The listed macroblocks' decisions, one workgroup (4 waves) per macroblock: the 16x16
// best_part (wave r searches reference r), then, when its cost exceeds the gate, the
// 16x8, 8x16 and (p8x8) 8x8 best_parts with the 16x16 vector as the extra seed, their
// (partition, reference) searches spread over the 4 waves; the first strict minimum
// of c16, ct + cb, cl + cr, lme4 + sum c8 is written to rec.
struct PdArgs {
    int32_t nrefs, gate, p8x8, lme4;
    int32_t refc[4];
};
// 6 waves per SIMD: 80 VGPRs without spills (the compiler's own 84 allow 5)
template <bool T8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
void k_pd_listed(rh264_search_params p, rh264_diamond dc, const uint8_t *__restrict__ cur, size_t cur_stride,
                 const uint8_t *__restrict__ planes, size_t planes_stride, PdArgs a, const int32_t *__restrict__ used,
                 const int32_t *__restrict__ list, const uint32_t *__restrict__ count,
                 uint32_t *__restrict__ next_count, int32_t *__restrict__ rec, const int32_t *__restrict__ usedf,
                 int32_t *__restrict__ events)
{
    __shared__ int32_t s_lr[RH264_LR_N], s_lrfp[RH264_LR_N];
    __shared__ int32_t s16[4][3], sp[8][4][3];
    if (threadIdx.x < RH264_LR_N) {
        s_lr[threadIdx.x] = p.lr[threadIdx.x];
        s_lrfp[threadIdx.x] = p.lr_fp[threadIdx.x];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0;
    __syncthreads();
    const int nbx = p.cw >> 4, nblk = nbx * (p.ch >> 4);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nrefs = a.nrefs;
    const int n = (int)__builtin_amdgcn_readfirstlane(*count);
    // 16x8 top, bottom; 8x16 left, right; 8x8 TL, TR, BL, BR: (x, y, w, h)
    constexpr int geo[8][4] = {{0, 0, 16, 8}, {0, 8, 16, 8}, {0, 0, 8, 16}, {8, 0, 8, 16},
                               {0, 0, 8, 8},  {8, 0, 8, 8},  {0, 8, 8, 8},  {8, 8, 8, 8}};
    for (int e = (int)blockIdx.x; e < n; e += (int)gridDim.x) { // workgroup-uniform
        const int i = __builtin_amdgcn_readfirstlane(list[e]);
        const int frame = i / nblk, b = i - frame * nblk;
        const int lx = (b % nbx) << 4, ly = (b / nbx) << 4;
        const uint8_t *fpx = cur + (size_t)frame * cur_stride;
        const int32_t *u = used + (size_t)i * 8;
        if (wave < nrefs) {
            int bx, by, bc, ev;
            const int resc = usedf ? __builtin_amdgcn_readfirstlane(usedf[(size_t)i * PD_SLOTS + wave]) : 0;
            ms_search<T8>(p, dc, resc, s_lr, s_lrfp, fpx, planes + (size_t)(frame + nrefs - 1 - wave) * planes_stride, lx, ly,
                          16, 16, u + 2 * wave, 1, lane, bx, by, bc, ev);
            if (events && lane == 0) events[(size_t)i * PD_SLOTS + wave] = ev;
            if (lane == 0) {
                s16[wave][0] = bx;
                s16[wave][1] = by;
                s16[wave][2] = bc + a.refc[wave];
            }
        }
        __syncthreads();
        int r16 = 0;
        for (int q = 1; q < nrefs; q++)
            if (s16[q][2] < s16[r16][2]) r16 = q;
        const int mvx16 = s16[r16][0], mvy16 = s16[r16][1], c16 = s16[r16][2];
        const bool split = c16 > a.gate;
        if (split) {
            const int nparts = a.p8x8 ? 8 : 4;
            for (int j = wave; j < nparts * nrefs; j += 4) { // wave-uniform
                const int part = j / nrefs, rf = j - part * nrefs;
                // seeds: the predictor against rf (pricing the rate), then the 16x16 vector
                const int32_t prl[4] = {u[2 * rf], u[2 * rf + 1], mvx16, mvy16};
                int bx, by, bc, ev;
                const int resc = usedf ? __builtin_amdgcn_readfirstlane(usedf[(size_t)i * PD_SLOTS + nrefs + j]) : 0;
                ms_search<T8>(p, dc, resc, s_lr, s_lrfp, fpx, planes + (size_t)(frame + nrefs - 1 - rf) * planes_stride,
                              lx + geo[part][0], ly + geo[part][1], geo[part][2], geo[part][3], prl, 2, lane, bx, by,
                              bc, ev);
                if (events && lane == 0) events[(size_t)i * PD_SLOTS + nrefs + j] = ev;
                if (lane == 0) {
                    sp[part][rf][0] = bx;
                    sp[part][rf][1] = by;
                    sp[part][rf][2] = bc + a.refc[rf];
                }
            }
        } else if (events && (int)threadIdx.x < (a.p8x8 ? 8 : 4) * nrefs) {
            events[(size_t)i * PD_SLOTS + nrefs + threadIdx.x] = 0; // no partition search ran this pass
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int mode = 0, best = c16;
            int qr[4] = {r16, r16, r16, r16}, qx[4] = {mvx16, mvx16, mvx16, mvx16}, qy[4] = {mvy16, mvy16, mvy16, mvy16};
            if (split) {
                int pr_[8];
                int pc[8];
                const int nparts = a.p8x8 ? 8 : 4;
                for (int k = 0; k < nparts; k++) {
                    int br = 0;
                    for (int q = 1; q < nrefs; q++)
                        if (sp[k][q][2] < sp[k][br][2]) br = q;
                    pr_[k] = br;
                    pc[k] = sp[k][br][2];
                }
                auto take = [&](const int (&map)[4]) {
                    for (int q = 0; q < 4; q++) {
                        qr[q] = pr_[map[q]];
                        qx[q] = sp[map[q]][pr_[map[q]]][0];
                        qy[q] = sp[map[q]][pr_[map[q]]][1];
                    }
                };
                if (pc[0] + pc[1] < best) {
                    best = pc[0] + pc[1];
                    mode = 1;
                    const int m1[4] = {0, 0, 1, 1};
                    take(m1);
                }
                if (pc[2] + pc[3] < best) {
                    best = pc[2] + pc[3];
                    mode = 2;
                    const int m2[4] = {2, 3, 2, 3};
                    take(m2);
                }
                if (a.p8x8) {
                    const int c8 = a.lme4 + pc[4] + pc[5] + pc[6] + pc[7];
                    if (c8 < best) {
                        best = c8;
                        mode = 3;
                        const int m3[4] = {4, 5, 6, 7};
                        take(m3);
                    }
                }
            }
            int32_t *o = rec + (size_t)PD_REC * i;
            o[0] = mode;
            o[1] = best;
            for (int q = 0; q < 4; q++) {
                o[2 + 3 * q] = qr[q];
                o[3 + 3 * q] = qx[q];
                o[4 + 3 * q] = qy[q];
            }
        }
        __syncthreads();
    }
}

// The luma predictions of decided P frames, one thread per 4 pixels of a row:
// the motion-compensated 16x16 prediction of each macroblock at its decided (ref,
// mv), and its P_Skip prediction from reference 0 at skip_mv (mb16.rs:1841-1853):
// (0, 0) when A or B is unavailable or is a zero vector of reference 0, else
// predict_mv(A, B, C, 0).  Samples are read from the plane sets with the operand
// table and coordinates clamped into the padded plane, which equals mc_luma
// (inter.rs:1193-1262) for any vector (mc_luma_cached, mb16.rs:4514-4541).
__device__ __forceinline__ uint32_t dp_pred4(const uint8_t *pl, int pw, int ph, int x, int y, int mvx, int mvy)
{
    const size_t np = (size_t)pw * ph;
    const uint32_t q = qtab((mvx & 3) * 4 + (mvy & 3));
    const uint8_t *pa = pl + np * (q & 15), *pb = pl + np * ((q >> 12) & 15);
    const int px = x + (mvx >> 2), py = y + (mvy >> 2);
    const uint8_t *ra = pa + (size_t)clampi(py + ((q >> 8) & 15), 0, ph - 1) * pw;
    const uint8_t *rb = pb + (size_t)clampi(py + ((q >> 20) & 15), 0, ph - 1) * pw;
    uint32_t o = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int v = ra[clampi(px + k + ((q >> 4) & 15), 0, pw - 1)];
        if ((mvx | mvy) & 1) v = (v + rb[clampi(px + k + ((q >> 16) & 15), 0, pw - 1)] + 1) >> 1;
        o |= (uint32_t)v << (8 * k);
    }
    return o;
}

__global__ __launch_bounds__(256) void k_decision_predict(int cw, int ch, int pad, int nrefs,
                                                          const uint8_t *__restrict__ planes, size_t planes_stride,
                                                          const int32_t *__restrict__ ref,
                                                          const int32_t *__restrict__ mv,
                                                          int32_t *__restrict__ skip_mv, uint8_t *__restrict__ pred,
                                                          uint8_t *__restrict__ skip_pred)
{
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x); // 4-pixel group of the frame
    const int frame = (int)blockIdx.y;
    if (g >= cw * ch / 4) return;
    const int y = (g * 4) / cw, x = g * 4 - y * cw;
    const int nbx = cw >> 4, nblk = nbx * (ch >> 4), mx = x >> 4, my = y >> 4, b = my * nbx + mx;
    const int32_t *fmv = mv + (size_t)frame * nblk * 2, *fref = ref + (size_t)frame * nblk;
    // P_Skip vector from the decided neighbours
    int sx = 0, sy = 0;
    const bool av_a = mx > 0, av_b = my > 0;
    if (av_a && av_b) {
        const int ia = b - 1, ib = b - nbx;
        const bool zero_a = fref[ia] == 0 && fmv[2 * ia] == 0 && fmv[2 * ia + 1] == 0;
        const bool zero_b = fref[ib] == 0 && fmv[2 * ib] == 0 && fmv[2 * ib + 1] == 0;
        if (!zero_a && !zero_b) ms_predict_mv(fmv, fref, nbx, mx, my, 0, sx, sy);
    }
    const int pw = cw + 2 * pad, ph = ch + 2 * pad;
    const size_t boff = (size_t)frame * nblk + b;
    if ((x & 15) == 0 && (y & 15) == 0) {
        skip_mv[2 * boff] = sx;
        skip_mv[2 * boff + 1] = sy;
    }
    const int r = clampi(fref[b], 0, nrefs - 1);
    const uint8_t *pl_r = planes + (size_t)(frame + nrefs - 1 - r) * planes_stride;
    const uint8_t *pl_0 = planes + (size_t)(frame + nrefs - 1) * planes_stride;
    const size_t o = (size_t)frame * cw * ch + (size_t)y * cw + x;
    *(uint32_t *)(pred + o) = dp_pred4(pl_r, pw, ph, x + pad, y + pad, fmv[2 * b], fmv[2 * b + 1]);
    *(uint32_t *)(skip_pred + o) = dp_pred4(pl_0, pw, ph, x + pad, y + pad, sx, sy);
}

// ---------------------------------------------------------------------------
// Rate-control frame complexity (lookahead::complexity, lookahead.rs:38-101), one
// wave per macroblock, with the refinement's matrix-core 4x4 Hadamards:
//  * IDR: the AC part of the sixteen transforms of the pixels (intra_activity,
//    55-72); the DC coefficient of each 4x4 is element 0 of lanes l % 4 == 0;
//  * P: the minimum over the nine full-pel candidates (0,0), (+-1,0), (0,+-1),
//    (+-2,0), (0,+-2) of the residual's sum |H d| (inter_activity, 76-101), the
//    prediction read from the reference frame with coordinates clamped into it.
// Each wave scores RC_MB_PER_WAVE macroblocks (strided over the frame), the four
// waves of a workgroup add their sums in LDS, and one atomic add per workgroup
// accumulates the frame's 64-bit total (integer, so the order of the adds does not
// matter); k_rc_finalize writes max(total, 1).
// ---------------------------------------------------------------------------

constexpr int RC_MB_PER_WAVE = 16;

__device__ __forceinline__ int rc_mb_score(const uint8_t *__restrict__ cf, const uint8_t *__restrict__ rf, int cw,
                                           int ch, int x0, int y0, int lane, const MxConst &K)
{
    const int r = lane & 15, qd = lane >> 4;
    const uint32_t src = *(const uint32_t *)(cf + (size_t)(y0 + r) * cw + x0 + 4 * qd);
    int score;
    if (!rf) {
        const mx_f4 y = mx_transform(src, K, mx_f4{0.0f, 0.0f, 0.0f, 0.0f});
        const float s = ((lane & 3) ? fabsf(y[0]) : 0.0f) + fabsf(y[1]) + (fabsf(y[2]) + fabsf(y[3]));
        score = mx_wave_sum((int)s);
    } else {
        const mx_f4 c2 = -mx_transform(src, K, mx_f4{0.0f, 0.0f, 0.0f, 0.0f});
        // v[0..4]: dy = 0 with dx = 0, +1, -1, +2, -2; v[5..8]: dx = 0 with dy = +1, -1, +2, -2
        uint32_t v[9];
        const int col = x0 + 4 * qd;
        if (x0 >= 4 && y0 >= 2 && x0 + 20 <= cw && y0 + 18 <= ch) {
            // aligned dwords: every byte read lies inside the frame
            const uint32_t *row = (const uint32_t *)(rf + (size_t)(y0 + r) * cw + col);
            const uint32_t d0 = row[-1], d1 = row[0], d2 = row[1];
            v[0] = d1;
            v[1] = __builtin_amdgcn_alignbyte(d2, d1, 1u);
            v[2] = __builtin_amdgcn_alignbyte(d1, d0, 3u);
            v[3] = __builtin_amdgcn_alignbyte(d2, d1, 2u);
            v[4] = __builtin_amdgcn_alignbyte(d1, d0, 2u);
            v[5] = *(const uint32_t *)(rf + (size_t)(y0 + r + 1) * cw + col);
            v[6] = *(const uint32_t *)(rf + (size_t)(y0 + r - 1) * cw + col);
            v[7] = *(const uint32_t *)(rf + (size_t)(y0 + r + 2) * cw + col);
            v[8] = *(const uint32_t *)(rf + (size_t)(y0 + r - 2) * cw + col);
        } else {
            constexpr int DX[9] = {0, 1, -1, 2, -2, 0, 0, 0, 0}, DY[9] = {0, 0, 0, 0, 0, 1, -1, 2, -2};
#pragma unroll
            for (int c = 0; c < 9; c++) v[c] = load4_clamped(rf, cw, ch, col + DX[c], y0 + r + DY[c]);
        }
        score = mx_satd(v[0], K, c2);
#pragma unroll
        for (int c = 1; c < 9; c++) score = min(score, mx_satd(v[c], K, c2));
    }
    return score;
}

__global__ __launch_bounds__(256) void k_rc_complexity(const uint8_t *__restrict__ cur, const uint8_t *__restrict__ ref,
                                                       size_t frame_stride, int cw, int ch,
                                                       unsigned long long *__restrict__ total)
{
    __shared__ unsigned long long s_sum[4];
    const int mbw = cw >> 4, nmb = mbw * (ch >> 4);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // logical workgroup ids renumbered per XCD, so that neighbouring macroblocks, whose
    // reference rows share cache lines, meet in one L2
    unsigned wgx, wgy;
    xcd_grid2(wgx, wgy);
    const int frame = (int)wgy;
    const uint8_t *cf = cur + (size_t)frame * frame_stride;
    const uint8_t *rf = ref ? ref + (size_t)frame * frame_stride : nullptr;
    const MxConst K = mx_const(lane);
    unsigned long long sum = 0; // wave-uniform
    for (int m = (int)wgx * 4 + wave; m < nmb; m += gridDim.x * 4)
        sum += (unsigned)rc_mb_score(cf, rf, cw, ch, (m % mbw) << 4, (m / mbw) << 4, lane, K);
    if (lane == 0) s_sum[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(total + frame, s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3]);
}

__global__ void k_rc_finalize(const unsigned long long *__restrict__ total, int n, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = total[i] < 1ull ? 1.0 : (double)total[i];
}

// ---------------------------------------------------------------------------
// batched multi-stream search
// ---------------------------------------------------------------------------
//
// A batch is a list of segments, one per (stream, reference) for one partition
// shape.  Segments have their own frame size and pointers; full-search lanes and
// blocks are numbered consecutively across segments, and each lane or lane group
// finds its segment by binary search over the segment start indices.

struct SegDesc {
    const uint8_t *cur;    // current frame of the stream (cw x ch)
    const uint8_t *planes; // plane set of this reference
    const int32_t *mvps;   // per-block rate predictors of this stream and shape, or null
    int32_t cw, ch, nbx, nblk;
    int64_t lane0;         // first integer-search lane of the segment
    int64_t blk0;          // first block of the segment in the key / per-reference arrays
};

struct SelDesc {           // one stream's reference selection for one shape
    int64_t ref_blk0[4];   // blk0 of the stream's segment for each reference
    int64_t sel0;          // first block of the stream in the selection numbering
    int32_t *out;          // {ref, mvx, mvy, cost} per block
    int32_t nblk, nrefs;
    int32_t lr_ref[4];     // (int64)(lambda * ref_bits(r, nrefs)), mb16.rs:4951-4956, 5967-5980
};

struct LrTab {
    int32_t v[RH264_LR_N];
};

__device__ __forceinline__ int seg_find_lane(const SegDesc *segs, int nseg, long gi)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].lane0 <= gi) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int seg_find_blk(const SegDesc *segs, int nseg, long gb)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].blk0 <= gb) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Integer search over all segments.  persistent != 0: the grid is sized to the
// machine and blocks take 256-lane tiles from a work counter until none is left;
// otherwise block i does tile i.
template <int BW, int BH, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RH264_FS_WAVES)))
void k_fs_multi(const SegDesc *__restrict__ segs, int nseg, long total_lanes, int R, int pad, LrTab lr_fp,
                uint64_t *__restrict__ keys, unsigned *__restrict__ counter, int persistent)
{
    __shared__ int32_t s_lr[RH264_LR_N];
    __shared__ unsigned s_tile;
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = lr_fp.v[threadIdx.x];
    const int NG = fs_groups<V>(2 * R + 1);
    for (unsigned it = 0;; it++) {
        unsigned tile;
        if (persistent) {
            __syncthreads(); // s_tile of the previous round has been read by every thread
            if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
            __syncthreads();
            tile = s_tile;
        } else {
            __syncthreads();
            if (it > 0) break;
            tile = xcd_linear(blockIdx.x, gridDim.x); // consecutive tiles share one L2
        }
        if ((long)tile * 256 >= total_lanes) break;
        const long gi = (long)tile * 256 + threadIdx.x;
        const bool inrange = gi < total_lanes;
        const int si = seg_find_lane(segs, nseg, inrange ? gi : total_lanes - 1);
        const SegDesc S = segs[si];
        const long local = (inrange ? gi : total_lanes - 1) - S.lane0;
        const int b = (int)(local / NG);
        const bool active = inrange && b < S.nblk;
        const int bs = min(b, S.nblk - 1);
        FsJob J;
        J.cw = S.cw;
        J.pw = S.cw + 2 * pad;
        J.ph = S.ch + 2 * pad;
        J.pad = pad;
        J.R = R;
        J.g = (int)(local - (long)b * NG);
        J.lx = (bs % S.nbx) * BW;
        J.ly = (bs / S.nbx) * BH;
        J.cx = 0;
        J.cy = 0;
        J.mpx = S.mvps ? S.mvps[2 * bs] : 0;
        J.mpy = S.mvps ? S.mvps[2 * bs + 1] : 0;
        J.F = S.planes;
        J.srcp = S.cur + (size_t)J.ly * S.cw + J.lx;
        const uint64_t key = fs_key<BW, BH, V>(J, s_lr);
        fs_commit(key, S.blk0 + bs, active, keys + S.blk0 + bs);
    }
}

// ---------------------------------------------------------------------------
// fused integer search of the macroblock shapes
// ---------------------------------------------------------------------------
//
// The 16x16, 16x8, 8x16, 8x8 and 4x4 searches of a batch scan the same +-R window of
// the same macroblocks, so one pass serves all of them.  A row of 16 lanes owns one
// unit (macroblock, column group g); lane sb of the row owns the 4x4 sub-block whose
// position bits are sb = x0 | y0 << 1 | x1 << 2 | y1 << 3 (x0, y0: the 4x4 inside its
// 8x8 quadrant; x1, y1: the quadrant).  Each lane computes its sub-block's quad-SADs
// for candidate columns 4g..4g+3 row by row, as the 4x4 search does.  When a
// candidate row completes, DPP adds inside the 16-lane row form the larger
// partitions' SADs: the xor-1 and xor-2 quad permutations sum a quadrant (8x8), the
// half-row mirror pairs quadrants 0-1 and 2-3 (16x8), rotation by 8 pairs 0-2 and
// 1-3 (8x16), and both together give the 16x16.  The four u16 SADs of a lane sit two
// to a dword and no partial sum exceeds 256 * 255 < 2^16, so plain dword adds never
// carry from one into the other.  Every lane prices its 4x4 block; nine lanes also
// price one larger partition (sb 0: the 16x16; sb 1, 5, 9, 13: the 8x8 of their
// quadrant; sb 2 / 10: the top / bottom 16x8; sb 3 / 7: the left / right 8x16).
// Costs, predictors and the tie order are those of each shape's own search (SAD +
// lr_fp[bits(x) + bits(y)], first minimum in raster candidate order), so every
// shape's keys equal what its own k_fs_multi launch writes.  A lane keeps its two
// running minima as 32-bit keys cost << 10 | t << 2 | k (t: candidate row, k: the
// lane's column), which needs cost < 2^22: a batch fuses only when every lr_fp entry
// is at most 2^21.  Columns past 2R (the last group's lanes) price against the
// sentinel entries 128.. of the LDS table (2^21 + 2^16, above every real cost).

constexpr int FU_NS = 5;               // fused shapes, in this order: 16x16, 16x8, 8x16, 8x8, 4x4
constexpr int FU_LR_MAX = 1 << 21;     // largest lr_fp entry a fused batch accepts
constexpr int FU_SENTINEL = (1 << 21) + (1 << 16);

struct FuSeg {                         // one (stream, reference) of a fused batch
    const uint8_t *cur;
    const uint8_t *planes;
    const int32_t *mvps[FU_NS];        // the stream's block predictors per shape, or null
    int64_t unit0;                     // first unit of the segment
    int64_t blk0[FU_NS];               // first block of the segment in each shape's keys; -1: not searched
    int32_t cw, ch, mbw, pad;
};

struct FuKeys {
    uint64_t *k[FU_NS];
};

struct FuLane {
    uint32_t src[4];
    uint64_t acc[4];
    uint32_t bestA, bestB;             // 4x4 block; the lane's larger partition
    int rxA[4], rxB[4];                // 4 x column bits per candidate column (4 x 128: past 2R)
    int mpyA, mpyB;
    uint32_t mhs, mvs;                 // all-ones where the lane's mirror / row-rotation partner adds this lane's sum
    int ryc;                           // without predictors: the y bits lrc was read for (wave-uniform)
    uint32_t lrc[4];
    const uint8_t *row;                // next reference row (the aligned dword left of the lane's window)
};

template <bool MVP>
__device__ __forceinline__ void fu_price(FuLane &L, uint64_t a4, int t, int R, const int32_t *s_lr)
{
    const uint32_t lo = (uint32_t)a4, hi = (uint32_t)(a4 >> 32);
    uint32_t l8 = lo + (uint32_t)__builtin_amdgcn_mov_dpp((int)lo, 0xB1, 0xF, 0xF, true);
    uint32_t h8 = hi + (uint32_t)__builtin_amdgcn_mov_dpp((int)hi, 0xB1, 0xF, 0xF, true);
    l8 += (uint32_t)__builtin_amdgcn_mov_dpp((int)l8, 0x4E, 0xF, 0xF, true); // the lane's quadrant (8x8)
    h8 += (uint32_t)__builtin_amdgcn_mov_dpp((int)h8, 0x4E, 0xF, 0xF, true);
    // Each lane sends its sum to its half-row mirror partner (the quadrant beside)
    // only where that partner pairs horizontally, then the pair sums to the lane 8
    // apart (the quadrant below / above) only where that one pairs vertically: sb 0
    // gets all four quadrants, 2 / 10 the top / bottom pair, 3 / 7 the left / right
    // pair, the others their own quadrant.
    const uint32_t lp = l8 + (uint32_t)__builtin_amdgcn_mov_dpp((int)(l8 & L.mhs), 0x141, 0xF, 0xF, true);
    const uint32_t hp = h8 + (uint32_t)__builtin_amdgcn_mov_dpp((int)(h8 & L.mhs), 0x141, 0xF, 0xF, true);
    const uint32_t lB = lp + (uint32_t)__builtin_amdgcn_mov_dpp((int)(lp & L.mvs), 0x128, 0xF, 0xF, true);
    const uint32_t hB = hp + (uint32_t)__builtin_amdgcn_mov_dpp((int)(hp & L.mvs), 0x128, 0xF, 0xF, true);
    // without predictors |4 (t - R)| <= 512, so mvbits' clamp is not needed
    const int dyA = 4 * (t - R) - L.mpyA;
    const int ryA = MVP ? 4 * mvbits(dyA)
                        : 4 * (1 + 2 * (31 - __clz((int)((dyA > 0 ? 2u * dyA - 1u : -2u * dyA) + 1u))));
    const int ryB = MVP ? 4 * mvbits(4 * (t - R) - L.mpyB) : ryA;
    const char *lr = (const char *)s_lr; // byte offsets: 4 x (bits(x) + bits(y))
    if (!MVP && ryA != L.ryc) {
        // no predictors: the y bits depend on t alone, so the wave re-reads its four
        // rates only when they change
#pragma unroll
        for (int k = 0; k < 4; k++) L.lrc[k] = *(const uint32_t *)(lr + L.rxA[k] + ryA);
        L.ryc = ryA;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t tk = ((uint32_t)t << 2) | (uint32_t)k;
        const uint32_t sA = ((k < 2 ? lo : hi) >> (16 * (k & 1))) & 0xffffu;
        const uint32_t sB = ((k < 2 ? lB : hB) >> (16 * (k & 1))) & 0xffffu;
        const uint32_t lrA = MVP ? *(const uint32_t *)(lr + L.rxA[k] + ryA) : L.lrc[k];
        const uint32_t lrB = MVP ? *(const uint32_t *)(lr + L.rxB[k] + ryB) : lrA;
        L.bestA = min(L.bestA, ((sA + lrA) << 10) | tk);
        L.bestB = min(L.bestB, ((sB + lrB) << 10) | tk);
    }
}

// Rows jb .. jb+3 of the lane's window (MODE as in fs_rows: 1 = first ring block,
// 2 = last, 0 = any).  FAST: every lane of the wave has its whole window inside the
// padded plane and walks L.row down it; otherwise every lane reads clamped.
template <bool MVP, int MODE, bool FAST>
__device__ __forceinline__ void fu_rows(FuLane &L, const uint8_t *__restrict__ F, int pw, int ph, int px0, int py0,
                                        int shift, int jb, int NJ, int R, const int32_t *s_lr)
{
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
        const int j = jb + jj;
        if (MODE == 0 && j >= NJ) break;
        uint32_t w0, w1;
        const int ry = py0 + j;
        if (FAST) {
            const uint32_t *rp = (const uint32_t *)L.row;
            const uint32_t a0 = rp[0], a1 = rp[1], a2 = rp[2];
            L.row += pw;
            w0 = __builtin_amdgcn_alignbyte(a1, a0, shift);
            w1 = __builtin_amdgcn_alignbyte(a2, a1, shift);
        } else {
            w0 = load4_clamped(F, pw, ph, px0, ry);
            w1 = load4_clamped(F, pw, ph, px0 + 4, ry);
        }
        const uint64_t ref8 = ((uint64_t)w1 << 32) | w0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (MODE == 1 && k > jj) continue;
            if (MODE == 2 && k < jj) continue;
            const int slot = (jj - k + 4) % 4;
            L.acc[slot] = __builtin_amdgcn_qsad_pk_u16_u8(ref8, L.src[k], L.acc[slot]);
        }
        const int t = j - 3;
        const int sd = (jj + 1) % 4;
        if (t >= 0) fu_price<MVP>(L, L.acc[sd], t, R, s_lr);
        L.acc[sd] = 0;
    }
}

#ifndef RH264_FU_WAVES
#define RH264_FU_WAVES 8 // 8 waves per SIMD (64 VGPRs): 16.2 ms against 19.6 ms at the compiler's 101
#endif
template <bool MVP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RH264_FU_WAVES)))
void k_fs_fused(const FuSeg *__restrict__ segs, int nseg, long total_units, int R, LrTab lr_fp, FuKeys keys)
{
    __shared__ int32_t s_lr[2 * RH264_LR_N];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = lr_fp.v[threadIdx.x];
    else s_lr[threadIdx.x] = FU_SENTINEL; // 256 threads: entries 128..255
    __syncthreads();
    const int NC = 2 * R + 1, NG = (NC + 3) / 4;
    const int lane = threadIdx.x & 63, sb = threadIdx.x & 15;
    const long u0 = ((long)blockIdx.x * 256 + threadIdx.x) >> 4;
    const bool inrange = u0 < total_units;
    const long unit = inrange ? u0 : total_units - 1;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].unit0 <= unit) lo = mid; else hi = mid - 1;
    }
    const FuSeg &S = segs[lo];
    const long local = unit - S.unit0;
    const int mb = (int)(local / NG), g = (int)(local - (long)mb * NG);
    const int mbw = S.mbw, mbx = mb % mbw, mby = mb / mbw;
    const int x4 = (sb & 1) | ((sb >> 1) & 2), y4 = ((sb >> 1) & 1) | ((sb >> 2) & 2);
    const int lx = 16 * mbx + 4 * x4, ly = 16 * mby + 4 * y4;
    const int cw = S.cw, pad = S.pad, pw = cw + 2 * pad, ph = S.ch + 2 * pad;
    // the lane's larger partition (shape index and block index in that shape's grid)
    int shB = -1, bB = 0;
    if (sb == 0) {
        shB = 0;
        bB = mb;
    } else if ((sb & 3) == 1) {
        shB = 3;
        bB = (2 * mby + (sb >> 3)) * (2 * mbw) + 2 * mbx + ((sb >> 2) & 1);
    } else if (sb == 2 || sb == 10) {
        shB = 1;
        bB = (2 * mby + (sb >> 3)) * mbw + mbx;
    } else if (sb == 3 || sb == 7) {
        shB = 2;
        bB = mby * (2 * mbw) + 2 * mbx + (sb >> 2);
    }
    const int bA = (4 * mby + y4) * (4 * mbw) + 4 * mbx + x4;
    FuLane L;
    {
        // partners: half-row mirror (sb & 8) | (7 - (sb & 7)), row rotation sb ^ 8;
        // horizontal pairs: sb 0, 2, 10 and 8 (which carries the bottom pair to sb 0);
        // vertical pairs: sb 0, 3, 7
        const int mir = (sb & 8) | (7 - (sb & 7)), rot = sb ^ 8;
        L.mhs = (mir == 0 || mir == 2 || mir == 8 || mir == 10) ? ~0u : 0u;
        L.mvs = (rot == 0 || rot == 3 || rot == 7) ? ~0u : 0u;
    }
    L.ryc = -1;
    for (int k = 0; k < 4; k++) L.lrc[k] = 0;
    int mpxA = 0, mpxB = 0;
    L.mpyA = 0;
    L.mpyB = 0;
    if (MVP) {
        const int32_t *pa = S.mvps[FU_NS - 1];
        if (pa) {
            mpxA = pa[2 * bA];
            L.mpyA = pa[2 * bA + 1];
        }
        const int32_t *pb = shB >= 0 ? S.mvps[shB] : nullptr;
        if (pb) {
            mpxB = pb[2 * bB];
            L.mpyB = pb[2 * bB + 1];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) L.src[k] = *(const uint32_t *)(S.cur + (size_t)(ly + k) * cw + lx);
    const int dx0 = -R + 4 * g;
    const int px0 = lx + pad + dx0, py0 = ly + pad - R;
    const int wx_lo = lx + pad - R, wx_hi = wx_lo + 2 * R + 3;
    const bool fast = wx_lo >= 0 && wx_hi <= pw - 1 && py0 >= 0 && py0 + 2 * R + 3 <= ph - 1;
    const int shift = px0 & 3;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool valid = 4 * g + k < NC;
        L.rxA[k] = 4 * (valid ? mvbits(4 * (dx0 + k) - mpxA) : RH264_LR_N);
        L.rxB[k] = 4 * (valid ? mvbits(4 * (dx0 + k) - mpxB) : RH264_LR_N);
        L.acc[k] = 0;
    }
    L.bestA = 0xffffffffu;
    L.bestB = 0xffffffffu;
    const uint8_t *F = S.planes;
    L.row = F + (size_t)py0 * pw + (px0 & ~3);
    const int NJ = 2 * R + 4;
    // one branch for the wave: the DPP sums need every lane on the same path
    if (__all(fast)) {
        if ((2 * R) % 4 == 0) {
            fu_rows<MVP, 1, true>(L, F, pw, ph, px0, py0, shift, 0, NJ, R, s_lr);
            for (int jb = 4; jb < NJ - 4; jb += 4) fu_rows<MVP, 0, true>(L, F, pw, ph, px0, py0, shift, jb, NJ, R, s_lr);
            fu_rows<MVP, 2, true>(L, F, pw, ph, px0, py0, shift, NJ - 4, NJ, R, s_lr);
        } else {
            for (int jb = 0; jb < NJ; jb += 4) fu_rows<MVP, 0, true>(L, F, pw, ph, px0, py0, shift, jb, NJ, R, s_lr);
        }
    } else {
        if ((2 * R) % 4 == 0) {
            fu_rows<MVP, 1, false>(L, F, pw, ph, px0, py0, shift, 0, NJ, R, s_lr);
            for (int jb = 4; jb < NJ - 4; jb += 4)
                fu_rows<MVP, 0, false>(L, F, pw, ph, px0, py0, shift, jb, NJ, R, s_lr);
            fu_rows<MVP, 2, false>(L, F, pw, ph, px0, py0, shift, NJ - 4, NJ, R, s_lr);
        } else {
            for (int jb = 0; jb < NJ; jb += 4) fu_rows<MVP, 0, false>(L, F, pw, ph, px0, py0, shift, jb, NJ, R, s_lr);
        }
    }
    // the keys the separate searches write: cost << 32 | (t * NC + c)
    uint64_t kA = ((uint64_t)(L.bestA >> 10) << 32) | (uint32_t)(((L.bestA >> 2) & 255u) * NC + 4 * g + (L.bestA & 3u));
    uint64_t kB = ((uint64_t)(L.bestB >> 10) << 32) | (uint32_t)(((L.bestB >> 2) & 255u) * NC + 4 * g + (L.bestB & 3u));
    // minimum over the wave's units of the same macroblock (segments start on whole
    // macroblocks, so unit / NG numbers the macroblocks of the batch), then one
    // atomic per macroblock run and partition
    const int gmb = (int)(unit / NG);
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
        const uint64_t oA = __shfl_down(kA, off, 64), oB = __shfl_down(kB, off, 64);
        const int og = __shfl_down(gmb, off, 64);
        if (lane + off < 64 && og == gmb) {
            kA = oA < kA ? oA : kA;
            kB = oB < kB ? oB : kB;
        }
    }
    const int up = __shfl_up(gmb, 16, 64);
    if (!inrange || (lane >= 16 && up == gmb)) return;
    if (S.blk0[FU_NS - 1] >= 0) atomicMin((unsigned long long *)(keys.k[FU_NS - 1] + S.blk0[FU_NS - 1] + bA), kA);
    if (shB >= 0 && S.blk0[shB] >= 0) atomicMin((unsigned long long *)(keys.k[shB] + S.blk0[shB] + bB), kB);
}

// Refinement of every block of every segment from its integer result; writes
// {mvx, mvy, cost} per block into res.
template <int BW, int BH, bool USE_LDS>
__global__ __launch_bounds__(256) void k_refine_multi(const SegDesc *__restrict__ segs, int nseg, long total_blocks,
                                                      int R, int pad, int pat, LrTab lr,
                                                      const uint64_t *__restrict__ keys, int32_t *__restrict__ res)
{
    constexpr int NSB = RefShape<BW, BH>::NSB, GPW = RefShape<BW, BH>::GPW;
    using W = RefineWin<BW, BH>;
    __shared__ int32_t s_lr[RH264_LR_N];
    __shared__ __align__(16) uint8_t s_win[USE_LDS ? GPW * W::BYTES : 16];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = lr.v[threadIdx.x];
    const int grp = threadIdx.x / NSB, li = threadIdx.x % NSB;
    // without LDS windows, workgroups are renumbered per XCD: consecutive block strips,
    // which read the same plane rows, share one L2 (4x4: 44.5 -> 0.9 GB fetched past L2
    // per launch); the staged forms run 3-4 % faster in dispatch order
    const long gb0 = (long)(USE_LDS ? blockIdx.x : xcd_linear(blockIdx.x, gridDim.x)) * GPW + grp;
    const bool active = gb0 < total_blocks;
    const long gb = active ? gb0 : total_blocks - 1;
    const int si = seg_find_blk(segs, nseg, gb);
    const SegDesc S = segs[si];
    const int b = (int)(gb - S.blk0);
    const int lx = (b % S.nbx) * BW, ly = (b / S.nbx) * BH;
    const int sbx = 4 * (li % (BW / 4)), sby = 4 * (li / (BW / 4));
    RefineGeom G;
    G.pw = S.cw + 2 * pad;
    G.ph = S.ch + 2 * pad;
    G.pad = pad;
    G.n = (size_t)G.pw * G.ph;
    G.pl = S.planes;
    uint32_t s[4];
    load_src4<BW, BH>(S.cur + (size_t)(ly + sby) * S.cw + lx + sbx, S.cw, s);
    const int mpx = S.mvps ? S.mvps[2 * b] : 0, mpy = S.mvps ? S.mvps[2 * b + 1] : 0;
    const uint64_t key = keys[gb];
    const int NC = 2 * R + 1;
    const uint32_t idx = (uint32_t)key;
    int bx = 4 * ((int)(idx % (uint32_t)NC) - R), by = 4 * ((int)(idx / (uint32_t)NC) - R);
    uint8_t *win = s_win + (USE_LDS ? grp * W::BYTES : 0);
    const int sxi = bx >> 2, syi = by >> 2;
    if (USE_LDS) stage_window<BW, BH>(win, G, lx + sxi + pad - 4, ly + syi + pad - 3, li, NSB);
    __syncthreads();
    if (!active) return;
    int best, evals;
    refine_walk<BW, BH, USE_LDS>(G, win, s, lx, ly, sbx, sby, sxi, syi, bx, by, mpx, mpy, pat, s_lr, best, evals);
    if (li == 0) {
        res[3 * gb] = bx;
        res[3 * gb + 1] = by;
        res[3 * gb + 2] = best;
    }
}

// best_part's reference loop (mb16.rs:4938-4957): cost + lr_ref[r], first minimum
// over references in order.
__global__ void k_select_ref(const SelDesc *__restrict__ sels, int nsel, long total, const int32_t *__restrict__ res)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int lo = 0, hi = nsel - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sels[mid].sel0 <= i) lo = mid; else hi = mid - 1;
    }
    const SelDesc &S = sels[lo];
    const int b = (int)(i - S.sel0);
    long long bc = 0x7fffffffffffffffLL;
    int br = 0, bmx = 0, bmy = 0;
    for (int rr = 0; rr < S.nrefs; rr++) {
        const long g = S.ref_blk0[rr] + b;
        const long long c = (long long)res[3 * g + 2] + S.lr_ref[rr];
        if (c < bc) {
            bc = c;
            br = rr;
            bmx = res[3 * g];
            bmy = res[3 * g + 1];
        }
    }
    int32_t *o = S.out + 4 * (size_t)b;
    o[0] = br;
    o[1] = bmx;
    o[2] = bmy;
    o[3] = (int32_t)bc;
}

__global__ void k_fill_u32(unsigned *p, size_t n, unsigned v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// batched block metrics (the accel per-block family)
// ---------------------------------------------------------------------------

__device__ int satd4x4_sum_abs(const int d[16])
{
    int m[16];
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
        int a = d[rr * 4], b = d[rr * 4 + 1], c = d[rr * 4 + 2], e = d[rr * 4 + 3];
        m[rr * 4] = a + b + c + e; m[rr * 4 + 1] = a + b - c - e;
        m[rr * 4 + 2] = a - b - c + e; m[rr * 4 + 3] = a - b + c - e;
    }
    int s = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int a = m[c], b = m[4 + c], cc = m[8 + c], e = m[12 + c];
        s += abs(a + b + cc + e) + abs(a + b - cc - e) + abs(a - b - cc + e) + abs(a - b + cc - e);
    }
    return s;
}

// sum |H8 d H8^T| of an 8x8 difference, H8 the order-8 Sylvester Hadamard: three
// butterfly stages on the rows, three on the columns (even: every coefficient has
// the parity of sum d)
__device__ __forceinline__ int satd8x8_sum_abs(int d[64])
{
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
        const int st = pass ? 8 : 1, ln = pass ? 1 : 8; // element stride, line stride
#pragma unroll
        for (int l = 0; l < 8; l++)
#pragma unroll
            for (int half = 4; half >= 1; half >>= 1)
#pragma unroll
                for (int i = 0; i < 8; i++)
                    if (!(i & half)) {
                        const int a = d[l * ln + i * st], b = d[l * ln + (i + half) * st];
                        d[l * ln + i * st] = a + b;
                        d[l * ln + (i + half) * st] = a - b;
                    }
    }
    int s = 0;
#pragma unroll
    for (int k = 0; k < 64; k++) s += abs(d[k]);
    return s;
}

__global__ void k_block_metric(int kind, int w, int h, int n, const uint8_t *__restrict__ src, int ss,
                               const int64_t *__restrict__ src_off, const uint8_t *__restrict__ ref,
                               int rs, const int64_t *__restrict__ ref_off,
                               const int64_t *__restrict__ ref_off2, int32_t *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t *a = src + src_off[i];
    const uint8_t *b = ref + ref_off[i];
    const uint8_t *b2 = kind == RH264_METRIC_SATD_AVG ? ref + ref_off2[i] : nullptr;
    int total = 0;
    if (kind == RH264_METRIC_SAD) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) total += abs((int)a[y * ss + x] - (int)b[(size_t)y * rs + x]);
    } else if (kind == RH264_METRIC_SATD8) {
        for (int by = 0; by < h; by += 8)
            for (int bx = 0; bx < w; bx += 8) {
                int d[64];
                for (int y = 0; y < 8; y++)
                    for (int x = 0; x < 8; x++)
                        d[y * 8 + x] = (int)a[(by + y) * ss + bx + x] - (int)b[(size_t)(by + y) * rs + bx + x];
                total += satd8x8_sum_abs(d) >> 1;
            }
    } else {
        for (int by = 0; by < h; by += 4)
            for (int bx = 0; bx < w; bx += 4) {
                int d[16];
                for (int y = 0; y < 4; y++)
                    for (int x = 0; x < 4; x++) {
                        size_t o = (size_t)(by + y) * rs + bx + x;
                        int pv = b2 ? (((int)b[o] + (int)b2[o] + 1) >> 1) : (int)b[o];
                        d[y * 4 + x] = (int)a[(by + y) * ss + bx + x] - pv;
                    }
                int s = satd4x4_sum_abs(d);
                total += kind == RH264_METRIC_SATD_WELS ? (s + 1) >> 1 : s;
            }
    }
    out[i] = total;
}

// ---------------------------------------------------------------------------
// block prediction from a plane set
// ---------------------------------------------------------------------------

__global__ void k_predict(const uint8_t *__restrict__ planes, int pw, int ph, int pad, int bw, int bh,
                          int n, const int32_t *__restrict__ xy, const int32_t *__restrict__ mv,
                          uint8_t *__restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n * bw * bh;
    if (t >= total) return;
    const int i = (int)(t / (bw * bh));
    const int pix = (int)(t - (long)i * bw * bh);
    const int yy = pix / bw, xx = pix - yy * bw;
    const int mvx = mv[2 * i], mvy = mv[2 * i + 1];
    const int x = xy[2 * i] + (mvx >> 2) + pad + xx, y = xy[2 * i + 1] + (mvy >> 2) + pad + yy;
    const size_t np = (size_t)pw * ph;
    const uint32_t q = qtab((mvx & 3) * 4 + (mvy & 3));
    const uint8_t *pa = planes + np * (q & 15);
    int v = pa[(size_t)clampi(y + ((q >> 8) & 15), 0, ph - 1) * pw + clampi(x + ((q >> 4) & 15), 0, pw - 1)];
    if ((mvx | mvy) & 1) {
        const uint8_t *pb = planes + np * ((q >> 12) & 15);
        int u = pb[(size_t)clampi(y + ((q >> 20) & 15), 0, ph - 1) * pw + clampi(x + ((q >> 16) & 15), 0, pw - 1)];
        v = (v + u + 1) >> 1;
    }
    out[t] = (uint8_t)v;
}

// ---------------------------------------------------------------------------
// mb-tree lookahead (crates/rusty_h264-encoder/src/mbtree.rs), HalfRes mode
// ---------------------------------------------------------------------------

// downsample2x, mbtree.rs:61-74: (a+b+c+d+2)/4, one thread per 4 output pixels.
__global__ void k_downsample2x(const uint8_t *__restrict__ frames, size_t stride, int cw, int ch,
                               uint8_t *__restrict__ half)
{
    const int hw = cw / 2, hh = ch / 2;
    const long i4 = (long)blockIdx.x * blockDim.x + threadIdx.x; // index of a 4-pixel output group
    const int f = blockIdx.y;
    if (i4 * 4 >= (long)hw * hh) return;
    const int j = (int)(i4 * 4 / hw), i = (int)(i4 * 4 - (long)j * hw);
    const uint8_t *y = frames + (size_t)f * stride;
    const uint8_t *r0 = y + (size_t)(2 * j) * cw + 2 * i, *r1 = r0 + cw;
    const uint2 a = *(const uint2 *)r0, b = *(const uint2 *)r1; // 8 source bytes per row
    uint32_t o = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t aw = k < 2 ? a.x : a.y, bw = k < 2 ? b.x : b.y;
        const int sh = 16 * (k & 1);
        const uint32_t sum = ((aw >> sh) & 255) + ((aw >> (sh + 8)) & 255) + ((bw >> sh) & 255) +
                             ((bw >> (sh + 8)) & 255);
        o |= ((sum + 2) / 4) << (8 * k);
    }
    *(uint32_t *)(half + (size_t)f * hw * hh + (size_t)j * hw + i) = o;
}

// Per-lane rows of one 8x8 block: lane li (0..15) = 4x4 sub-block li>>2, row li&3.
__device__ __forceinline__ uint32_t mbt_row_clamped(const uint8_t *plane, int w, int h, int x, int y)
{
    const uint8_t *row = plane + (size_t)clampi(y, 0, h - 1) * w;
    uint32_t v = 0;
#pragma unroll
    for (int m = 0; m < 4; m++) v |= (uint32_t)row[clampi(x + m, 0, w - 1)] << (8 * m);
    return v;
}


// intra_cost, mbtree.rs:78-93, 8x8 blocks of the half-res plane: sum over the four
// 4x4 Hadamards of |coeff| without the DC coefficient, clamped to >= 1.  Frame 0
// (the IDR) also gets inter = intra and mv = 0 (mbtree.rs:256).
__global__ void k_mbtree_intra(const uint8_t *__restrict__ half, int mb_w, int mb_h, int32_t *__restrict__ intra,
                               int32_t *__restrict__ inter, int32_t *__restrict__ mv)
{
    const int f = blockIdx.y;
    const int nm = mb_w * mb_h;
    const int m0 = blockIdx.x * 16 + (threadIdx.x >> 4);
    const bool active = m0 < nm;
    const int m = active ? m0 : nm - 1;
    const int li = threadIdx.x & 15, bi = li >> 2, r = li & 3;
    const int cwh = mb_w * 8;
    const uint8_t *pl = half + (size_t)f * cwh * (mb_h * 8);
    const int bx0 = (m % mb_w) * 8 + 4 * (bi & 1), by0 = (m / mb_w) * 8 + 4 * (bi >> 1) + r;
    const uint32_t s = *(const uint32_t *)(pl + (size_t)by0 * cwh + bx0);
    int d0, d1, d2, d3;
    diff4(s, 0u, d0, d1, d2, d3);
    // row transform, column transform across the quad (as hadamard_row_quad) but
    // keeping the coefficient that is the 4x4 block's DC (lane r = 0, column 0)
    int e0 = d0 + d1 + d2 + d3, e1 = d0 + d1 - d2 - d3, e2 = d0 - d1 - d2 + d3, e3 = d0 - d1 + d2 - d3;
    const bool hi2 = (r & 2) != 0, odd = (r & 1) != 0;
    int y0 = __builtin_amdgcn_mov_dpp(e0, 0x4E, 0xF, 0xF, true), y1 = __builtin_amdgcn_mov_dpp(e1, 0x4E, 0xF, 0xF, true);
    int y2 = __builtin_amdgcn_mov_dpp(e2, 0x4E, 0xF, 0xF, true), y3 = __builtin_amdgcn_mov_dpp(e3, 0x4E, 0xF, 0xF, true);
    int s0 = hi2 ? y0 - e0 : e0 + y0, s1 = hi2 ? y1 - e1 : e1 + y1, s2 = hi2 ? y2 - e2 : e2 + y2,
        s3 = hi2 ? y3 - e3 : e3 + y3;
    int z0 = __builtin_amdgcn_mov_dpp(s0, 0xB1, 0xF, 0xF, true), z1 = __builtin_amdgcn_mov_dpp(s1, 0xB1, 0xF, 0xF, true);
    int z2 = __builtin_amdgcn_mov_dpp(s2, 0xB1, 0xF, 0xF, true), z3 = __builtin_amdgcn_mov_dpp(s3, 0xB1, 0xF, 0xF, true);
    int c0 = odd ? z0 - s0 : s0 + z0, c1 = odd ? z1 - s1 : s1 + z1, c2 = odd ? z2 - s2 : s2 + z2,
        c3 = odd ? z3 - s3 : s3 + z3;
    int v = (r == 0 ? 0 : abs(c0)) + abs(c1) + abs(c2) + abs(c3); // lane r=0, c0 = sum of all 16 = DC
    v = group_sum<16>(v);
    if (active && li == 0) {
        const int ic = max(v, 1);
        const size_t g = (size_t)f * nm + m;
        intra[g] = ic;
        if (f == 0) {
            inter[g] = ic;
            mv[2 * g] = 0;
            mv[2 * g + 1] = 0;
        }
    }
}

// Four candidate SATDs at once (one per 16-lane slot) for the 8x8 block at
// (bx0, by0); candidate of slot s = (mvx[s], mvy[s]), full-pel, reads clamped to
// the plane (mc_luma's full-pel copy, inter.rs:1214-1243).  Returns the four
// totals through scalars.
__device__ __forceinline__ void mbt_eval4(const uint8_t *cur, const uint8_t *ref, int w, int h, int bx0, int by0,
                                          uint32_t srow, int li, int slot, int mvx, int mvy, int out[4])
{
    const int bi = li >> 2, r = li & 3;
    const int x = bx0 + (mvx >> 2) + 4 * (bi & 1), y = by0 + (mvy >> 2) + 4 * (bi >> 1) + r;
    const uint32_t p = mbt_row_clamped(ref, w, h, x, y);
    int d0, d1, d2, d3;
    diff4(srow, p, d0, d1, d2, d3);
    int v = group_sum<16>(hadamard_row_quad(d0, d1, d2, d3, r));
    out[0] = __builtin_amdgcn_readlane(v, 0);
    out[1] = __builtin_amdgcn_readlane(v, 16);
    out[2] = __builtin_amdgcn_readlane(v, 32);
    out[3] = __builtin_amdgcn_readlane(v, 48);
}

// inter_cost (mbtree.rs:154-197) for one block with one wave; the cascade diamond
// evaluates the remaining candidates of a ring pass speculatively in parallel and
// keeps exactly the sequential decisions (first improving candidate in ring order
// moves the centre; later candidates of that round are discarded).
__device__ __forceinline__ void mbt_inter(const uint8_t *cur, const uint8_t *ref, int w, int h, int bx0, int by0,
                                          int seedx, int seedy, int max_step, int &cost, int &bmx, int &bmy)
{
    const int lane = threadIdx.x & 63, slot = lane >> 4, li = lane & 15;
    const int bi = li >> 2, r = li & 3;
    const uint32_t srow = *(const uint32_t *)(cur + (size_t)(by0 + 4 * (bi >> 1) + r) * w + bx0 + 4 * (bi & 1));
    int c4[4];
    // slot 0: (0,0); slot 1: the seed
    mbt_eval4(cur, ref, w, h, bx0, by0, srow, li, slot, slot == 1 ? seedx : 0, slot == 1 ? seedy : 0, c4);
    int best = c4[0];
    bmx = 0;
    bmy = 0;
    if (best != 0) {
        if ((seedx != 0 || seedy != 0) && c4[1] < best) {
            best = c4[1];
            bmx = seedx;
            bmy = seedy;
        }
        const int rx[4] = {1, -1, 0, 0}, ry[4] = {0, 0, 1, -1};
        for (int step = max_step; step >= 1; step >>= 1) {
            for (;;) {
                bool moved = false;
                int i = 0;
                while (i < 4) {
                    const int k = min(i + slot, 3); // slots past the ring end repeat the last candidate
                    mbt_eval4(cur, ref, w, h, bx0, by0, srow, li, slot, bmx + rx[k] * step * 4, bmy + ry[k] * step * 4,
                              c4);
                    int acc = -1;
                    for (int sidx = 0; sidx < 4 - i; sidx++)
                        if (c4[sidx] < best) { acc = sidx; break; }
                    if (acc < 0) break;
                    const int kk = i + acc;
                    best = c4[acc];
                    bmx += rx[kk] * step * 4;
                    bmy += ry[kk] * step * 4;
                    moved = true;
                    i = kk + 1;
                }
                if (!moved) break;
            }
        }
    }
    cost = best;
}

// Column 0 of each inter frame: one wave walks rows 0..mb_h-1 (seed = the block
// above's vector, mbtree.rs:224-231).
__global__ __launch_bounds__(64) void k_mbtree_col0(const uint8_t *__restrict__ half, int mb_w, int mb_h,
                                                    const int32_t *__restrict__ intra, int32_t *__restrict__ inter,
                                                    int32_t *__restrict__ mv)
{
    const int f = blockIdx.x + 1;
    const int cwh = mb_w * 8, chh = mb_h * 8, nm = mb_w * mb_h;
    const uint8_t *cur = half + (size_t)f * cwh * chh, *ref = cur - (size_t)cwh * chh;
    int sfx = 0, sfy = 0;
    for (int y = 0; y < mb_h; y++) {
        const size_t g = (size_t)f * nm + (size_t)y * mb_w;
        int c, mx, my;
        mbt_inter(cur, ref, cwh, chh, 0, y * 8, sfx / 2, sfy / 2, 8, c, mx, my);
        const int ic = intra[g];
        sfx = mx * 2;
        sfy = my * 2;
        if ((threadIdx.x & 63) == 0) {
            inter[g] = min(c, ic);
            mv[2 * g] = sfx;
            mv[2 * g + 1] = sfy;
        }
    }
}

// Rows of each inter frame: one wave per (frame, row) walks x = 1..mb_w-1 from the
// column-0 result (seed = the left neighbour's vector).
__global__ __launch_bounds__(64) void k_mbtree_rows(const uint8_t *__restrict__ half, int mb_w, int mb_h,
                                                    const int32_t *__restrict__ intra, int32_t *__restrict__ inter,
                                                    int32_t *__restrict__ mv)
{
    const int f = blockIdx.y + 1, y = blockIdx.x;
    const int cwh = mb_w * 8, chh = mb_h * 8, nm = mb_w * mb_h;
    const uint8_t *cur = half + (size_t)f * cwh * chh, *ref = cur - (size_t)cwh * chh;
    const size_t g0 = (size_t)f * nm + (size_t)y * mb_w;
    int sfx = mv[2 * g0], sfy = mv[2 * g0 + 1];
    for (int x = 1; x < mb_w; x++) {
        const size_t g = g0 + x;
        int c, mx, my;
        mbt_inter(cur, ref, cwh, chh, x * 8, y * 8, sfx / 2, sfy / 2, 8, c, mx, my);
        const int ic = intra[g];
        sfx = mx * 2;
        sfy = my * 2;
        if ((threadIdx.x & 63) == 0) {
            inter[g] = min(c, ic);
            mv[2 * g] = sfx;
            mv[2 * g + 1] = sfy;
        }
    }
}

// Build-defined int32 propagation step for frame f (mirrors propagate_to,
// mbtree.rs:268-297, in integers): prop[f] = saturate(acc[f]); amount =
// ((intra + prop) * (intra - inter)) / intra; shares (amount * wx * wy) >> 8 are
// added into acc[f-1] -- or acc[f - refs[f][mb]] for the multi-reference costs --
// with 64-bit integer atomics (order-independent).
__global__ void k_mbtree_prop_i32(int f, int mb_w, int mb_h, const int32_t *__restrict__ intra,
                                  const int32_t *__restrict__ inter, const int32_t *__restrict__ mv,
                                  const int32_t *__restrict__ refs, unsigned long long *__restrict__ acc,
                                  int32_t *__restrict__ prop)
{
    const int nm = mb_w * mb_h;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const size_t g = (size_t)f * nm + m;
    const long long a = (long long)acc[g];
    const int32_t pf = a > 0x7fffffffLL ? 0x7fffffff : (int32_t)a;
    prop[g] = pf;
    if (f == 0) return;
    const int ic = intra[g];
    const long long amount = ((long long)ic + pf) * (long long)(ic - inter[g]) / ic;
    if (amount <= 0) return;
    const int x = m % mb_w, y = m / mb_w;
    const int rx = clampi(x * 16 + (mv[2 * g] >> 2), 0, (mb_w - 1) * 16);
    const int ry = clampi(y * 16 + (mv[2 * g + 1] >> 2), 0, (mb_h - 1) * 16);
    const int cx0 = rx / 16, cy0 = ry / 16, fx = rx % 16, fy = ry % 16;
    const int wys[2] = {16 - fy, fy}, wxs[2] = {16 - fx, fx};
    unsigned long long *prev = acc + (size_t)(f - (refs ? refs[g] : 1)) * nm;
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
        if (wys[dy] <= 0) continue;
        const int cy = min(cy0 + dy, mb_h - 1);
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            if (wxs[dx] <= 0) continue;
            const int cx = min(cx0 + dx, mb_w - 1);
            atomicAdd(&prev[cy * mb_w + cx], (unsigned long long)((amount * wxs[dx] * wys[dy]) >> 8));
        }
    }
}

// The network's proposals: 8 per block, bins of a 17 x 17 grid of 4-pel steps
constexpr int LS_NCAND = 8;
constexpr int LS_BINS = 17;

// ---------------------------------------------------------------------------
// Multi-reference lookahead (build-defined: the 4K configuration's 3 refs, +-64;
// the reference's inter_cost walks a seeded diamond against one reference).
// The half-res frames are copied into edge-replicated planes with a pad of
// range + 8 (rounded up to 4), so every search window and every winner's 8x8
// block lies inside a plane and the clamped reads of the definition are plain
// reads: the integer search is the full search's lane ring (fs_key) at zero rate,
// one workgroup row per (frame, reference); k_la_cost then prices each
// reference's winner with the 8x8 Hadamard and keeps the cheapest.
// ---------------------------------------------------------------------------

__global__ void k_la_pad(const uint8_t *__restrict__ half, int hw, int hh, int P, uint8_t *__restrict__ planes,
                         size_t pstride)
{
    const int pw = hw + 2 * P, ph = hh + 2 * P;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x; // dword of the padded plane
    if (i >= (long)pw * ph / 4) return;
    const int y = (int)(i / (pw / 4)), x = (int)(i - (long)y * (pw / 4)) * 4;
    const uint8_t *row = half + (size_t)blockIdx.y * hw * hh + (size_t)clampi(y - P, 0, hh - 1) * hw;
    uint32_t v = 0;
#pragma unroll
    for (int m = 0; m < 4; m++) v |= (uint32_t)row[clampi(x + m - P, 0, hw - 1)] << (8 * m);
    *(uint32_t *)(planes + (size_t)blockIdx.y * pstride + (size_t)y * pw + x) = v;
}

template <int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RH264_FS_WAVES)))
void k_la_search(const uint8_t *__restrict__ half, int mb_w, int mb_h, int nrefs, int R, int P,
                 const uint8_t *__restrict__ planes, size_t pstride, uint64_t *__restrict__ keys)
{
    __shared__ int32_t s_lr[RH264_LR_N];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = 0; // zero rate: the raster-first SAD minimum
    __syncthreads();
    const int f = 1 + (int)blockIdx.y / nrefs, k = 1 + (int)blockIdx.y % nrefs;
    if (k > f) return; // no such reference (the whole workgroup)
    const int hw = mb_w * 8, hh = mb_h * 8, nblk = mb_w * mb_h;
    const int NG = fs_groups<V>(2 * R + 1);
    const long gi = (long)blockIdx.x * 256 + threadIdx.x;
    const int b = (int)(gi / NG);
    const bool active = b < nblk;
    const int bs = active ? b : nblk - 1;
    FsJob J;
    J.cw = hw;
    J.pw = hw + 2 * P;
    J.ph = hh + 2 * P;
    J.pad = P;
    J.R = R;
    J.g = (int)(gi - (long)b * NG);
    J.lx = (bs % mb_w) * 8;
    J.ly = (bs / mb_w) * 8;
    J.cx = J.cy = J.mpx = J.mpy = 0;
    J.F = planes + (size_t)(f - k) * pstride;
    J.srcp = half + (size_t)f * hw * hh + (size_t)J.ly * hw + J.lx;
    const uint64_t key = fs_key<8, 8, V>(J, s_lr);
    fs_commit(key, b, active, keys + ((size_t)(f - 1) * nrefs + (k - 1)) * nblk + bs);
}

// One thread per (frame, block): intra = max(1, AC sum|H8 x H8| >> 1), and for
// frames > 0 the cheapest reference's 8x8 SATD (ties to the nearest reference).
// The winners come from the exhaustive search's keys, or (learned-seeded form) as
// vectors wmv[((f-1)*nrefs + k-1)*nm + m] with fall-back flags wfb at the same index.
__global__ __launch_bounds__(256) void k_la_cost(const uint8_t *__restrict__ half, int mb_w, int mb_h, int nframes,
                                                 int nrefs, int R, int P, const uint8_t *__restrict__ planes,
                                                 size_t pstride, const uint64_t *__restrict__ keys,
                                                 const int32_t *__restrict__ wmv, const int32_t *__restrict__ wfb,
                                                 int32_t *__restrict__ intra, int32_t *__restrict__ inter,
                                                 int32_t *__restrict__ mv, int32_t *__restrict__ ref,
                                                 int32_t *__restrict__ fbmask)
{
    const int nm = mb_w * mb_h;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nframes * nm) return;
    const int f = (int)(i / nm), m = (int)(i - (long)f * nm);
    const int hw = mb_w * 8, hh = mb_h * 8, pw = hw + 2 * P;
    const int bx0 = (m % mb_w) * 8, by0 = (m / mb_w) * 8;
    const uint8_t *cur = half + (size_t)f * hw * hh + (size_t)by0 * hw + bx0;
    uint32_t src[16]; // the block's rows as packed dwords
#pragma unroll
    for (int j = 0; j < 8; j++) {
        src[2 * j] = *(const uint32_t *)(cur + (size_t)j * hw);
        src[2 * j + 1] = *(const uint32_t *)(cur + (size_t)j * hw + 4);
    }
    int d[64];
    int dc = 0;
#pragma unroll
    for (int k = 0; k < 64; k++) {
        d[k] = (src[k >> 2] >> (8 * (k & 3))) & 255;
        dc += d[k];
    }
    // the DC coefficient of H8 x H8 is the block sum; the AC part is the rest
    const int ic = max(1, (satd8x8_sum_abs(d) - dc) >> 1);
    int best = 0x7fffffff, bdx = 0, bdy = 0, bk = 0, fbm = 0;
    const int NC = 2 * R + 1;
    for (int k = 1; k <= nrefs && k <= f; k++) {
        const size_t w = ((size_t)(f - 1) * nrefs + (k - 1)) * nm + m;
        int dx, dy;
        if (wmv) {
            dx = wmv[2 * w];
            dy = wmv[2 * w + 1];
            fbm |= wfb[w] << (k - 1);
        } else {
            const uint32_t idx = (uint32_t)keys[w];
            dy = (int)(idx / (uint32_t)NC) - R;
            dx = (int)(idx % (uint32_t)NC) - R;
        }
        const uint8_t *pl = planes + (size_t)(f - k) * pstride + (size_t)(P + by0 + dy) * pw + P + bx0 + dx;
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
            for (int b = 0; b < 8; b++)
                d[j * 8 + b] = (int)((src[(j * 8 + b) >> 2] >> (8 * (b & 3))) & 255) - (int)pl[(size_t)j * pw + b];
        const int c = satd8x8_sum_abs(d) >> 1;
        if (c < best) {
            best = c;
            bdx = dx;
            bdy = dy;
            bk = k;
        }
    }
    intra[i] = ic;
    inter[i] = f == 0 ? ic : min(best, ic);
    mv[2 * i] = 8 * bdx;
    mv[2 * i + 1] = 8 * bdy;
    ref[i] = bk;
    if (fbmask) fbmask[i] = fbm;
}

// Learned-seeded lookahead: one lane per (block, proposal) searches the proposal's
// +-RL window of one reference exhaustively (8x8 SADs: a quad-SAD group for columns
// 0..3, v_sad_u8 for the rest, as k_cand_search_m); the block's winner is the first
// minimum in proposal order, then raster order.  A block whose best SAD exceeds
// fb_sad is appended to the fall-back list (entry = its (frame, reference, block)
// index) and its argmin slot cleared.  grid.y = (frame - 1) * nrefs + reference - 1.
template <int RL>
__global__ __launch_bounds__(256) void k_la_cand(const uint8_t *__restrict__ half, int mb_w, int mb_h, int nframes,
                                                 int nrefs, int P, const uint8_t *__restrict__ planes, size_t pstride,
                                                 const int32_t *__restrict__ top8, int scaled, int cmax, int fb_sad,
                                                 int32_t *__restrict__ wmv, int32_t *__restrict__ wfb,
                                                 unsigned *__restrict__ count, int32_t *__restrict__ list,
                                                 uint64_t *__restrict__ lkeys)
{
    constexpr int NC = 2 * RL + 1, NQ = NC / 4, REM = NC - 4 * NQ, BPW = 256 / LS_NCAND, NJ = 8 + 2 * RL;
    constexpr int NW = (NC + 7 + 3) / 4; // realigned window dwords per row (bytes 0 .. NC + 6)
    __shared__ unsigned long long s_key[BPW];
    if (threadIdx.x < BPW) s_key[threadIdx.x] = ~0ull;
    __syncthreads();
    unsigned wgx, wgy;
    xcd_grid2(wgx, wgy);
    const int f = 1 + (int)wgy / nrefs, k = 1 + (int)wgy % nrefs;
    if (k > f) return; // no such reference (the whole workgroup)
    const int hw = mb_w * 8, hh = mb_h * 8, nm = mb_w * mb_h, pw = hw + 2 * P;
    const int slot = threadIdx.x / LS_NCAND, c = threadIdx.x - slot * LS_NCAND;
    const int m = (int)wgx * BPW + slot;
    // proposals of (frame f, reference f - k): the network's run over the pairs (j + k, j),
    // or (scaled) its run over the pairs (j + 1, j) for frame f, times the distance k
    const int32_t *t8 = scaled ? top8 + ((size_t)(f - 1) * nm) * LS_NCAND
                               : top8 + (((size_t)(k - 1) * nframes + (f - k)) * nm) * LS_NCAND;
    if (m < nm) {
        const int bx0 = (m % mb_w) * 8, by0 = (m / mb_w) * 8;
        const int bin = t8[(size_t)m * LS_NCAND + c];
        // bin -> full-resolution vector ((bin % 17) - 8, (bin / 17) - 8) * 4 pels = *2 half-res pels
        const int sk = scaled ? k : 1;
        const int cx = clampi(((bin % LS_BINS) - LS_BINS / 2) * 2 * sk, -cmax, cmax);
        const int cy = clampi(((bin / LS_BINS) - LS_BINS / 2) * 2 * sk, -cmax, cmax);
        const uint8_t *src = half + (size_t)f * hw * hh + (size_t)by0 * hw + bx0;
        uint32_t sr[8][2];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            sr[r][0] = *(const uint32_t *)(src + (size_t)r * hw);
            sr[r][1] = *(const uint32_t *)(src + (size_t)r * hw + 4);
        }
        const uint8_t *F = planes + (size_t)(f - k) * pstride;
        // |cx|, |cy| <= cmax = P - 8 - RL: the window lies inside the padded plane
        const int x = P + bx0 + cx - RL, y0 = P + by0 + cy - RL;
        const int sh = x & 3;
        uint64_t aq[NQ > 0 ? NQ : 1][NC];
        uint32_t ar[REM > 0 ? REM : 1][NC];
#pragma unroll
        for (int t = 0; t < NC; t++) {
#pragma unroll
            for (int g = 0; g < NQ; g++) aq[g][t] = 0;
#pragma unroll
            for (int q = 0; q < REM; q++) ar[q][t] = 0;
        }
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const uint32_t *rp = (const uint32_t *)(F + (size_t)(y0 + j) * pw + (x & ~3));
            uint32_t a[NW + 1], w[NW];
#pragma unroll
            for (int i = 0; i < NW + 1; i++) a[i] = rp[i];
#pragma unroll
            for (int i = 0; i < NW; i++) w[i] = __builtin_amdgcn_alignbyte(a[i + 1], a[i], (uint32_t)sh);
#pragma unroll
            for (int t = 0; t < NC; t++) {
                const int r = j - t;
                if (r < 0 || r >= 8) continue;
#pragma unroll
                for (int g = 0; g < NQ; g++)
#pragma unroll
                    for (int i = 0; i < 2; i++)
                        aq[g][t] = __builtin_amdgcn_qsad_pk_u16_u8(((uint64_t)w[g + i + 1] << 32) | w[g + i], sr[r][i],
                                                                   aq[g][t]);
#pragma unroll
                for (int q = 0; q < REM; q++) {
                    const int col = 4 * NQ + q;
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        const uint32_t ref4 = (col & 3) == 0 ? w[(col >> 2) + i]
                                                             : __builtin_amdgcn_alignbyte(w[(col >> 2) + i + 1],
                                                                                          w[(col >> 2) + i],
                                                                                          (uint32_t)(col & 3));
                        ar[q][t] = __builtin_amdgcn_sad_u8(sr[r][i], ref4, ar[q][t]);
                    }
                }
            }
        }
        uint64_t best = ~0ull;
#pragma unroll
        for (int t = 0; t < NC; t++)
#pragma unroll
            for (int col = 0; col < NC; col++) {
                const uint32_t sad = col < 4 * NQ ? (uint32_t)(aq[col >> 2][t] >> (16 * (col & 3))) & 0xffffu
                                                  : ar[col - 4 * NQ][t];
                const uint64_t key = ((uint64_t)sad << 32) | ((uint64_t)c << 8) | (uint64_t)(t * NC + col);
                best = key < best ? key : best;
            }
        atomicMin(&s_key[slot], (unsigned long long)best);
    }
    __syncthreads();
    if (threadIdx.x < BPW) {
        const int mm = (int)wgx * BPW + threadIdx.x;
        if (mm < nm) {
            const uint64_t key = s_key[threadIdx.x];
            const int kk = (int)((key >> 8) & 255), idx = (int)(key & 255);
            const int bin = t8[(size_t)mm * LS_NCAND + kk];
            const size_t w = ((size_t)(f - 1) * nrefs + (k - 1)) * nm + mm;
            const int sk = scaled ? k : 1; // the winning proposal's centre, as the search placed it
            wmv[2 * w] = clampi(((bin % LS_BINS) - LS_BINS / 2) * 2 * sk, -cmax, cmax) + idx % NC - RL;
            wmv[2 * w + 1] = clampi(((bin / LS_BINS) - LS_BINS / 2) * 2 * sk, -cmax, cmax) + idx / NC - RL;
            const bool fb = (int64_t)(key >> 32) > (int64_t)fb_sad;
            wfb[w] = fb ? 1 : 0;
            if (fb) {
                const unsigned e = atomicAdd(count, 1u);
                list[e] = (int32_t)w;
                lkeys[e] = ~0ull;
            }
        }
    }
}

// Exhaustive +-R search of the fall-back entries (w = ((f-1)*nrefs + k-1)*nm + m): the
// full search's lane ring at zero rate; the grid strides over the list, whose length
// only the device knows.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RH264_FS_WAVES)))
void k_la_search_list(const uint8_t *__restrict__ half, int mb_w, int mb_h, int nrefs, int R, int P,
                      const uint8_t *__restrict__ planes, size_t pstride, const int32_t *__restrict__ list,
                      const unsigned *__restrict__ count, uint64_t *__restrict__ lkeys)
{
    __shared__ int32_t s_lr[RH264_LR_N];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = 0;
    __syncthreads();
    const int hw = mb_w * 8, hh = mb_h * 8, nm = mb_w * mb_h;
    const int NG = fs_groups<0>(2 * R + 1);
    const long n = (long)*count;
    for (long chunk = blockIdx.x; chunk * 256 < n * NG; chunk += gridDim.x) {
        const long gi = chunk * 256 + threadIdx.x;
        const long e = gi / NG;
        const bool active = e < n;
        const long w = list[active ? e : n - 1];
        const int m = (int)(w % nm), fk = (int)(w / nm), f = 1 + fk / nrefs, k = 1 + fk % nrefs;
        FsJob J;
        J.cw = hw;
        J.pw = hw + 2 * P;
        J.ph = hh + 2 * P;
        J.pad = P;
        J.R = R;
        J.g = (int)(gi - e * NG);
        J.lx = (m % mb_w) * 8;
        J.ly = (m / mb_w) * 8;
        J.cx = J.cy = J.mpx = J.mpy = 0;
        J.F = planes + (size_t)(f - k) * pstride;
        J.srcp = half + (size_t)f * hw * hh + (size_t)J.ly * hw + J.lx;
        const uint64_t key = fs_key<8, 8, 0>(J, s_lr);
        fs_commit(key, e, active, lkeys + (active ? e : n - 1));
    }
}

// The fall-back entries take the exhaustive search's winner.
__global__ void k_la_fb_merge(const int32_t *__restrict__ list, const unsigned *__restrict__ count,
                              const uint64_t *__restrict__ lkeys, int R, int32_t *__restrict__ wmv)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)*count) return;
    const long w = list[e];
    const uint32_t idx = (uint32_t)lkeys[e];
    const int NC = 2 * R + 1;
    wmv[2 * w] = (int)(idx % (uint32_t)NC) - R;
    wmv[2 * w + 1] = (int)(idx / (uint32_t)NC) - R;
}

// ---------------------------------------------------------------------------
// learned-proposal search (the network lives in rh264_mvnet.hip)
// ---------------------------------------------------------------------------


// Candidate centres from the network's top bins: bin c -> ((c%17)-8, (c/17)-8) * 4
// pels, as quarter-pel MVs.  centres: [k][block][2].
__global__ void k_cand_centres(const int32_t *__restrict__ top8, long nb, int32_t *__restrict__ centres)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * LS_NCAND) return;
    const long b = i / LS_NCAND;
    const int k = (int)(i - b * LS_NCAND);
    const int c = top8[i];
    centres[2 * ((size_t)k * nb + b)] = ((c % LS_BINS) - LS_BINS / 2) * 16;
    centres[2 * ((size_t)k * nb + b) + 1] = ((c / LS_BINS) - LS_BINS / 2) * 16;
}

// The exact costs of the candidates' local searches decide the start vector (first
// minimum in candidate order); blocks whose best cost exceeds fb_cost join the
// fall-back list for a full search.
__global__ void k_cand_select(const uint64_t *__restrict__ keys8, const int32_t *__restrict__ centres, long nb, int Rl,
                              int32_t fb_cost, int32_t *__restrict__ start, int32_t *__restrict__ flag,
                              unsigned *__restrict__ count, int32_t *__restrict__ list)
{
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    uint64_t best = ~0ull;
    int bk = 0;
    for (int k = 0; k < LS_NCAND; k++) {
        const uint64_t key = keys8[(size_t)k * nb + b];
        if ((key >> 32) < (best >> 32)) {
            best = key;
            bk = k;
        }
    }
    const int NC = 2 * Rl + 1;
    const uint32_t idx = (uint32_t)best;
    const int32_t *c = centres + 2 * ((size_t)bk * nb + b);
    start[2 * b] = c[0] + 4 * ((int)(idx % (uint32_t)NC) - Rl);
    start[2 * b + 1] = c[1] + 4 * ((int)(idx / (uint32_t)NC) - Rl);
    const bool fb = (int64_t)(best >> 32) > (int64_t)fb_cost;
    flag[b] = fb ? 1 : 0;
    if (fb) list[atomicAdd(count, 1u)] = (int32_t)b;
}

// All candidates' local searches in one pass.  Lane (block, candidate k, window
// column dx) keeps the block's 16 source rows in registers, realigns each of the
// 16 + 2RL reference rows of its column once and adds it into the SADs of the
// 2RL + 1 window rows that contain it.  A block's key is the minimum over its
// lanes of (cost << 32 | k << 8 | raster index): the lowest cost, then the
// earliest candidate, then the first raster position of that candidate's window
// -- the order of the per-candidate searches followed by k_cand_select.
template <int RL, bool FAST>
__device__ __forceinline__ void cand_rows(const uint32_t (&s)[16][4], const uint8_t *__restrict__ F, int pw, int ph,
                                          int x, int y0, uint32_t (&acc)[2 * RL + 1])
{
    constexpr int NP = 2 * RL + 1, NJ = 16 + 2 * RL;
    const int sh = x & 3;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        uint32_t w[4];
        if (FAST) {
            const uint32_t *rp = (const uint32_t *)(F + (size_t)(y0 + j) * pw + (x & ~3));
            uint32_t a[5];
#pragma unroll
            for (int i = 0; i < 5; i++) a[i] = rp[i];
#pragma unroll
            for (int i = 0; i < 4; i++) w[i] = __builtin_amdgcn_alignbyte(a[i + 1], a[i], (uint32_t)sh);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) w[i] = load4_clamped(F, pw, ph, x + 4 * i, y0 + j);
        }
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int r = j - t;
            if (r < 0 || r >= 16) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) acc[t] = __builtin_amdgcn_sad_u8(s[r][i], w[i], acc[t]);
        }
    }
}

template <int RL>
__global__ __launch_bounds__(256) void k_cand_search(rh264_search_params p, const uint8_t *__restrict__ cur,
                                                     size_t cur_stride, const uint8_t *__restrict__ planes,
                                                     size_t planes_stride, const int32_t *__restrict__ top8,
                                                     const int32_t *__restrict__ mvps, long nb, int32_t fb_cost,
                                                     int32_t *__restrict__ start, int32_t *__restrict__ flag,
                                                     unsigned *__restrict__ count, int32_t *__restrict__ list)
{
    constexpr int NP = 2 * RL + 1, LPB = LS_NCAND * NP, BPW = 256 / LPB, NJ = 16 + 2 * RL;
    __shared__ int32_t s_lr[RH264_LR_N];
    __shared__ unsigned long long s_key[BPW];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr_fp[threadIdx.x];
    if (threadIdx.x < BPW) s_key[threadIdx.x] = ~0ull;
    __syncthreads();
    const int slot = threadIdx.x / LPB, li = threadIdx.x - slot * LPB;
    const long gb0 = (long)blockIdx.x * BPW + slot;
    const bool active = slot < BPW && gb0 < nb;
    if (active) {
        const long gb = gb0;
        const int k = li / NP, dxi = li - k * NP;
        const int nbx = p.cw / 16, nblk = nbx * (p.ch / 16);
        const int frame = (int)(gb / nblk), b = (int)(gb - (long)frame * nblk);
        const int lx = (b % nbx) * 16, ly = (b / nbx) * 16;
        const int c = top8[gb * LS_NCAND + k];
        const int cx = ((c % LS_BINS) - LS_BINS / 2) * 16, cy = ((c / LS_BINS) - LS_BINS / 2) * 16;
        const int mpx = mvps ? mvps[2 * gb] : 0, mpy = mvps ? mvps[2 * gb + 1] : 0;
        const int pw = p.cw + 2 * p.pad, ph = p.ch + 2 * p.pad;
        const uint8_t *F = planes + (size_t)frame * planes_stride;
        const uint8_t *src = cur + (size_t)frame * cur_stride + (size_t)ly * p.cw + lx;
        uint32_t s[16][4];
#pragma unroll
        for (int r = 0; r < 16; r++)
#pragma unroll
            for (int i = 0; i < 4; i++) s[r][i] = *(const uint32_t *)(src + (size_t)r * p.cw + 4 * i);
        const int x = lx + (cx >> 2) + p.pad + dxi - RL; // padded column of the lane's candidates
        const int y0 = ly + (cy >> 2) + p.pad - RL;      // padded row of the window's first row
        uint32_t acc[NP];
#pragma unroll
        for (int t = 0; t < NP; t++) acc[t] = 0;
        if (x >= 0 && (x & ~3) + 20 <= pw && y0 >= 0 && y0 + NJ <= ph)
            cand_rows<RL, true>(s, F, pw, ph, x, y0, acc);
        else
            cand_rows<RL, false>(s, F, pw, ph, x, y0, acc);
        const int rx = mvbits(cx + 4 * (dxi - RL) - mpx);
        uint64_t best = ~0ull;
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const uint32_t cost = acc[t] + (uint32_t)s_lr[rx + mvbits(cy + 4 * (t - RL) - mpy)];
            const uint64_t key = ((uint64_t)cost << 32) | ((uint64_t)k << 8) | (uint64_t)(t * NP + dxi);
            best = key < best ? key : best;
        }
        atomicMin(&s_key[slot], (unsigned long long)best);
    }
    __syncthreads();
    if (threadIdx.x < BPW) {
        const long g = (long)blockIdx.x * BPW + threadIdx.x;
        if (g < nb) {
            const uint64_t key = s_key[threadIdx.x];
            const int kk = (int)((key >> 8) & 255), idx = (int)(key & 255);
            const int ck = top8[g * LS_NCAND + kk];
            start[2 * g] = ((ck % LS_BINS) - LS_BINS / 2) * 16 + 4 * (idx % NP - RL);
            start[2 * g + 1] = ((ck / LS_BINS) - LS_BINS / 2) * 16 + 4 * (idx / NP - RL);
            const bool fb = (int64_t)(key >> 32) > (int64_t)fb_cost;
            flag[g] = fb ? 1 : 0;
            if (fb) list[atomicAdd(count, 1u)] = (int32_t)g;
        }
    }
}

// The same search with quad-SADs: lane (block slot, candidate k, column group g)
// owns window columns 4g..4g+3 (four packed u16 SADs per window row, one
// v_qsad_pk_u16_u8 per source dword and reference row); the blocks of a workgroup
// stage their sources in LDS.  Keys and tie-breaks are those of k_cand_search.
template <int RL, bool FAST>
__device__ __forceinline__ void cand_rows_q(const uint32_t (&s)[16][4], const uint8_t *__restrict__ F, int pw, int ph,
                                            int x, int y0, uint64_t (&acc)[2 * RL + 1])
{
    constexpr int NP = 2 * RL + 1, NJ = 16 + 2 * RL;
    const int sh = x & 3;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        uint32_t w[5];
        if (FAST) {
            const uint32_t *rp = (const uint32_t *)(F + (size_t)(y0 + j) * pw + (x & ~3));
            uint32_t a[6];
#pragma unroll
            for (int i = 0; i < 6; i++) a[i] = rp[i];
#pragma unroll
            for (int i = 0; i < 5; i++) w[i] = __builtin_amdgcn_alignbyte(a[i + 1], a[i], (uint32_t)sh);
        } else {
#pragma unroll
            for (int i = 0; i < 5; i++) w[i] = load4_clamped(F, pw, ph, x + 4 * i, y0 + j);
        }
#pragma unroll
        for (int t = 0; t < NP; t++) {
            const int r = j - t;
            if (r < 0 || r >= 16) continue;
#pragma unroll
            for (int i = 0; i < 4; i++)
                acc[t] = __builtin_amdgcn_qsad_pk_u16_u8(((uint64_t)w[i + 1] << 32) | w[i], s[r][i], acc[t]);
        }
    }
}

template <int RL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_cand_search_q(rh264_search_params p, const uint8_t *__restrict__ cur, size_t cur_stride,
                     const uint8_t *__restrict__ planes, size_t planes_stride, const int32_t *__restrict__ top8,
                     const int32_t *__restrict__ mvps, long nb, int32_t fb_cost, int32_t *__restrict__ start,
                     int32_t *__restrict__ flag, unsigned *__restrict__ count, int32_t *__restrict__ list)
{
    constexpr int NC = 2 * RL + 1, NG = (NC + 3) / 4, LPB = LS_NCAND * NG, BPW = 256 / LPB;
    static_assert(BPW * LPB == 256 && BPW % 4 == 0, "whole blocks per workgroup");
    __shared__ int32_t s_lr[RH264_LR_N];
    __shared__ unsigned long long s_key[BPW];
    __shared__ uint32_t s_src[BPW][64];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr_fp[threadIdx.x];
    if (threadIdx.x < BPW) s_key[threadIdx.x] = ~0ull;
    const unsigned wg = xcd_linear(blockIdx.x, gridDim.x);
    const int nbx = p.cw / 16, nblk = nbx * (p.ch / 16);
    // stage the sources: 64 threads per block, one dword each (row t/4, dword t%4)
#pragma unroll
    for (int s0 = 0; s0 < BPW; s0 += 4) {
        const int sl = s0 + (threadIdx.x >> 6), t = threadIdx.x & 63;
        const long g = (long)wg * BPW + sl;
        if (g < nb) {
            const int frame = (int)(g / nblk), b = (int)(g - (long)frame * nblk);
            const uint8_t *src = cur + (size_t)frame * cur_stride + (size_t)((b / nbx) * 16 + (t >> 2)) * p.cw +
                                 (b % nbx) * 16 + 4 * (t & 3);
            s_src[sl][t] = *(const uint32_t *)src;
        }
    }
    __syncthreads();
    const int slot = threadIdx.x / LPB, li = threadIdx.x - slot * LPB;
    const long gb = (long)wg * BPW + slot;
    if (gb < nb) {
        const int k = li / NG, g = li - k * NG;
        const int frame = (int)(gb / nblk), b = (int)(gb - (long)frame * nblk);
        const int lx = (b % nbx) * 16, ly = (b / nbx) * 16;
        const int c = top8[gb * LS_NCAND + k];
        const int cx = ((c % LS_BINS) - LS_BINS / 2) * 16, cy = ((c / LS_BINS) - LS_BINS / 2) * 16;
        const int mpx = mvps ? mvps[2 * gb] : 0, mpy = mvps ? mvps[2 * gb + 1] : 0;
        const int pw = p.cw + 2 * p.pad, ph = p.ch + 2 * p.pad;
        const uint8_t *F = planes + (size_t)frame * planes_stride;
        uint32_t sr[16][4];
#pragma unroll
        for (int r = 0; r < 16; r++)
#pragma unroll
            for (int i = 0; i < 4; i++) sr[r][i] = s_src[slot][4 * r + i];
        const int x = lx + (cx >> 2) + p.pad + 4 * g - RL; // padded column of the lane's first candidate
        const int y0 = ly + (cy >> 2) + p.pad - RL;        // padded row of the window's first row
        uint64_t acc[NC];
#pragma unroll
        for (int t = 0; t < NC; t++) acc[t] = 0;
        if (x >= 0 && (x & ~3) + 24 <= pw && y0 >= 0 && y0 + 16 + 2 * RL <= ph)
            cand_rows_q<RL, true>(sr, F, pw, ph, x, y0, acc);
        else
            cand_rows_q<RL, false>(sr, F, pw, ph, x, y0, acc);
        int rxb[4];
#pragma unroll
        for (int q = 0; q < 4; q++) rxb[q] = mvbits(cx + 4 * (4 * g + q - RL) - mpx);
        uint64_t best = ~0ull;
#pragma unroll
        for (int t = 0; t < NC; t++) {
            const int ryb = mvbits(cy + 4 * (t - RL) - mpy);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int col = 4 * g + q;
                const uint32_t cost = (uint32_t)(acc[t] >> (16 * q)) & 0xffffu;
                const uint64_t key = ((uint64_t)(cost + (uint32_t)s_lr[rxb[q] + ryb]) << 32) | ((uint64_t)k << 8) |
                                     (uint64_t)(t * NC + col);
                if (col < NC && key < best) best = key;
            }
        }
        atomicMin(&s_key[slot], (unsigned long long)best);
    }
    __syncthreads();
    if (threadIdx.x < BPW) {
        const long g = (long)wg * BPW + threadIdx.x;
        if (g < nb) {
            const uint64_t key = s_key[threadIdx.x];
            const int kk = (int)((key >> 8) & 255), idx = (int)(key & 255);
            const int ck = top8[g * LS_NCAND + kk];
            start[2 * g] = ((ck % LS_BINS) - LS_BINS / 2) * 16 + 4 * (idx % NC - RL);
            start[2 * g + 1] = ((ck / LS_BINS) - LS_BINS / 2) * 16 + 4 * (idx / NC - RL);
            const bool fb = (int64_t)(key >> 32) > (int64_t)fb_cost;
            flag[g] = fb ? 1 : 0;
            if (fb) list[atomicAdd(count, 1u)] = (int32_t)g;
        }
    }
}

template <int RL, bool FAST, typename SRC>
__device__ __forceinline__ void cand_rows_m(const SRC &src, const uint8_t *__restrict__ F, int pw, int ph,
                                            int x, int y0, uint64_t (&aq)[(2 * RL + 1) / 4 > 0 ? (2 * RL + 1) / 4 : 1][2 * RL + 1],
                                            uint32_t (&ar)[(2 * RL + 1) % 4 > 0 ? (2 * RL + 1) % 4 : 1][2 * RL + 1])
{
    constexpr int NC = 2 * RL + 1, NQ = NC / 4, REM = NC - 4 * NQ, NJ = 16 + 2 * RL;
    constexpr int NW = (NC + 15 + 3) / 4; // realigned window dwords per row (bytes 0 .. NC + 14)
    const int sh = x & 3;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        uint32_t w[NW];
        if (FAST) {
            const uint32_t *rp = (const uint32_t *)(F + (size_t)(y0 + j) * pw + (x & ~3));
            uint32_t a[NW + 1];
#pragma unroll
            for (int i = 0; i < NW + 1; i++) a[i] = rp[i];
#pragma unroll
            for (int i = 0; i < NW; i++) w[i] = __builtin_amdgcn_alignbyte(a[i + 1], a[i], (uint32_t)sh);
        } else {
#pragma unroll
            for (int i = 0; i < NW; i++) w[i] = load4_clamped(F, pw, ph, x + 4 * i, y0 + j);
        }
#pragma unroll
        for (int t = 0; t < NC; t++) {
            const int r = j - t;
            if (r < 0 || r >= 16) continue;
            const uint4 sv = src(r); // source row r
            const uint32_t s[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
            for (int g = 0; g < NQ; g++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    aq[g][t] = __builtin_amdgcn_qsad_pk_u16_u8(((uint64_t)w[g + i + 1] << 32) | w[g + i], s[i],
                                                               aq[g][t]);
#pragma unroll
            for (int q = 0; q < REM; q++) {
                const int col = 4 * NQ + q; // window bytes col + 4i .. col + 4i + 3
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t ref4 = (col & 3) == 0 ? w[(col >> 2) + i]
                                                         : __builtin_amdgcn_alignbyte(w[(col >> 2) + i + 1],
                                                                                      w[(col >> 2) + i],
                                                                                      (uint32_t)(col & 3));
                    ar[q][t] = __builtin_amdgcn_sad_u8(s[i], ref4, ar[q][t]);
                }
            }
        }
    }
}

// The same search with one lane per (block, candidate): the lane reads each of its
// window's 16 + 2RL reference rows once (one realigned span of 16 + 2RL bytes),
// prices the window's columns in groups of four with quad-SADs and the 2RL + 1 - 4NQ
// remaining columns with v_sad_u8 (no column beyond the window is computed), so a
// block costs half the loads of k_cand_search_q and no wasted quad-SAD columns.
// Keys and tie-breaks are those of k_cand_search.
template <int RL, bool SL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SL ? 4 : (RL <= 2 ? 3 : 2))))
void k_cand_search_m(rh264_search_params p, const uint8_t *__restrict__ cur, size_t cur_stride,
                     const uint8_t *__restrict__ planes, size_t planes_stride, const int32_t *__restrict__ top8,
                     const int32_t *__restrict__ mvps, long nb, int32_t fb_cost, int32_t *__restrict__ start,
                     int32_t *__restrict__ flag, unsigned *__restrict__ count, int32_t *__restrict__ list,
                     uint64_t *__restrict__ fkeys)
{
    constexpr int NC = 2 * RL + 1, NQ = NC / 4, REM = NC - 4 * NQ, BPW = 256 / LS_NCAND, NJ = 16 + 2 * RL;
    constexpr int NW = (NC + 15 + 3) / 4; // realigned window dwords per row (bytes 0 .. NC + 14)
    __shared__ int32_t s_lr[RH264_LR_N];
    __shared__ unsigned long long s_key[BPW];
    // 68 dwords a block: the 8 blocks one wave reads start 4 banks apart (a pitch of 64
    // would put them all on the same banks)
    __shared__ uint32_t s_src[BPW][68];
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr_fp[threadIdx.x];
    if (threadIdx.x < BPW) s_key[threadIdx.x] = ~0ull;
    const unsigned wg = xcd_linear(blockIdx.x, gridDim.x);
    const int nbx = p.cw / 16, nblk = nbx * (p.ch / 16);
    // stage the sources: 64 threads per block, one dword each (row t/4, dword t%4)
#pragma unroll
    for (int s0 = 0; s0 < BPW; s0 += 4) {
        const int sl = s0 + (threadIdx.x >> 6), t = threadIdx.x & 63;
        const long g = (long)wg * BPW + sl;
        if (g < nb) {
            const int frame = (int)(g / nblk), b = (int)(g - (long)frame * nblk);
            const uint8_t *src = cur + (size_t)frame * cur_stride + (size_t)((b / nbx) * 16 + (t >> 2)) * p.cw +
                                 (b % nbx) * 16 + 4 * (t & 3);
            s_src[sl][t] = *(const uint32_t *)src;
        }
    }
    __syncthreads();
    const int slot = threadIdx.x / LS_NCAND, k = threadIdx.x - slot * LS_NCAND;
    const long gb = (long)wg * BPW + slot;
    if (gb < nb) {
        const int frame = (int)(gb / nblk), b = (int)(gb - (long)frame * nblk);
        const int lx = (b % nbx) * 16, ly = (b / nbx) * 16;
        const int c = top8[gb * LS_NCAND + k];
        const int cx = ((c % LS_BINS) - LS_BINS / 2) * 16, cy = ((c / LS_BINS) - LS_BINS / 2) * 16;
        const int mpx = mvps ? mvps[2 * gb] : 0, mpy = mvps ? mvps[2 * gb + 1] : 0;
        const int pw = p.cw + 2 * p.pad, ph = p.ch + 2 * p.pad;
        const uint8_t *F = planes + (size_t)frame * planes_stride;
        // the source rows: read from LDS where they are used (SL), else held in registers
        uint32_t sr[SL ? 1 : 16][4];
        if (!SL) {
#pragma unroll
            for (int r = 0; r < 16; r++)
#pragma unroll
                for (int i = 0; i < 4; i++) sr[SL ? 0 : r][i] = s_src[slot][4 * r + i];
        }
        const uint32_t *lsrc = s_src[slot];
        const auto src = [&](int r) -> uint4 {
            if (SL) return *(const uint4 *)(lsrc + 4 * r);
            return make_uint4(sr[SL ? 0 : r][0], sr[SL ? 0 : r][1], sr[SL ? 0 : r][2], sr[SL ? 0 : r][3]);
        };
        const int x = lx + (cx >> 2) + p.pad - RL; // padded column of the window's first candidate
        const int y0 = ly + (cy >> 2) + p.pad - RL; // padded row of the window's first row
        const bool fast = x >= 0 && (x & ~3) + 4 * (NW + 1) <= pw && y0 >= 0 && y0 + NJ <= ph;
        const int sh = x & 3;
        uint64_t aq[NQ > 0 ? NQ : 1][NC];
        uint32_t ar[REM > 0 ? REM : 1][NC];
#pragma unroll
        for (int t = 0; t < NC; t++) {
#pragma unroll
            for (int g = 0; g < NQ; g++) aq[g][t] = 0;
#pragma unroll
            for (int q = 0; q < REM; q++) ar[q][t] = 0;
        }
        if (fast)
            cand_rows_m<RL, true>(src, F, pw, ph, x, y0, aq, ar);
        else
            cand_rows_m<RL, false>(src, F, pw, ph, x, y0, aq, ar);
        uint64_t best = ~0ull;
#pragma unroll
        for (int t = 0; t < NC; t++) {
            const int ryb = mvbits(cy + 4 * (t - RL) - mpy);
#pragma unroll
            for (int col = 0; col < NC; col++) {
                const uint32_t sad = col < 4 * NQ ? (uint32_t)(aq[col >> 2][t] >> (16 * (col & 3))) & 0xffffu
                                                  : ar[col - 4 * NQ][t];
                const int rxb = mvbits(cx + 4 * (col - RL) - mpx);
                const uint64_t key = ((uint64_t)(sad + (uint32_t)s_lr[rxb + ryb]) << 32) | ((uint64_t)k << 8) |
                                     (uint64_t)(t * NC + col);
                best = key < best ? key : best;
            }
        }
        atomicMin(&s_key[slot], (unsigned long long)best);
    }
    __syncthreads();
    if (threadIdx.x < BPW) {
        const long g = (long)wg * BPW + threadIdx.x;
        if (g < nb) {
            const uint64_t key = s_key[threadIdx.x];
            const int kk = (int)((key >> 8) & 255), idx = (int)(key & 255);
            const int ck = top8[g * LS_NCAND + kk];
            start[2 * g] = ((ck % LS_BINS) - LS_BINS / 2) * 16 + 4 * (idx % NC - RL);
            start[2 * g + 1] = ((ck / LS_BINS) - LS_BINS / 2) * 16 + 4 * (idx / NC - RL);
            const bool fb = (int64_t)(key >> 32) > (int64_t)fb_cost;
            flag[g] = fb ? 1 : 0;
            if (fb) {
                fkeys[g] = ~0ull; // the fall-back search's argmin slot
                list[atomicAdd(count, 1u)] = (int32_t)g;
            }
        }
    }
}

// Full search of the listed blocks only (frame-major global block indices).
template <int BW, int BH, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RH264_FS_WAVES)))
void k_full_search_list(rh264_search_params p, const uint8_t *__restrict__ cur, size_t cur_stride,
                        const uint8_t *__restrict__ planes, size_t planes_stride, const int32_t *__restrict__ list,
                        const unsigned *__restrict__ count, const int32_t *__restrict__ mvps,
                        uint64_t *__restrict__ keys)
{
    __shared__ int32_t s_lr[RH264_LR_N];
    const int NG = fs_groups<V>(2 * p.range + 1);
    const long n = (long)*count;
    // the list's length is known only here: workgroups with no chunk leave at once, the
    // others take chunks blockIdx.x, + gridDim.x, ... (the grid is capped, not sized
    // for a list of every block)
    if ((long)blockIdx.x * 256 >= n * NG) return;
    if (threadIdx.x < RH264_LR_N) s_lr[threadIdx.x] = p.lr_fp[threadIdx.x];
    __syncthreads();
    const int nbx = p.cw / BW, nblk = nbx * (p.ch / BH);
    for (long chunk = blockIdx.x; chunk * 256 < n * NG; chunk += gridDim.x) {
        const long gi = chunk * 256 + threadIdx.x;
        const long e = gi / NG;
        const bool active = e < n;
        const long gb = list[active ? e : n - 1];
        const int frame = (int)(gb / nblk), b = (int)(gb - (long)frame * nblk);
        FsJob J;
        J.cw = p.cw;
        J.pw = p.cw + 2 * p.pad;
        J.ph = p.ch + 2 * p.pad;
        J.pad = p.pad;
        J.R = p.range;
        J.g = (int)(gi - e * NG);
        J.lx = (b % nbx) * BW;
        J.ly = (b / nbx) * BH;
        J.cx = 0;
        J.cy = 0;
        J.mpx = mvps ? mvps[2 * gb] : 0;
        J.mpy = mvps ? mvps[2 * gb + 1] : 0;
        J.F = planes + (size_t)frame * planes_stride;
        J.srcp = cur + (size_t)frame * cur_stride + (size_t)J.ly * p.cw + J.lx;
        const uint64_t key = fs_key<BW, BH, V>(J, s_lr);
        fs_commit(key, e, active, keys + gb);
    }
}

// Fall-back blocks take the full search's vector as their start.
__global__ void k_fallback_merge(const int32_t *__restrict__ list, const unsigned *__restrict__ count,
                                 const uint64_t *__restrict__ keys, int R, int32_t *__restrict__ start)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)*count) return;
    const long gb = list[e];
    const uint32_t idx = (uint32_t)keys[gb];
    const int NC = 2 * R + 1;
    start[2 * gb] = 4 * ((int)(idx % (uint32_t)NC) - R);
    start[2 * gb + 1] = 4 * ((int)(idx / (uint32_t)NC) - R);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char *rh264_last_error(void) { return g_err.c_str(); }

int rh264_init(int device, rh264_ctx **out)
{
    if (!out) return fail("rh264_init: null out");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail("rh264_init: device %d of %d", device, ndev);
    HIP_TRY(hipSetDevice(device));
    rh264_ctx *c = new rh264_ctx();
    c->device = device;
    c->keys = nullptr;
    c->keys_cap = 0;
    c->noalloc_keys = nullptr;
    c->tmp_mv = nullptr;
    c->tmp_mv_cap = 0;
    c->prof_cap = c->prof_n = 0;
    c->ls_buf = nullptr;
    c->ls_cap = 0;
    c->la_buf = nullptr;
    c->la_cap = 0;
    c->lal_buf = nullptr;
    c->lal_cap = 0;
    c->ms_buf = nullptr;
    c->ms_cap = 0;
    c->decide_mode = 2;
    c->la_scaled = 1;
    c->fs_kernel = 0;
    c->refine_lds = 1;
    c->refine_mfma = 1;
    c->cand_quad = 2;
    c->cand_lds = 1;
    c->fallback_v = 1;
    c->fallback_wgs = 256 * 8;
    c->fallback_overlap = 0;
    c->planes_rw = 1;
    c->mvnet_fused = 2;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail("hipStreamCreate: %s", hipGetErrorString(e));
    }
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail("hipStreamCreate: %s", hipGetErrorString(e));
    }
    for (int i = 0; i < 16; i++) HIP_TRY(hipEventCreate(&c->ev[i]));
    HIP_TRY(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
    // this device's copy of the no-allocation scratch (its address differs per device)
    void *sp = nullptr;
    e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) {
        (void)hipStreamDestroy(c->stream);
        (void)hipStreamDestroy(c->side);
        delete c;
        return fail("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    c->noalloc_keys = (uint64_t *)sp;
    *out = c;
    return 0;
}

int rh264_destroy(rh264_ctx *ctx)
{
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->side);
    if (ctx->keys) (void)hipFree(ctx->keys);
    if (ctx->tmp_mv) (void)hipFree(ctx->tmp_mv);
    if (ctx->ls_buf) (void)hipFree(ctx->ls_buf);
    if (ctx->la_buf) (void)hipFree(ctx->la_buf);
    if (ctx->lal_buf) (void)hipFree(ctx->lal_buf);
    if (ctx->ms_buf) (void)hipFree(ctx->ms_buf);
    for (int i = 0; i < 16; i++) (void)hipEventDestroy(ctx->ev[i]);
    for (hipEvent_t e : ctx->prof) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ctx->fork_ev);
    (void)hipEventDestroy(ctx->join_ev);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->side);
    delete ctx;
    return 0;
}

int rh264_device_name(rh264_ctx *ctx, char *buf, int buflen)
{
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
    snprintf(buf, (size_t)buflen, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return 0;
}

int rh264_alloc(rh264_ctx *ctx, size_t bytes, void **dptr)
{
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMalloc(dptr, bytes ? bytes : 1));
    return 0;
}

int rh264_free(rh264_ctx *ctx, void *dptr)
{
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipFree(dptr));
    return 0;
}

int rh264_h2d(rh264_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int rh264_d2h(rh264_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int rh264_memset(rh264_ctx *ctx, void *dst, int value, size_t bytes)
{
    HIP_TRY(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return 0;
}

int rh264_sync(rh264_ctx *ctx)
{
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int rh264_event_record(rh264_ctx *ctx, int slot)
{
    if (slot < 0 || slot >= 16) return fail("event slot %d", slot);
    HIP_TRY(hipEventRecord(ctx->ev[slot], ctx->stream));
    return 0;
}

int rh264_event_elapsed_ms(rh264_ctx *ctx, int a, int b, float *ms)
{
    if (a < 0 || a >= 16 || b < 0 || b >= 16) return fail("event slot");
    HIP_TRY(hipEventSynchronize(ctx->ev[b]));
    HIP_TRY(hipEventElapsedTime(ms, ctx->ev[a], ctx->ev[b]));
    return 0;
}

void rh264_lambda_table(double lambda, int32_t *lr)
{
    for (int b = 0; b < RH264_LR_N; b++) {
        int64_t v = (int64_t)(lambda * (double)b); // `as i64` truncation, mb16.rs:2563
        lr[b] = v > INT32_MAX ? INT32_MAX : (int32_t)v;
    }
}

size_t rh264_planes_bytes(int cw, int ch, int pad)
{
    return (size_t)4 * (size_t)(cw + 2 * pad) * (size_t)(ch + 2 * pad) + 256;
}

static bool geom_ok(int cw, int ch, int pad)
{
    return cw > 0 && ch > 0 && cw % 16 == 0 && ch % 16 == 0 && pad >= 4 && pad <= 128 && pad % 4 == 0;
}

int rh264_build_planes(rh264_ctx *ctx, const uint8_t *frames, size_t frame_stride, int nframes, int cw,
                       int ch, int pad, uint8_t *planes)
{
    if (!geom_ok(cw, ch, pad)) return RH264_DECLINED;
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("build_planes: nframes %d > 65535", nframes);
    const int pw = cw + 2 * pad, ph = ch + 2 * pad;
    if (ctx->planes_rw) {
        dim3 grid((pw + PB_W - 1) / PB_W, (ph + PB_H - 1) / PB_H, nframes);
        hipLaunchKernelGGL(k_build_planes_rw, grid, dim3(256), 0, ctx->stream, frames, frame_stride, cw, ch, pad,
                           planes, rh264_planes_bytes(cw, ch, pad));
        LAUNCH_CHECK("k_build_planes_rw");
        return 0;
    }
    dim3 grid((pw + PT_W - 1) / PT_W, (ph + PT_H - 1) / PT_H, nframes);
    hipLaunchKernelGGL(k_build_planes, grid, dim3(256), 0, ctx->stream, frames, frame_stride, cw, ch, pad,
                       planes, rh264_planes_bytes(cw, ch, pad));
    LAUNCH_CHECK("k_build_planes");
    return 0;
}

static bool shape_ok(int w, int h)
{
    return (w == 16 && (h == 16 || h == 8)) || (w == 8 && (h == 16 || h == 8 || h == 4)) ||
           (w == 4 && (h == 8 || h == 4));
}

static int check_strides(size_t cur_stride, size_t planes_stride)
{
    if (cur_stride % 4 || planes_stride % 4) return fail("frame and plane-set strides must be multiples of 4");
    return 0;
}

static int check_params(const rh264_search_params *p)
{
    if (!p) return fail("null params");
    if (!geom_ok(p->cw, p->ch, p->pad) || !shape_ok(p->bw, p->bh)) return RH264_DECLINED;
    if (p->range < 1 || p->range > 64) return RH264_DECLINED;
    if (p->subpel_pat < -1 || p->subpel_pat > 4) return RH264_DECLINED;
    if (p->hadamard != 0 && p->hadamard != 4 && p->hadamard != 8) return fail("hadamard %d: 4 or 8", p->hadamard);
    if (p->hadamard == 8 && (p->bw != 16 || p->bh != 16)) return RH264_DECLINED;
    for (int i = 0; i < RH264_LR_N; i++)
        if (p->lr[i] < 0 || p->lr_fp[i] < 0 || p->lr[i] > (1 << 26) || p->lr_fp[i] > (1 << 26))
            return fail("lambda table entry %d out of range", i);
    return 0;
}

} // extern "C"

template <int BW, int BH, int V>
static void launch_full_v(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                          size_t cur_stride, const uint8_t *planes, size_t planes_stride, const int32_t *centres,
                          const int32_t *mvps, uint64_t *keys)
{
    const long nblk = (long)(p->cw / BW) * (p->ch / BH);
    const long lanes = nblk * fs_groups<V>(2 * p->range + 1);
    dim3 grid((unsigned)((lanes + 255) / 256), nframes);
    hipLaunchKernelGGL((k_full_search<BW, BH, V>), grid, dim3(256), 0, ctx->stream, *p, cur, cur_stride, planes,
                       planes_stride, centres, mvps, keys);
}

template <int BW, int BH>
static int launch_full(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                       size_t cur_stride, const uint8_t *planes, size_t planes_stride,
                       const int32_t *centres, const int32_t *mvps, uint64_t *keys)
{
    if (ctx->fs_kernel == 1)
        launch_full_v<BW, BH, 1>(ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps, keys);
    else if (ctx->fs_kernel == 2)
        launch_full_v<BW, BH, 2>(ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps, keys);
    else
        launch_full_v<BW, BH, 0>(ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps, keys);
    LAUNCH_CHECK("k_full_search");
    return 0;
}

template <int BW, int BH>
static int launch_refine(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                         size_t cur_stride, const uint8_t *planes, size_t planes_stride,
                         const uint64_t *keys, const int32_t *centres, const int32_t *start_mv,
                         const int32_t *mvps, int32_t *int_mv, int32_t *int_cost, int32_t *mv,
                         int32_t *cost, int32_t *evals, const int32_t *fflag = nullptr,
                         const uint64_t *fkeys = nullptr, int fskip = 0)
{
    constexpr int GPW = RefShape<BW, BH>::GPW;
    const int nblk = (p->cw / BW) * (p->ch / BH);
    if constexpr (BW == 16 && BH == 16) {
        // the 8x8 transform exists only in the matrix-core kernels
        const bool t8 = p->hadamard == 8;
        if (ctx->refine_mfma || t8) {
            const dim3 grid((unsigned)((nblk + 3) / 4), nframes);
#define MX_LAUNCH(PT)                                                                                                \
    do {                                                                                                             \
        if (t8)                                                                                                      \
            hipLaunchKernelGGL((k_refine_mx16<PT, true>), grid, dim3(256), 0, ctx->stream, *p, cur, cur_stride,      \
                               planes, planes_stride, keys, centres, start_mv, mvps, int_mv, int_cost, mv, cost,     \
                               evals);                                                                               \
        else                                                                                                         \
            hipLaunchKernelGGL((k_refine_mx16<PT, false>), grid, dim3(256), 0, ctx->stream, *p, cur, cur_stride,     \
                               planes, planes_stride, keys, centres, start_mv, mvps, int_mv, int_cost, mv, cost,     \
                               evals);                                                                               \
    } while (0)
            if (p->subpel_pat == 4) {
                if (t8)
                    hipLaunchKernelGGL(k_refine_wn16<true>, grid, dim3(256), 0, ctx->stream, *p, cur, cur_stride,
                                       planes, planes_stride, keys, centres, start_mv, mvps, int_mv, int_cost, mv,
                                       cost, evals, fflag, fkeys, fskip);
                else
                    hipLaunchKernelGGL(k_refine_wn16<false>, grid, dim3(256), 0, ctx->stream, *p, cur, cur_stride,
                                       planes, planes_stride, keys, centres, start_mv, mvps, int_mv, int_cost, mv,
                                       cost, evals, fflag, fkeys, fskip);
                LAUNCH_CHECK("k_refine_wn16");
                return 0;
            }
            switch (p->subpel_pat) {
            case 0: MX_LAUNCH(0); break;
            case 1: MX_LAUNCH(1); break;
            case 2: MX_LAUNCH(2); break;
            case 3: MX_LAUNCH(3); break;
            default: MX_LAUNCH(-1); break;
            }
#undef MX_LAUNCH
            LAUNCH_CHECK("k_refine_mx16");
            return 0;
        }
    }
    dim3 grid((unsigned)((nblk + GPW - 1) / GPW), nframes);
    if (RefShape<BW, BH>::NSB >= 8 && ctx->refine_lds)
        hipLaunchKernelGGL((k_subpel_refine<BW, BH, RefShape<BW, BH>::NSB >= 8>), grid, dim3(256), 0, ctx->stream, *p,
                           cur, cur_stride, planes, planes_stride, keys, centres, start_mv, mvps, int_mv, int_cost, mv,
                           cost, evals);
    else
        hipLaunchKernelGGL((k_subpel_refine<BW, BH, false>), grid, dim3(256), 0, ctx->stream, *p, cur, cur_stride,
                           planes, planes_stride, keys, centres, start_mv, mvps, int_mv, int_cost, mv, cost, evals);
    LAUNCH_CHECK("k_subpel_refine");
    return 0;
}

extern "C" {

#define SHAPE_DISPATCH(FN, ...)                                                               \
    switch (p->bw * 100 + p->bh) {                                                            \
    case 1616: return FN<16, 16>(__VA_ARGS__);                                                \
    case 1608: return FN<16, 8>(__VA_ARGS__);                                                 \
    case 816: return FN<8, 16>(__VA_ARGS__);                                                  \
    case 808: return FN<8, 8>(__VA_ARGS__);                                                   \
    case 804: return FN<8, 4>(__VA_ARGS__);                                                   \
    case 408: return FN<4, 8>(__VA_ARGS__);                                                   \
    case 404: return FN<4, 4>(__VA_ARGS__);                                                   \
    default: return RH264_DECLINED;                                                           \
    }

// The integer search into `keys` (nframes x blocks argmin keys, cleared here by a
// fill kernel) and the refinement from them; neither allocates.
static int full_search_into(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                            size_t cur_stride, const uint8_t *planes, size_t planes_stride,
                            const int32_t *centres, const int32_t *mvps, uint64_t *keys)
{
    const size_t nk = (size_t)(p->cw / p->bw) * (p->ch / p->bh) * (size_t)nframes;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, ctx->stream, keys, nk, ~0ull);
    LAUNCH_CHECK("k_fill_u64");
    SHAPE_DISPATCH(launch_full, ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps, keys);
}

static int refine_from(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                       size_t cur_stride, const uint8_t *planes, size_t planes_stride, const uint64_t *keys,
                       const int32_t *centres, const int32_t *mvps, int32_t *int_mv, int32_t *int_cost, int32_t *mv,
                       int32_t *cost)
{
    SHAPE_DISPATCH(launch_refine, ctx, p, nframes, cur, cur_stride, planes, planes_stride, keys, centres, nullptr,
                   mvps, int_mv, int_cost, mv, cost, nullptr);
}

static int full_search_keys(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                            size_t cur_stride, const uint8_t *planes, size_t planes_stride,
                            const int32_t *centres, const int32_t *mvps)
{
    const size_t nk = (size_t)(p->cw / p->bw) * (p->ch / p->bh) * (size_t)nframes;
    int rc = ensure_keys(ctx, nk);
    if (rc) return rc;
    return full_search_into(ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps, ctx->keys);
}

static int refine_keys(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                       size_t cur_stride, const uint8_t *planes, size_t planes_stride, const int32_t *centres,
                       const int32_t *mvps, int32_t *int_mv, int32_t *int_cost, int32_t *mv, int32_t *cost)
{
    return refine_from(ctx, p, nframes, cur, cur_stride, planes, planes_stride, ctx->keys, centres, mvps, int_mv,
                       int_cost, mv, cost);
}

int rh264_full_search(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                      size_t cur_stride, const uint8_t *planes, size_t planes_stride, const int32_t *centres,
                      const int32_t *mvps, int32_t *int_mv, int32_t *int_cost)
{
    int rc = check_params(p);
    if (rc) return rc;
    rc = check_strides(cur_stride, planes_stride);
    if (rc) return rc;
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    if (!int_mv || !int_cost) return fail("full_search: null outputs");
    rc = full_search_keys(ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps);
    if (rc) return rc;
    // decode keys with a no-ring refine pass writing only the integer outputs
    const size_t nblk = (size_t)(p->cw / p->bw) * (p->ch / p->bh) * nframes;
    rc = ensure_tmp_mv(ctx, 3 * nblk);
    if (rc) return rc;
    rh264_search_params q = *p;
    q.subpel_pat = -1;
    SHAPE_DISPATCH(launch_refine, ctx, &q, nframes, cur, cur_stride, planes, planes_stride, ctx->keys, centres,
                   nullptr, mvps, int_mv, int_cost, ctx->tmp_mv, ctx->tmp_mv + 2 * nblk, nullptr);
}

int rh264_subpel_refine(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                        size_t cur_stride, const uint8_t *planes, size_t planes_stride, const int32_t *start_mv,
                        const int32_t *mvps, int32_t *mv, int32_t *cost, int32_t *evals)
{
    int rc = check_params(p);
    if (rc) return rc;
    rc = check_strides(cur_stride, planes_stride);
    if (rc) return rc;
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    if (!start_mv || !mv || !cost) return fail("subpel_refine: null pointers");
    SHAPE_DISPATCH(launch_refine, ctx, p, nframes, cur, cur_stride, planes, planes_stride, nullptr, nullptr,
                   start_mv, mvps, nullptr, nullptr, mv, cost, evals);
}

// The next free profiling slot (nullptr when profiling is off or full); the caller
// records nphases + 1 events into it.
static hipEvent_t *prof_slot(rh264_ctx *ctx, int nphases)
{
    if (ctx->prof_n >= ctx->prof_cap) return nullptr;
    ctx->prof_np[(size_t)ctx->prof_n] = nphases;
    return &ctx->prof[(size_t)RH264_PROF_EV * ctx->prof_n++];
}

int rh264_plain_search(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                       size_t cur_stride, const uint8_t *planes, size_t planes_stride, const int32_t *centres,
                       const int32_t *mvps, int32_t *int_mv, int32_t *int_cost, int32_t *mv, int32_t *cost)
{
    int rc = check_params(p);
    if (rc) return rc;
    rc = check_strides(cur_stride, planes_stride);
    if (rc) return rc;
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    if (!mv || !cost) return fail("plain_search: null outputs");
    hipEvent_t *ev = prof_slot(ctx, 2);
    if (ev) HIP_TRY(hipEventRecord(ev[0], ctx->stream));
    rc = full_search_keys(ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps);
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[1], ctx->stream));
    rc = refine_keys(ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps, int_mv, int_cost,
                     mv, cost);
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[2], ctx->stream));
    return 0;
}

int rh264_plain_search_noalloc(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                               size_t cur_stride, const uint8_t *planes, size_t planes_stride, const int32_t *centres,
                               const int32_t *mvps, int32_t *int_mv, int32_t *int_cost, int32_t *mv, int32_t *cost)
{
    int rc = check_params(p);
    if (rc) return rc;
    if (!shape_ok(p->bw, p->bh)) return RH264_DECLINED;
    rc = check_strides(cur_stride, planes_stride);
    if (rc) return rc;
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    if (!mv || !cost) return fail("plain_search_noalloc: null outputs");
    const size_t nk = (size_t)(p->cw / p->bw) * (p->ch / p->bh) * (size_t)nframes;
    if (nk > (size_t)RH264_NOALLOC_KEYS)
        return fail("plain_search_noalloc: %zu blocks exceed the static scratch's %u", nk, (unsigned)RH264_NOALLOC_KEYS);
    if (!ctx->noalloc_keys) return fail("plain_search_noalloc: no scratch address for this context");
    hipEvent_t *ev = prof_slot(ctx, 2);
    if (ev) HIP_TRY(hipEventRecord(ev[0], ctx->stream));
    rc = full_search_into(ctx, p, nframes, cur, cur_stride, planes, planes_stride, centres, mvps, ctx->noalloc_keys);
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[1], ctx->stream));
    rc = refine_from(ctx, p, nframes, cur, cur_stride, planes, planes_stride, ctx->noalloc_keys, centres, mvps, int_mv,
                     int_cost, mv, cost);
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[2], ctx->stream));
    return 0;
}

int rh264_set_profiling(rh264_ctx *ctx, int capacity)
{
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (hipEvent_t e : ctx->prof) HIP_TRY(hipEventDestroy(e));
    ctx->prof.clear();
    ctx->prof_np.clear();
    ctx->prof_cap = ctx->prof_n = 0;
    if (capacity <= 0) return 0;
    ctx->prof.resize((size_t)RH264_PROF_EV * capacity);
    ctx->prof_np.assign((size_t)capacity, 0);
    for (auto &e : ctx->prof) HIP_TRY(hipEventCreate(&e));
    ctx->prof_cap = capacity;
    return 0;
}

int rh264_profile_phases(rh264_ctx *ctx, int i, float *ms, int cap, int *nphases)
{
    if (i < 0 || i >= ctx->prof_n) return fail("profile slot %d of %d", i, ctx->prof_n);
    if (!ms || !nphases) return fail("rh264_profile_phases: null pointers");
    const int np = ctx->prof_np[(size_t)i];
    hipEvent_t *ev = &ctx->prof[(size_t)RH264_PROF_EV * i];
    HIP_TRY(hipEventSynchronize(ev[np]));
    for (int k = 0; k < np && k < cap; k++) HIP_TRY(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
    *nphases = np;
    return 0;
}

int rh264_profile_times(rh264_ctx *ctx, int i, float *ms, int cap, int *nevents)
{
    if (i < 0 || i >= ctx->prof_n) return fail("profile slot %d of %d", i, ctx->prof_n);
    if (!ms || !nevents) return fail("rh264_profile_times: null pointers");
    const int np = ctx->prof_np[(size_t)i];
    hipEvent_t *ev = &ctx->prof[(size_t)RH264_PROF_EV * i];
    HIP_TRY(hipEventSynchronize(ev[np]));
    for (int k = 0; k < np && k < cap; k++) HIP_TRY(hipEventElapsedTime(&ms[k], ev[0], ev[k + 1]));
    *nevents = np;
    return 0;
}

int rh264_profile_count(rh264_ctx *ctx) { return ctx->prof_n; }

int rh264_tune(rh264_ctx *ctx, const char *knob, int value)
{
    if (!knob) return fail("rh264_tune: null knob");
    if (!strcmp(knob, "fs_kernel")) {
        if (value < 0 || value > 2) return fail("rh264_tune: fs_kernel %d", value);
        ctx->fs_kernel = value;
        return 0;
    }
    if (!strcmp(knob, "refine_lds")) {
        ctx->refine_lds = value ? 1 : 0;
        return 0;
    }
    if (!strcmp(knob, "planes_rw")) {
        ctx->planes_rw = value ? 1 : 0;
        return 0;
    }
    if (!strcmp(knob, "la_scaled")) {
        if (value < 0 || value > 1) return fail("la_scaled %d: 0 or 1", value);
        ctx->la_scaled = value;
        return 0;
    }
    if (!strcmp(knob, "decide_mode")) {
        if (value < 0 || value > 2) return fail("decide_mode %d: 0, 1 or 2", value);
        ctx->decide_mode = value;
        return 0;
    }
    if (!strcmp(knob, "cand_lds")) {
        if (value < 0 || value > 1) return fail("cand_lds %d: 0 or 1", value);
        ctx->cand_lds = value;
        return 0;
    }
    if (!strcmp(knob, "fallback_v")) {
        if (value < 0 || value > 2) return fail("fallback_v %d: 0, 1 or 2", value);
        ctx->fallback_v = value;
        return 0;
    }
    if (!strcmp(knob, "fallback_wgs")) {
        if (value < 1 || value > 65536) return fail("fallback_wgs %d: 1..65536", value);
        ctx->fallback_wgs = value;
        return 0;
    }
    if (!strcmp(knob, "fallback_overlap")) {
        if (value < 0 || value > 1) return fail("fallback_overlap %d: 0 or 1", value);
        ctx->fallback_overlap = value;
        return 0;
    }
    if (!strcmp(knob, "cand_quad")) {
        if (value < 0 || value > 2) return fail("cand_quad %d: 0, 1 or 2", value);
        ctx->cand_quad = value;
        return 0;
    }
    if (!strcmp(knob, "refine_mfma")) {
        ctx->refine_mfma = value ? 1 : 0;
        return 0;
    }
    if (!strcmp(knob, "mvnet_fused")) {
        ctx->mvnet_fused = value < 0 ? 0 : value > 2 ? 2 : value;
        return 0;
    }
    return fail("rh264_tune: unknown knob '%s'", knob);
}

int rh264_profile_get(rh264_ctx *ctx, int i, float *search_ms, float *refine_ms)
{
    if (i < 0 || i >= ctx->prof_n) return fail("profile slot %d of %d", i, ctx->prof_n);
    const int np = ctx->prof_np[(size_t)i];
    hipEvent_t *ev = &ctx->prof[(size_t)RH264_PROF_EV * i];
    HIP_TRY(hipEventSynchronize(ev[np]));
    HIP_TRY(hipEventElapsedTime(search_ms, ev[0], ev[np - 1]));
    HIP_TRY(hipEventElapsedTime(refine_ms, ev[np - 1], ev[np]));
    return 0;
}

int rh264_block_metric(rh264_ctx *ctx, int kind, int w, int h, int n, const uint8_t *src, int ss,
                       const int64_t *src_off, const uint8_t *ref, int rs, const int64_t *ref_off,
                       const int64_t *ref_off2, int32_t *out)
{
    if (kind < 0 || kind > 4) return fail("block_metric: kind %d", kind);
    if (!(shape_ok(w, h) || (w == 4 && h == 4))) return RH264_DECLINED;
    if (kind == RH264_METRIC_SATD8 && (w % 8 || h % 8)) return RH264_DECLINED;
    if (kind == RH264_METRIC_SATD_AVG && !ref_off2) return fail("block_metric: SATD_AVG needs ref_off2");
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_block_metric, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, kind, w, h, n,
                       src, ss, src_off, ref, rs, ref_off, ref_off2, out);
    LAUNCH_CHECK("k_block_metric");
    return 0;
}

int rh264_predict_blocks(rh264_ctx *ctx, const uint8_t *planes, int cw, int ch, int pad, int bw, int bh, int n,
                         const int32_t *xy, const int32_t *mv, uint8_t *out)
{
    if (!geom_ok(cw, ch, pad)) return RH264_DECLINED;
    if (bw < 1 || bw > 16 || bh < 1 || bh > 16) return RH264_DECLINED;
    if (n <= 0) return 0;
    const long total = (long)n * bw * bh;
    hipLaunchKernelGGL(k_predict, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, planes,
                       cw + 2 * pad, ch + 2 * pad, pad, bw, bh, n, xy, mv, out);
    LAUNCH_CHECK("k_predict");
    return 0;
}


int rh264_mbtree_downsample(rh264_ctx *ctx, const uint8_t *frames, size_t frame_stride, int nframes, int cw, int ch,
                            uint8_t *half)
{
    if (cw <= 0 || ch <= 0 || cw % 16 || ch % 16) return RH264_DECLINED;
    if (frame_stride % 8) return fail("mbtree_downsample: frame stride must be a multiple of 8");
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    const long groups = (long)(cw / 2) * (ch / 2) / 4;
    hipLaunchKernelGGL(k_downsample2x, dim3((unsigned)((groups + 255) / 256), nframes), dim3(256), 0, ctx->stream,
                       frames, frame_stride, cw, ch, half);
    LAUNCH_CHECK("k_downsample2x");
    return 0;
}

int rh264_mbtree_costs(rh264_ctx *ctx, const uint8_t *half, int nframes, int mb_w, int mb_h, int32_t *intra,
                       int32_t *inter, int32_t *mv)
{
    if (mb_w <= 0 || mb_h <= 0) return RH264_DECLINED;
    if (nframes <= 0) return 0;
    if (nframes > 65535 || mb_h > 65535) return fail("mbtree_costs: nframes/mb_h too large");
    const int nm = mb_w * mb_h;
    hipLaunchKernelGGL(k_mbtree_intra, dim3((unsigned)((nm + 15) / 16), nframes), dim3(256), 0, ctx->stream, half,
                       mb_w, mb_h, intra, inter, mv);
    LAUNCH_CHECK("k_mbtree_intra");
    if (nframes == 1) return 0;
    hipLaunchKernelGGL(k_mbtree_col0, dim3(nframes - 1), dim3(64), 0, ctx->stream, half, mb_w, mb_h, intra, inter, mv);
    LAUNCH_CHECK("k_mbtree_col0");
    if (mb_w > 1) {
        hipLaunchKernelGGL(k_mbtree_rows, dim3(mb_h, nframes - 1), dim3(64), 0, ctx->stream, half, mb_w, mb_h, intra,
                           inter, mv);
        LAUNCH_CHECK("k_mbtree_rows");
    }
    return 0;
}

static int check_diamond(const rh264_diamond *dc, const char *who)
{
    if (!dc) return fail("%s: null diamond", who);
    if (dc->nsteps < 0 || dc->nsteps > 8) return RH264_DECLINED;
    for (int i = 0; i < dc->nsteps; i++)
        if (dc->steps[i] < 1 || dc->steps[i] > 1024) return fail("%s: diamond step %d", who, dc->steps[i]);
    const int32_t flags[] = {dc->sadfp, dc->snap, dc->fc, dc->sad_only, dc->mv_cost, dc->rescue, dc->rescue_fast};
    for (int32_t f : flags)
        if (f != 0 && f != 1) return fail("%s: a diamond flag is %d (0 or 1)", who, f);
    if (dc->mv_cost) {
        if (!dc->mv_tab) return fail("%s: mv_cost 1 without mv_tab", who);
        if (!(dc->lambda >= 0.0 && dc->lambda <= 4096.0) || !(dc->lambda_fp >= 0.0 && dc->lambda_fp <= 4096.0))
            return fail("%s: lambda %g / %g outside 0..4096", who, dc->lambda, dc->lambda_fp);
    }
    if (dc->rescue && (dc->rescue_range < 0 || dc->rescue_range > 64 || dc->rescue_var < 0))
        return fail("%s: rescue range %d / variance bound %d", who, dc->rescue_range, dc->rescue_var);
    return 0;
}

void rh264_mv_cost_smooth(uint16_t *tab)
{
    for (int a = 0; a < 4096; a++)
        tab[a] = (uint16_t)std::lround(4.0 * (2.0 * std::log2((double)(a + 1)) + 0.718 + (a != 0 ? 1.0 : 0.0)));
}

int rh264_motion_search(rh264_ctx *ctx, const rh264_search_params *p, const rh264_diamond *dc, int nframes,
                        const uint8_t *cur, size_t cur_stride, const uint8_t *planes, size_t planes_stride,
                        const int32_t *preds, int npreds, int32_t *mv, int32_t *cost, int32_t *events)
{
    int rc = check_params(p);
    if (rc) return rc;
    rc = check_diamond(dc, "motion_search");
    if (rc) return rc;
    if (!shape_ok(p->bw, p->bh)) return RH264_DECLINED;
    if (p->hadamard == 8 && (p->bw != 16 || p->bh != 16)) return RH264_DECLINED;
    if (npreds < 0 || npreds > 8) return RH264_DECLINED;
    if (npreds > 0 && !preds) return fail("motion_search: null predictors");
    if (!mv || !cost) return fail("motion_search: null output");
    rc = check_strides(cur_stride, planes_stride);
    if (rc) return rc;
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    const int nblk = (p->cw / p->bw) * (p->ch / p->bh);
    const dim3 grid((unsigned)((nblk + 3) / 4), nframes);
    if (p->hadamard == 8)
        hipLaunchKernelGGL(k_motion_search<true>, grid, dim3(256), 0, ctx->stream, *p, *dc, cur, cur_stride, planes,
                           planes_stride, preds, npreds, mv, cost, events);
    else
        hipLaunchKernelGGL(k_motion_search<false>, grid, dim3(256), 0, ctx->stream, *p, *dc, cur, cur_stride, planes,
                           planes_stride, preds, npreds, mv, cost, events);
    LAUNCH_CHECK("k_motion_search");
    return 0;
}

static int ref_bits(int r, int nrefs);
static int ensure_buf(rh264_ctx *ctx, void **buf, size_t *cap, size_t bytes);

#define MS_WF_LAUNCH(T8, RW)                                                                                  \
    hipLaunchKernelGGL((k_ms_wavefront<T8, RW>), grid, dim3(256), 0, ctx->stream, *p, *dc, cur, cur_stride, planes, \
                       planes_stride, nrefs, refc, t, ref, mv, cost)

int rh264_motion_search_wavefront(rh264_ctx *ctx, const rh264_search_params *p, const rh264_diamond *dc,
                                  int nframes, int nrefs, double lambda_ref, const uint8_t *cur, size_t cur_stride,
                                  const uint8_t *planes, size_t planes_stride, int32_t *ref, int32_t *mv,
                                  int32_t *cost)
{
    int rc = check_params(p);
    if (rc) return rc;
    rc = check_diamond(dc, "motion_search_wavefront");
    if (rc) return rc;
    if (p->bw != 16 || p->bh != 16) return RH264_DECLINED;
    if (nrefs < 1 || nrefs > 4) return RH264_DECLINED;
    if (!(lambda_ref >= 0.0 && lambda_ref < 1e8)) return fail("motion_search_wavefront: lambda_ref %g", lambda_ref);
    const bool resc = dc->rescue && !dc->sad_only;
    if (resc && (dc->rescue_learn < 0 || dc->rescue_payoff < 0 || dc->rescue_payoff > 100))
        return fail("motion_search_wavefront: rescue window %d / payoff %d", dc->rescue_learn, dc->rescue_payoff);
    if (!ref || !mv || !cost) return fail("motion_search_wavefront: null output");
    rc = check_strides(cur_stride, planes_stride);
    if (rc) return rc;
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    MsRefCost refc{};
    for (int r = 0; r < nrefs; r++) refc.c[r] = (int32_t)(int64_t)(lambda_ref * (double)ref_bits(r, nrefs));
    const int rw = nrefs == 1 ? 1 : nrefs == 2 ? 2 : 4, per_wg = 4 / rw;
    const int nbx = p->cw / 16, nby = p->ch / 16, ndiag = (nbx - 1) + 2 * (nby - 1) + 1;
    const bool t8 = p->hadamard == 8;
    // mode 2 picks by measurement at 1080p: the fixed point for one or two
    // references, where few passes settle it, the anti-diagonals for more
    if (resc || ctx->decide_mode == 1 || (ctx->decide_mode == 2 && nrefs <= 2)) {
        // fixed-point passes: S starts at (ref 0, mv 0) everywhere and no predictor
        // matches the sentinel, so the first pass searches every macroblock
        const size_t ntot = (size_t)nframes * nbx * nby;
        if (ntot > (size_t)INT32_MAX / 8) return fail("motion_search_wavefront: %zu macroblocks", ntot);
        // used [ntot][8] | list [ntot] | counts [2] (+pad) | with the rescue: usedf
        // [ntot][4] | events [ntot][4] | fstate [nframes][2]
        const size_t base_words = ntot * 9 + 64;
        const size_t words = base_words + (resc ? ntot * 8 + (size_t)nframes * 2 : 0);
        rc = ensure_buf(ctx, &ctx->ms_buf, &ctx->ms_cap, words * 4);
        if (rc) return rc;
        int32_t *used = (int32_t *)ctx->ms_buf, *list = used + ntot * 8;
        uint32_t *counts = (uint32_t *)(list + ntot); // [2], alternating between passes
        int32_t *usedf = resc ? (int32_t *)ctx->ms_buf + base_words : nullptr;
        int32_t *events = resc ? usedf + ntot * 4 : nullptr, *fstate = resc ? events + ntot * 4 : nullptr;
        const unsigned fg = (unsigned)((ntot * 8 + 255) / 256), ng = (unsigned)((ntot + 255) / 256);
        if (resc) {
            // switches unknown (a sentinel: every first check differs), no events, every
            // frame's window open
            hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, ctx->stream, (unsigned *)usedf, ntot * 4,
                               0x7fffffffu);
            hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, ctx->stream, (unsigned *)events, ntot * 4, 0u);
            hipLaunchKernelGGL(k_ms_latch_init, dim3((unsigned)((nframes + 255) / 256)), dim3(256), 0, ctx->stream,
                               nframes, fstate);
        }
        hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, ctx->stream, (unsigned *)used, ntot * 8, 0x80000000u);
        hipLaunchKernelGGL(k_fill_u32, dim3(ng), dim3(256), 0, ctx->stream, (unsigned *)ref, ntot, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(2 * ng), dim3(256), 0, ctx->stream, (unsigned *)mv, ntot * 2, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, counts, (size_t)2, 0u);
        LAUNCH_CHECK("k_fill_u32");
        // enough workgroups to fill the chip on the first (full) pass
        const unsigned sg = (unsigned)std::min<size_t>((ntot + per_wg - 1) / per_wg, 256 * 16);
        const int chunk = 8;
        // the anti-diagonals bound the passes; with the rescue a switch depends on every
        // earlier search of its frame, which bounds them by the searches in a frame
        const long kmax = ndiag + (resc ? (long)nbx * nby * nrefs : 0);
        for (long k = 0; k <= kmax;) {
            for (int c = 0; c < chunk && k <= kmax; c++, k++) {
                uint32_t *cnt = counts + (k & 1), *nxt = counts + ((k + 1) & 1);
                hipLaunchKernelGGL(k_ms_check, dim3(ng), dim3(256), 0, ctx->stream, nbx, nbx * nby, (int)ntot, nrefs,
                                   (const int32_t *)ref, (const int32_t *)mv, used, list, cnt, usedf,
                                   (const int32_t *)fstate);
                LAUNCH_CHECK("k_ms_check");
#define MS_FP_LAUNCH(T8, RW)                                                                                          \
    hipLaunchKernelGGL((k_ms_listed<T8, RW>), dim3(sg), dim3(256), 0, ctx->stream, *p, *dc, cur, cur_stride, planes,   \
                       planes_stride, nrefs, refc, (const int32_t *)used, (const int32_t *)list,                        \
                       (const uint32_t *)cnt, nxt, ref, mv, cost, (const int32_t *)usedf, events)
                if (rw == 1) {
                    if (t8) MS_FP_LAUNCH(true, 1); else MS_FP_LAUNCH(false, 1);
                } else if (rw == 2) {
                    if (t8) MS_FP_LAUNCH(true, 2); else MS_FP_LAUNCH(false, 2);
                } else {
                    if (t8) MS_FP_LAUNCH(true, 4); else MS_FP_LAUNCH(false, 4);
                }
#undef MS_FP_LAUNCH
                LAUNCH_CHECK("k_ms_listed");
                if (resc) {
                    hipLaunchKernelGGL(k_ms_latch, dim3((unsigned)nframes), dim3(64), 0, ctx->stream, nbx * nby, nrefs,
                                       4, dc->rescue_learn, dc->rescue_payoff, (const int32_t *)events, fstate);
                    LAUNCH_CHECK("k_ms_latch");
                }
            }
            // the last check's count (the search clears only the other counter)
            uint32_t last = 1;
            HIP_TRY(hipMemcpyAsync(&last, counts + ((k - 1) & 1), 4, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            if (last == 0) return 0;
        }
        return fail("motion_search_wavefront: no fixed point after %ld passes", kmax + 1);
    }
    // anti-diagonals x + 2y = t: each macroblock's left, above and above-right
    // (or above-left) neighbours lie on earlier ones
    for (int t = 0; t < ndiag; t++) {
        const int y_lo = std::max(0, (t - (nbx - 1) + 1) / 2), y_hi = std::min(nby - 1, t / 2);
        const int n_t = y_hi - y_lo + 1;
        if (n_t <= 0) continue;
        const dim3 grid((unsigned)((n_t + per_wg - 1) / per_wg), nframes);
        if (rw == 1) {
            if (t8) MS_WF_LAUNCH(true, 1); else MS_WF_LAUNCH(false, 1);
        } else if (rw == 2) {
            if (t8) MS_WF_LAUNCH(true, 2); else MS_WF_LAUNCH(false, 2);
        } else {
            if (t8) MS_WF_LAUNCH(true, 4); else MS_WF_LAUNCH(false, 4);
        }
        LAUNCH_CHECK("k_ms_wavefront");
    }
    return 0;
}

int rh264_partition_decision(rh264_ctx *ctx, const rh264_search_params *p, const rh264_diamond *dc, int nframes,
                             int nrefs, double lambda_ref, int32_t split_gate, int p8x8, const uint8_t *cur,
                             size_t cur_stride, const uint8_t *planes, size_t planes_stride, int32_t *rec)
{
    int rc = check_params(p);
    if (rc) return rc;
    rc = check_diamond(dc, "partition_decision");
    if (rc) return rc;
    if (p->bw != 16 || p->bh != 16 || p->hadamard == 8) return RH264_DECLINED;
    if (nrefs < 1 || nrefs > 4) return RH264_DECLINED;
    if (!(lambda_ref >= 0.0 && lambda_ref < 1e6)) return fail("partition_decision: lambda %g", lambda_ref);
    const bool resc = dc->rescue && !dc->sad_only;
    if (resc && (dc->rescue_learn < 0 || dc->rescue_payoff < 0 || dc->rescue_payoff > 100))
        return fail("partition_decision: rescue window %d / payoff %d", dc->rescue_learn, dc->rescue_payoff);
    if (p8x8 != 0 && p8x8 != 1) return fail("partition_decision: p8x8 %d", p8x8);
    if (!rec) return fail("partition_decision: null output");
    rc = check_strides(cur_stride, planes_stride);
    if (rc) return rc;
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    PdArgs a{};
    a.nrefs = nrefs;
    a.gate = split_gate;
    a.p8x8 = p8x8;
    a.lme4 = (int32_t)(int64_t)(lambda_ref * 4.0);
    for (int r = 0; r < nrefs; r++) a.refc[r] = (int32_t)(int64_t)(lambda_ref * (double)ref_bits(r, nrefs));
    const int nbx = p->cw / 16, nby = p->ch / 16, ndiag = (nbx - 1) + 2 * (nby - 1) + 1;
    const size_t ntot = (size_t)nframes * nbx * nby;
    if (ntot > (size_t)INT32_MAX / PD_SLOTS) return fail("partition_decision: %zu macroblocks", ntot);
    // used [ntot][8] | list [ntot] | counts [2] (+pad) | with the rescue: usedf and
    // events [ntot][PD_SLOTS] | fstate [nframes][2]
    const size_t base_words = ntot * 9 + 64;
    const size_t words = base_words + (resc ? ntot * 2 * PD_SLOTS + (size_t)nframes * 2 : 0);
    rc = ensure_buf(ctx, &ctx->ms_buf, &ctx->ms_cap, words * 4);
    if (rc) return rc;
    int32_t *used = (int32_t *)ctx->ms_buf, *list = used + ntot * 8;
    uint32_t *counts = (uint32_t *)(list + ntot);
    int32_t *usedf = resc ? (int32_t *)ctx->ms_buf + base_words : nullptr;
    int32_t *events = resc ? usedf + ntot * PD_SLOTS : nullptr, *fstate = resc ? events + ntot * PD_SLOTS : nullptr;
    const int slots = (1 + (p8x8 ? 8 : 4)) * nrefs;
    if (resc) {
        const unsigned sg2 = (unsigned)((ntot * PD_SLOTS + 255) / 256);
        hipLaunchKernelGGL(k_fill_u32, dim3(sg2), dim3(256), 0, ctx->stream, (unsigned *)usedf, ntot * PD_SLOTS,
                           0x7fffffffu);
        hipLaunchKernelGGL(k_fill_u32, dim3(sg2), dim3(256), 0, ctx->stream, (unsigned *)events, ntot * PD_SLOTS, 0u);
        hipLaunchKernelGGL(k_ms_latch_init, dim3((unsigned)((nframes + 255) / 256)), dim3(256), 0, ctx->stream,
                           nframes, fstate);
    }
    const unsigned fg = (unsigned)((ntot * 8 + 255) / 256), rg = (unsigned)((ntot * PD_REC + 255) / 256);
    const unsigned ng = (unsigned)((ntot + 255) / 256);
    // S starts at (16x16, ref 0, mv 0) everywhere; no predictor matches the sentinel,
    // so the first pass decides every macroblock
    hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, ctx->stream, (unsigned *)used, ntot * 8, 0x80000000u);
    hipLaunchKernelGGL(k_fill_u32, dim3(rg), dim3(256), 0, ctx->stream, (unsigned *)rec, ntot * PD_REC, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, counts, (size_t)2, 0u);
    LAUNCH_CHECK("k_fill_u32");
    const unsigned sg = (unsigned)std::min<size_t>(ntot, 256 * 16); // one macroblock per workgroup
    const int chunk = 8;
    const long kmax = ndiag + (resc ? (long)nbx * nby * slots : 0);
    for (long k = 0; k <= kmax;) {
        for (int c = 0; c < chunk && k <= kmax; c++, k++) {
            uint32_t *cnt = counts + (k & 1), *nxt = counts + ((k + 1) & 1);
            hipLaunchKernelGGL(k_pd_check, dim3(ng), dim3(256), 0, ctx->stream, nbx, nbx * nby, (int)ntot, nrefs,
                               (const int32_t *)rec, used, list, cnt, slots, usedf, (const int32_t *)fstate);
            LAUNCH_CHECK("k_pd_check");
            hipLaunchKernelGGL(k_pd_listed<false>, dim3(sg), dim3(256), 0, ctx->stream, *p, *dc, cur, cur_stride, planes,
                               planes_stride, a, (const int32_t *)used, (const int32_t *)list, (const uint32_t *)cnt,
                               nxt, rec, (const int32_t *)usedf, events);
            LAUNCH_CHECK("k_pd_listed");
            if (resc) {
                hipLaunchKernelGGL(k_ms_latch, dim3((unsigned)nframes), dim3(64), 0, ctx->stream, nbx * nby, slots,
                                   PD_SLOTS, dc->rescue_learn, dc->rescue_payoff, (const int32_t *)events, fstate);
                LAUNCH_CHECK("k_ms_latch");
            }
        }
        uint32_t last = 1;
        HIP_TRY(hipMemcpyAsync(&last, counts + ((k - 1) & 1), 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (last == 0) return 0;
    }
    return fail("partition_decision: no fixed point after %ld passes", kmax + 1);
}

int32_t rh264_split_gate(int qp)
{
    static const int32_t qs[6] = {10, 11, 13, 14, 16, 18};
    if (qp < 0 || qp > 51) return -1;
    const int64_t q = (int64_t)qs[qp % 6] << (qp / 6);
    return (int32_t)(((30 * (q + 160)) >> 3) * 2);
}

int rh264_decision_predict(rh264_ctx *ctx, int cw, int ch, int pad, int nframes, int nrefs, const uint8_t *planes,
                           size_t planes_stride, const int32_t *ref, const int32_t *mv, int32_t *skip_mv,
                           uint8_t *pred, uint8_t *skip_pred)
{
    if (cw <= 0 || ch <= 0 || cw % 16 || ch % 16) return RH264_DECLINED;
    if (pad < 4 || pad > 128 || pad % 4) return RH264_DECLINED;
    if (nrefs < 1 || nrefs > 16) return RH264_DECLINED;
    if (!planes || !ref || !mv || !skip_mv || !pred || !skip_pred) return fail("decision_predict: null pointer");
    if (planes_stride < rh264_planes_bytes(cw, ch, pad)) return fail("decision_predict: planes stride too small");
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    if ((size_t)cw * ch / 4 > (size_t)INT32_MAX) return fail("decision_predict: frame too large");
    const unsigned groups = (unsigned)((size_t)cw * ch / 4);
    hipLaunchKernelGGL(k_decision_predict, dim3((groups + 255) / 256, nframes), dim3(256), 0, ctx->stream, cw, ch,
                       pad, nrefs, planes, planes_stride, ref, mv, skip_mv, pred, skip_pred);
    LAUNCH_CHECK("k_decision_predict");
    return 0;
}

// ---- mb-tree QP-offset policy (gop_qp_offsets, mbtree.rs:369-593), host f64 ----
// The f64 propagation is order-dependent (frames descending, macroblocks in raster
// order, four targets per block), so it runs on the host in exactly that order.
static void mbt_propagate_to(double *prev, int mb_w, int mb_h, int mb_x, int mb_y, int mvx, int mvy, double amount)
{
#pragma clang fp contract(off) // every product and sum rounded on its own, as the reference's
    if (amount <= 0.0) return; // mbtree.rs:268-297
    const int rx = std::min(std::max(mb_x * 16 + (mvx >> 2), 0), (mb_w - 1) * 16);
    const int ry = std::min(std::max(mb_y * 16 + (mvy >> 2), 0), (mb_h - 1) * 16);
    const int cx0 = rx / 16, cy0 = ry / 16;
    const double fx = (double)(rx % 16), fy = (double)(ry % 16);
    const double wys[2] = {16.0 - fy, fy}, wxs[2] = {16.0 - fx, fx};
    for (int dy = 0; dy < 2; dy++) {
        if (wys[dy] <= 0.0) continue;
        const int cy = std::min(cy0 + dy, mb_h - 1);
        for (int dx = 0; dx < 2; dx++) {
            if (wxs[dx] <= 0.0) continue;
            const int cx = std::min(cx0 + dx, mb_w - 1);
            prev[cy * mb_w + cx] += amount * (wxs[dx] * wys[dy]) / 256.0;
        }
    }
}

// The grain signature (signals.rs:528-533): three frame-level probes of the GOP's
// frames 1 and 0, each over the interior macroblock walk the reference uses
// (columns and rows 2 .. mb-3 in raster order, every stride-th).
static uint32_t gs_sad16(const uint8_t *cur, const uint8_t *ref, int cw, int bx, int by, int rx, int ry)
{
    uint32_t s = 0;
    for (int y = 0; y < 16; y++) {
        const uint8_t *a = cur + (size_t)(by + y) * cw + bx, *b = ref + (size_t)(ry + y) * cw + rx;
        for (int x = 0; x < 16; x++) s += (uint32_t)std::abs((int)a[x] - (int)b[x]);
    }
    return s;
}

int rh264_grain_signature(const uint8_t *cur, const uint8_t *ref, int cw, int ch, int32_t *grain, double *stats)
{
#pragma clang fp contract(off)
    if (!cur || !ref || !grain) return fail("grain_signature: null pointer");
    if (cw <= 0 || ch <= 0 || cw % 16 || ch % 16) return fail("grain_signature: %dx%d is not a macroblock grid", cw, ch);
    const int mbw = cw / 16, mbh = ch / 16;
    // median macroblock variance over the even rows and columns (signals.rs:247-266)
    std::vector<int64_t> vars;
    vars.reserve((size_t)mbw * mbh);
    for (int my = 0; my < mbh; my++)
        for (int mx = 0; mx < mbw; mx++) {
            int64_t sum = 0, sq = 0;
            for (int r = 0; r < 16; r += 2) {
                const uint8_t *row = cur + (size_t)(my * 16 + r) * cw + mx * 16;
                for (int c = 0; c < 16; c += 2) {
                    sum += row[c];
                    sq += (int64_t)row[c] * row[c];
                }
            }
            vars.push_back((sq - sum * sum / 64) / 64);
        }
    std::sort(vars.begin(), vars.end());
    const int64_t median_var = vars[vars.size() / 2];
    double floor_v = 0.0, mgain = 0.0;
    if (mbw >= 6 && mbh >= 6) {
        const int inner = (mbw - 4) * (mbh - 4);
        // residual floor: the lower quartile of ~48 zero-motion SADs per pixel (signals.rs:311-338)
        std::vector<double> fl;
        for (int i = 0, st = std::max(inner / 48, 1); i < inner; i += st) {
            const int bx = (2 + i % (mbw - 4)) * 16, by = (2 + i / (mbw - 4)) * 16;
            fl.push_back((double)gs_sad16(cur, ref, cw, bx, by, bx, by) / 256.0);
        }
        std::sort(fl.begin(), fl.end());
        floor_v = fl[fl.size() / 4];
        // motion gain: what a +-8 step-4 full-pel grid saves on ~24 blocks (signals.rs:61-129)
        double acc = 0.0;
        unsigned cnt = 0;
        for (int i = 0, st = std::max(inner / 24, 1); i < inner; i += st) {
            const int bx = (2 + i % (mbw - 4)) * 16, by = (2 + i / (mbw - 4)) * 16;
            const uint32_t s0 = gs_sad16(cur, ref, cw, bx, by, bx, by);
            uint32_t best = s0;
            for (int dy = -8; dy <= 8; dy += 4)
                for (int dx = -8; dx <= 8; dx += 4) {
                    const int rx = bx + dx, ry = by + dy;
                    if (rx >= 0 && ry >= 0 && rx + 16 <= cw && ry + 16 <= ch)
                        best = std::min(best, gs_sad16(cur, ref, cw, bx, by, rx, ry));
                }
            acc += (double)(s0 - best) / (double)(s0 + 1);
            cnt++;
        }
        mgain = cnt ? acc / (double)cnt : 0.0;
    }
    *grain = median_var < 200 && floor_v > 5.0 && mgain < 0.1;
    if (stats) {
        stats[0] = (double)median_var;
        stats[1] = floor_v;
        stats[2] = mgain;
    }
    return 0;
}

int rh264_mbtree_qp_offsets(int nframes, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                            const int32_t *mv, double strength, double res_min, double sd_min, int grain,
                            int32_t *offsets, double *stats)
{
#pragma clang fp contract(off)
    if (nframes < 0 || mb_w < 0 || mb_h < 0) return fail("mbtree_qp_offsets: negative size");
    const size_t nm = (size_t)mb_w * mb_h, n = (size_t)nframes;
    if (n == 0 || nm == 0) return 0;
    if (!intra || !inter || !mv || !offsets) return fail("mbtree_qp_offsets: null pointer");
    for (size_t i = 0; i < n * nm; i++)
        if (intra[i] < 1 || inter[i] < 0 || inter[i] > intra[i])
            return fail("mbtree_qp_offsets: costs need 1 <= intra and 0 <= inter <= intra (block %zu)", i);
    std::fill(offsets, offsets + n * nm, 0);
    double st[3] = {0.0, 0.0, 0.0}; // residual fraction, spread, effective strength
    auto done = [&]() {
        if (stats) std::copy(st, st + 3, stats);
        return 0;
    };
    if (strength <= 0.0) return done();
    // the grain latch (mbtree.rs:401-420): a grainy GOP of two or more frames gets no offsets
    if (grain && n >= 2) return done();
    // backward propagation (mbtree.rs:440-455)
    std::vector<double> prop(n * nm, 0.0);
    for (size_t f = n - 1; f >= 1; f--) {
        const double *cur = prop.data() + f * nm;
        double *prev = prop.data() + (f - 1) * nm;
        for (int y = 0; y < mb_h; y++)
            for (int x = 0; x < mb_w; x++) {
                const size_t m = (size_t)y * mb_w + x, g = f * nm + m;
                const double total = (double)intra[g] + cur[m];
                const double frac = (double)(intra[g] - inter[g]) / (double)intra[g];
                mbt_propagate_to(prev, mb_w, mb_h, x, y, mv[2 * g], mv[2 * g + 1], total * frac);
            }
    }
    // predictability back-off: off below res_min (mbtree.rs:478-489)
    double fsum = 0.0, fc = 0.0;
    for (size_t g = nm; g < n * nm; g++) {
        fsum += (double)(intra[g] - inter[g]) / (double)intra[g];
        fc += 1.0;
    }
    st[0] = 1.0 - (fc > 0.0 ? fsum / fc : 0.0);
    st[2] = (res_min > 0.0 && st[0] < res_min) ? 0.0 : strength;
    if (st[2] == 0.0) return done();
    // offsets -s log2(total / intra), centred on the GOP mean (mbtree.rs:506-527)
    std::vector<double> offs(n * nm);
    for (size_t g = 0; g < n * nm; g++) {
        const double in = (double)intra[g];
        offs[g] = -st[2] * std::log2((in + prop[g]) / in);
    }
    const double cnt = (double)(n * nm);
    double sum = 0.0;
    for (size_t g = 0; g < n * nm; g++) sum += offs[g];
    const double mean = sum / cnt;
    for (size_t g = 0; g < n * nm; g++) offs[g] -= mean;
    // differentiation latch on the strength-invariant spread (mbtree.rs:528-583)
    double sq = 0.0;
    for (size_t g = 0; g < n * nm; g++) sq += offs[g] * offs[g];
    st[1] = std::sqrt(sq / cnt) / std::max(st[2], 1e-9);
    if (st[1] < sd_min) return done();
    for (size_t g = 0; g < n * nm; g++) // round half away from zero, clamp to +-6 (mbtree.rs:584-592)
        offsets[g] = (int32_t)std::min(std::max(std::round(offs[g]), -6.0), 6.0);
    return done();
}

int rh264_rc_complexity(rh264_ctx *ctx, const uint8_t *cur, const uint8_t *ref, size_t frame_stride, int nframes,
                        int cw, int ch, double *out)
{
    if (cw <= 0 || ch <= 0 || cw % 16 || ch % 16) return RH264_DECLINED;
    if (frame_stride % 4 || frame_stride < (size_t)cw * ch)
        return fail("rc_complexity: frame stride must be a multiple of 4 and at least cw*ch");
    if (!cur || !out) return fail("rc_complexity: null frame or output");
    if (nframes <= 0) return 0;
    if (nframes > 65535) return fail("nframes %d > 65535", nframes);
    int rc = ensure_keys(ctx, (size_t)nframes);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nframes + 255) / 256)), dim3(256), 0, ctx->stream, ctx->keys,
                       (size_t)nframes, 0ull);
    LAUNCH_CHECK("k_fill_u64");
    const int nmb = (cw / 16) * (ch / 16);
    const int groups = (nmb + 4 * RC_MB_PER_WAVE - 1) / (4 * RC_MB_PER_WAVE);
    hipLaunchKernelGGL(k_rc_complexity, dim3((unsigned)groups, nframes), dim3(256), 0, ctx->stream, cur, ref,
                       frame_stride, cw, ch, (unsigned long long *)ctx->keys);
    LAUNCH_CHECK("k_rc_complexity");
    hipLaunchKernelGGL(k_rc_finalize, dim3((unsigned)((nframes + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const unsigned long long *)ctx->keys, nframes, out);
    LAUNCH_CHECK("k_rc_finalize");
    return 0;
}

static int mbtree_propagate(rh264_ctx *ctx, int nframes, int mb_w, int mb_h, const int32_t *intra,
                            const int32_t *inter, const int32_t *mv, const int32_t *refs, int32_t *prop)
{
    if (mb_w <= 0 || mb_h <= 0) return RH264_DECLINED;
    if (nframes <= 0) return 0;
    const size_t nm = (size_t)mb_w * mb_h, n = nm * nframes;
    int rc = ensure_keys(ctx, n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->keys, n, 0ull);
    LAUNCH_CHECK("k_fill_u64");
    for (int f = nframes - 1; f >= 0; f--) {
        hipLaunchKernelGGL(k_mbtree_prop_i32, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, ctx->stream, f, mb_w,
                           mb_h, intra, inter, mv, refs, (unsigned long long *)ctx->keys, prop);
        LAUNCH_CHECK("k_mbtree_prop_i32");
    }
    return 0;
}

int rh264_mbtree_propagate_i32(rh264_ctx *ctx, int nframes, int mb_w, int mb_h, const int32_t *intra,
                               const int32_t *inter, const int32_t *mv, int32_t *prop)
{
    return mbtree_propagate(ctx, nframes, mb_w, mb_h, intra, inter, mv, nullptr, prop);
}

int rh264_mbtree_propagate_multi_i32(rh264_ctx *ctx, int nframes, int mb_w, int mb_h, const int32_t *intra,
                                     const int32_t *inter, const int32_t *mv, const int32_t *ref, int32_t *prop)
{
    if (!ref) return fail("mbtree_propagate_multi_i32: null ref");
    return mbtree_propagate(ctx, nframes, mb_w, mb_h, intra, inter, mv, ref, prop);
}

static int ensure_buf(rh264_ctx *ctx, void **buf, size_t *cap, size_t bytes);

int rh264_mbtree_costs_multi(rh264_ctx *ctx, const uint8_t *half, int nframes, int mb_w, int mb_h, int nrefs,
                             int range, int32_t *intra, int32_t *inter, int32_t *mv, int32_t *ref)
{
    if (mb_w <= 0 || mb_h <= 0) return RH264_DECLINED;
    if (nrefs < 1 || nrefs > 4 || range < 1 || range > 64) return RH264_DECLINED;
    if (nframes <= 0) return 0;
    if (nframes > 16384) return fail("mbtree_costs_multi: nframes %d > 16384", nframes);
    const int hw = mb_w * 8, hh = mb_h * 8, nm = mb_w * mb_h;
    const int P = (range + 8 + 3) & ~3; // every window and every winner's block inside the padded plane
    const int pw = hw + 2 * P, ph = hh + 2 * P;
    const size_t pstride = (size_t)pw * ph;
    int rc = ensure_buf(ctx, &ctx->la_buf, &ctx->la_cap, pstride * nframes + 256);
    if (rc) return rc;
    uint8_t *planes = (uint8_t *)ctx->la_buf;
    hipLaunchKernelGGL(k_la_pad, dim3((unsigned)((pstride / 4 + 255) / 256), nframes), dim3(256), 0, ctx->stream, half,
                       hw, hh, P, planes, pstride);
    LAUNCH_CHECK("k_la_pad");
    if (nframes > 1) {
        const size_t nk = (size_t)(nframes - 1) * nrefs * nm;
        rc = ensure_keys(ctx, nk);
        if (rc) return rc;
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, ctx->stream, ctx->keys, nk,
                           ~0ull);
        LAUNCH_CHECK("k_fill_u64");
        const dim3 grid_y(1, (unsigned)((nframes - 1) * nrefs));
        if (ctx->fs_kernel == 1) {
            const long lanes = (long)nm * fs_groups<1>(2 * range + 1);
            hipLaunchKernelGGL(k_la_search<1>, dim3((unsigned)((lanes + 255) / 256), grid_y.y), dim3(256), 0,
                               ctx->stream, half, mb_w, mb_h, nrefs, range, P, planes, pstride, ctx->keys);
        } else if (ctx->fs_kernel == 2) {
            const long lanes = (long)nm * fs_groups<2>(2 * range + 1);
            hipLaunchKernelGGL(k_la_search<2>, dim3((unsigned)((lanes + 255) / 256), grid_y.y), dim3(256), 0,
                               ctx->stream, half, mb_w, mb_h, nrefs, range, P, planes, pstride, ctx->keys);
        } else {
            const long lanes = (long)nm * fs_groups<0>(2 * range + 1);
            hipLaunchKernelGGL(k_la_search<0>, dim3((unsigned)((lanes + 255) / 256), grid_y.y), dim3(256), 0,
                               ctx->stream, half, mb_w, mb_h, nrefs, range, P, planes, pstride, ctx->keys);
        }
        LAUNCH_CHECK("k_la_search");
    }
    const long nb = (long)nframes * nm;
    hipLaunchKernelGGL(k_la_cost, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx->stream, half, mb_w, mb_h,
                       nframes, nrefs, range, P, planes, pstride, ctx->keys, (const int32_t *)nullptr,
                       (const int32_t *)nullptr, intra, inter, mv, ref, (int32_t *)nullptr);
    LAUNCH_CHECK("k_la_cost");
    return 0;
}

int rh264_mbtree_costs_multi_learned(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *frames, size_t frame_stride,
                                     const uint8_t *half, int nframes, int cw, int ch, int nrefs, int range,
                                     int local_range, int32_t fallback_sad, int32_t *intra, int32_t *inter,
                                     int32_t *mv, int32_t *ref, int32_t *fallback)
{
    if (!net) return fail("mbtree_costs_multi_learned: null network");
    if (cw <= 0 || ch <= 0 || cw % 16 || ch % 16) return RH264_DECLINED;
    if (nrefs < 1 || nrefs > 4 || range < 1 || range > 64 || local_range < 1 || local_range > 4)
        return RH264_DECLINED;
    if (nframes <= 0) return 0;
    if (nframes > 16384) return fail("mbtree_costs_multi_learned: nframes %d > 16384", nframes);
    const int mb_w = cw / 16, mb_h = ch / 16, hw = mb_w * 8, hh = mb_h * 8, nm = mb_w * mb_h;
    // pad: the exhaustive window, and every proposal (|v| <= 16 half-res pels; scaled
    // ones are clamped to P - 8 - local_range) with its +-local_range window
    const int P = (std::max(range, 16 + local_range) + 8 + 3) & ~3;
    const int pw = hw + 2 * P, ph = hh + 2 * P;
    const size_t pstride = (size_t)pw * ph;
    int rc = ensure_buf(ctx, &ctx->la_buf, &ctx->la_cap, pstride * nframes + 256);
    if (rc) return rc;
    uint8_t *planes = (uint8_t *)ctx->la_buf;
    hipLaunchKernelGGL(k_la_pad, dim3((unsigned)((pstride / 4 + 255) / 256), nframes), dim3(256), 0, ctx->stream, half,
                       hw, hh, P, planes, pstride);
    LAUNCH_CHECK("k_la_pad");
    const size_t nw = (size_t)std::max(nframes - 1, 1) * nrefs * nm; // (frame, reference, block) entries
    const size_t o_t8 = 0, o_mv = o_t8 + (size_t)nrefs * nframes * nm * LS_NCAND * 4;
    const size_t o_fb = o_mv + nw * 8, o_list = o_fb + nw * 4, o_keys = (o_list + nw * 4 + 15) & ~(size_t)15;
    const size_t o_cnt = o_keys + nw * 8, total = o_cnt + 256;
    rc = ensure_buf(ctx, &ctx->lal_buf, &ctx->lal_cap, total);
    if (rc) return rc;
    uint8_t *base = (uint8_t *)ctx->lal_buf;
    int32_t *top8 = (int32_t *)(base + o_t8), *wmv = (int32_t *)(base + o_mv), *wfb = (int32_t *)(base + o_fb);
    int32_t *list = (int32_t *)(base + o_list);
    uint64_t *lkeys = (uint64_t *)(base + o_keys);
    unsigned *count = (unsigned *)(base + o_cnt);
    if (nframes > 1) {
        // proposals for every (frame, reference) pair: the network over the pairs (j + k, j)
        // for each distance k, or (la_scaled) over the pairs (j + 1, j) only
        const int scaled = ctx->la_scaled, cmax = P - 8 - local_range;
        for (int k = 1; k <= (scaled ? 1 : nrefs) && k < nframes; k++) {
            rc = rh264_mvnet_forward(ctx, net, frames + (size_t)k * frame_stride, frames, frame_stride, nframes - k,
                                     cw, ch, nullptr, top8 + (size_t)(k - 1) * nframes * nm * LS_NCAND);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, count, (size_t)1, 0u);
        const dim3 grid((unsigned)((nm + 256 / LS_NCAND - 1) / (256 / LS_NCAND)), (unsigned)((nframes - 1) * nrefs));
#define LC_LAUNCH(RL)                                                                                             \
    hipLaunchKernelGGL(k_la_cand<RL>, grid, dim3(256), 0, ctx->stream, half, mb_w, mb_h, nframes, nrefs, P, planes, \
                       pstride, top8, scaled, cmax, (int)fallback_sad, wmv, wfb, count, list, lkeys)
        switch (local_range) {
        case 1: LC_LAUNCH(1); break;
        case 2: LC_LAUNCH(2); break;
        case 3: LC_LAUNCH(3); break;
        default: LC_LAUNCH(4); break;
        }
#undef LC_LAUNCH
        LAUNCH_CHECK("k_la_cand");
        const long lanes = (long)nw * fs_groups<0>(2 * range + 1);
        hipLaunchKernelGGL(k_la_search_list, dim3((unsigned)std::min<long>((lanes + 255) / 256, 2048)), dim3(256), 0,
                           ctx->stream, half, mb_w, mb_h, nrefs, range, P, planes, pstride, list, count, lkeys);
        LAUNCH_CHECK("k_la_search_list");
        hipLaunchKernelGGL(k_la_fb_merge, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, ctx->stream, list, count,
                           lkeys, range, wmv);
        LAUNCH_CHECK("k_la_fb_merge");
    }
    const long nb = (long)nframes * nm;
    hipLaunchKernelGGL(k_la_cost, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx->stream, half, mb_w, mb_h,
                       nframes, nrefs, range, P, planes, pstride, (const uint64_t *)nullptr, wmv, wfb, intra, inter,
                       mv, ref, fallback);
    LAUNCH_CHECK("k_la_cost");
    return 0;
}


// ---------------------------------------------------------------------------
// batched multi-stream search: host side
// ---------------------------------------------------------------------------

static int ref_bits(int r, int nrefs) // mb16.rs:5967-5980
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

size_t rh264_batch_out_len(int cw, int ch, const int32_t *shapes, int nshapes)
{
    size_t n = 0;
    for (int s = 0; s < nshapes; s++) n += (size_t)4 * (cw / shapes[2 * s]) * (ch / shapes[2 * s + 1]);
    return n;
}

size_t rh264_batch_mvp_len(int cw, int ch, const int32_t *shapes, int nshapes)
{
    return rh264_batch_out_len(cw, ch, shapes, nshapes) / 2;
}

struct BatchShape {
    int bw, bh;
    int nseg, nsel;
    long lanes[3];     // integer-search lanes per variant
    long blocks;       // blocks over all segments
    long sel_blocks;   // blocks over all streams
    SegDesc *d_segs;
    SelDesc *d_sels;
    int variant, persist; // tuned kernel choice: variant 0..2, persist 0 = one tile per block, k = k*4 blocks per CU
    int fused;            // index among the fused shapes (16x16, 16x8, 8x16, 8x8, 4x4), -1: searched on its own
    size_t koff;          // first key of the shape in the fused search's key array
};

struct rh264_batch {
    int pad, range, pat, ncu;
    LrTab lr_fp, lr;
    std::vector<BatchShape> shapes;
    uint64_t *keys;
    int32_t *res;
    unsigned *counter;
    FuSeg *d_fu;          // fused search: one descriptor per (stream, reference), or null
    int fu_nseg, fu_mvp;
    long fu_units;
    size_t fu_keys;
    hipGraph_t graph;
    hipGraphExec_t exec;
    bool has_graph;
};

int rh264_batch_destroy(rh264_batch *b)
{
    if (!b) return 0;
    if (b->has_graph) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
    }
    for (auto &S : b->shapes) {
        if (S.d_segs) (void)hipFree(S.d_segs);
        if (S.d_sels) (void)hipFree(S.d_sels);
    }
    if (b->keys) (void)hipFree(b->keys);
    if (b->d_fu) (void)hipFree(b->d_fu);
    if (b->res) (void)hipFree(b->res);
    if (b->counter) (void)hipFree(b->counter);
    delete b;
    return 0;
}

int rh264_batch_create(rh264_ctx *ctx, const rh264_stream *streams, int nstreams, const int32_t *shapes, int nshapes,
                       int pad, int range, int subpel_pat, const int32_t *lr_fp, const int32_t *lr, double lambda_ref,
                       rh264_batch **out)
{
    if (!out || !streams || !shapes || !lr_fp || !lr) return fail("batch_create: null argument");
    *out = nullptr;
    if (nstreams <= 0 || nshapes <= 0 || nshapes > 8) return fail("batch_create: %d streams, %d shapes", nstreams, nshapes);
    if (range < 1 || range > 64 || subpel_pat < -1 || subpel_pat > 4 || pad < 4 || pad > 128 || pad % 4)
        return RH264_DECLINED;
    for (int s = 0; s < nshapes; s++)
        if (!shape_ok(shapes[2 * s], shapes[2 * s + 1])) return RH264_DECLINED;
    for (int i = 0; i < nstreams; i++) {
        const rh264_stream &st = streams[i];
        if (!geom_ok(st.cw, st.ch, pad) || st.nrefs < 1 || st.nrefs > 4) return RH264_DECLINED;
        if (!st.cur || !st.planes || !st.out) return fail("batch_create: stream %d has a null pointer", i);
    }
    for (int i = 0; i < RH264_LR_N; i++)
        if (lr[i] < 0 || lr_fp[i] < 0 || lr[i] > (1 << 28) || lr_fp[i] > (1 << 28))
            return fail("lambda table entry %d out of range", i);
    rh264_batch *b = new rh264_batch();
    b->pad = pad;
    b->range = range;
    b->pat = subpel_pat;
    b->keys = nullptr;
    b->res = nullptr;
    b->counter = nullptr;
    b->d_fu = nullptr;
    b->fu_nseg = 0;
    b->fu_mvp = 0;
    b->fu_units = 0;
    b->fu_keys = 0;
    b->has_graph = false;
    memcpy(b->lr_fp.v, lr_fp, sizeof(b->lr_fp.v));
    memcpy(b->lr.v, lr, sizeof(b->lr.v));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) {
        delete b;
        return fail("hipGetDeviceProperties failed");
    }
    b->ncu = prop.multiProcessorCount;
    const int NC = 2 * range + 1;
    long max_blocks = 1;
    size_t mvp_off = 0, out_off = 0;
    for (int s = 0; s < nshapes; s++) {
        BatchShape S;
        S.bw = shapes[2 * s];
        S.bh = shapes[2 * s + 1];
        S.d_segs = nullptr;
        S.d_sels = nullptr;
        S.variant = 0;
        S.persist = 0;
        S.fused = -1;
        S.koff = 0;
        std::vector<SegDesc> segs;
        std::vector<SelDesc> sels;
        long lane = 0, blk = 0, sel = 0;
        const int ng0 = fs_groups<0>(NC);
        for (int i = 0; i < nstreams; i++) {
            const rh264_stream &st = streams[i];
            const int nbx = st.cw / S.bw, nblk = nbx * (st.ch / S.bh);
            // offsets of this shape's records inside the stream's out / mvp arrays
            size_t o_off = 0, m_off = 0;
            for (int t = 0; t < s; t++) {
                const size_t nb = (size_t)(st.cw / shapes[2 * t]) * (st.ch / shapes[2 * t + 1]);
                o_off += 4 * nb;
                m_off += 2 * nb;
            }
            SelDesc sd;
            memset(&sd, 0, sizeof(sd));
            sd.sel0 = sel;
            sd.out = st.out + o_off;
            sd.nblk = nblk;
            sd.nrefs = st.nrefs;
            const size_t pbytes = rh264_planes_bytes(st.cw, st.ch, pad);
            for (int r = 0; r < st.nrefs; r++) {
                SegDesc g;
                g.cur = st.cur;
                g.planes = st.planes + (size_t)r * pbytes;
                g.mvps = st.mvps ? st.mvps + m_off : nullptr;
                g.cw = st.cw;
                g.ch = st.ch;
                g.nbx = nbx;
                g.nblk = nblk;
                g.lane0 = lane;
                g.blk0 = blk;
                sd.ref_blk0[r] = blk;
                sd.lr_ref[r] = (int32_t)(int64_t)(lambda_ref * (double)ref_bits(r, st.nrefs));
                segs.push_back(g);
                lane += (long)nblk * ng0;
                blk += nblk;
            }
            sels.push_back(sd);
            sel += nblk;
        }
        (void)mvp_off;
        (void)out_off;
        S.nseg = (int)segs.size();
        S.nsel = (int)sels.size();
        S.blocks = blk;
        S.sel_blocks = sel;
        // lane numbering differs per variant (lanes per block); store each variant's table
        // by recomputing lane0 on the device side from blocks -- keep the quad-SAD
        // numbering here and per-variant lane counts for the launch.
        for (int v = 0; v < 3; v++) S.lanes[v] = 0;
        const int ngv[3] = {fs_groups<0>(NC), fs_groups<1>(NC), fs_groups<2>(NC)};
        for (int v = 0; v < 3; v++) S.lanes[v] = blk * ngv[v];
        // the device tables: one SegDesc array per variant (lane0 depends on the variant)
        std::vector<SegDesc> all;
        for (int v = 0; v < 3; v++) {
            long l = 0;
            for (auto g : segs) {
                g.lane0 = l;
                l += (long)g.nblk * ngv[v];
                all.push_back(g);
            }
        }
        hipError_t e = hipMalloc(&S.d_segs, all.size() * sizeof(SegDesc));
        if (e == hipSuccess) e = hipMemcpy(S.d_segs, all.data(), all.size() * sizeof(SegDesc), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&S.d_sels, sels.size() * sizeof(SelDesc));
        if (e == hipSuccess) e = hipMemcpy(S.d_sels, sels.data(), sels.size() * sizeof(SelDesc), hipMemcpyHostToDevice);
        b->shapes.push_back(S);
        if (e != hipSuccess) {
            rh264_batch_destroy(b);
            return fail("batch_create: %s", hipGetErrorString(e));
        }
        max_blocks = blk > max_blocks ? blk : max_blocks;
    }
    // The macroblock shapes search in one fused pass when the 4x4 and at least one
    // other of them are in the batch, every frame is whole macroblocks and the rate
    // table keeps the fused keys' costs below 2^22 (k_fs_fused).
    const int fu_w[FU_NS] = {16, 16, 8, 8, 4}, fu_h[FU_NS] = {16, 8, 16, 8, 4};
    int fu_idx[FU_NS], nfu = 0;
    for (int f = 0; f < FU_NS; f++) {
        fu_idx[f] = -1;
        for (int s = 0; s < nshapes && fu_idx[f] < 0; s++)
            if (shapes[2 * s] == fu_w[f] && shapes[2 * s + 1] == fu_h[f]) fu_idx[f] = s;
        nfu += fu_idx[f] >= 0;
    }
    bool fuse = fu_idx[FU_NS - 1] >= 0 && nfu >= 2;
    for (int i = 0; i < nstreams && fuse; i++) fuse = streams[i].cw % 16 == 0 && streams[i].ch % 16 == 0;
    for (int i = 0; i < RH264_LR_N && fuse; i++) fuse = lr_fp[i] <= FU_LR_MAX;
    size_t fu_total = 0;
    if (fuse) {
        for (int f = 0; f < FU_NS; f++)
            if (fu_idx[f] >= 0) {
                BatchShape &S = b->shapes[fu_idx[f]];
                S.fused = f;
                S.koff = fu_total;
                fu_total += (size_t)S.blocks;
            }
        std::vector<FuSeg> fu;
        long unit = 0;
        int64_t blk[FU_NS] = {0, 0, 0, 0, 0};
        const int ng = (NC + 3) / 4;
        for (int i = 0; i < nstreams; i++) {
            const rh264_stream &st = streams[i];
            const size_t pbytes = rh264_planes_bytes(st.cw, st.ch, pad);
            const int32_t *mv[FU_NS];
            int nb[FU_NS];
            for (int f = 0; f < FU_NS; f++) {
                nb[f] = (st.cw / fu_w[f]) * (st.ch / fu_h[f]);
                mv[f] = nullptr;
                if (fu_idx[f] < 0 || !st.mvps) continue;
                size_t m_off = 0;
                for (int t = 0; t < fu_idx[f]; t++) m_off += 2 * (size_t)(st.cw / shapes[2 * t]) * (st.ch / shapes[2 * t + 1]);
                mv[f] = st.mvps + m_off;
                b->fu_mvp = 1;
            }
            for (int r = 0; r < st.nrefs; r++) {
                FuSeg g;
                memset(&g, 0, sizeof(g));
                g.cur = st.cur;
                g.planes = st.planes + (size_t)r * pbytes;
                g.unit0 = unit;
                g.cw = st.cw;
                g.ch = st.ch;
                g.mbw = st.cw / 16;
                g.pad = pad;
                for (int f = 0; f < FU_NS; f++) {
                    g.mvps[f] = mv[f];
                    g.blk0[f] = fu_idx[f] >= 0 ? blk[f] : -1;
                    blk[f] += nb[f];
                }
                fu.push_back(g);
                unit += (long)(st.cw / 16) * (st.ch / 16) * ng;
            }
        }
        b->fu_nseg = (int)fu.size();
        b->fu_units = unit;
        b->fu_keys = fu_total;
        hipError_t e = hipMalloc(&b->d_fu, fu.size() * sizeof(FuSeg));
        if (e == hipSuccess) e = hipMemcpy(b->d_fu, fu.data(), fu.size() * sizeof(FuSeg), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            rh264_batch_destroy(b);
            return fail("batch_create: %s", hipGetErrorString(e));
        }
        if ((long)fu_total > max_blocks) max_blocks = (long)fu_total;
    }
    hipError_t e = hipMalloc(&b->keys, (size_t)max_blocks * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&b->res, (size_t)max_blocks * 3 * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&b->counter, 64);
    if (e != hipSuccess) {
        rh264_batch_destroy(b);
        return fail("batch_create: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

} // extern "C"

template <int BW, int BH, int V>
static void launch_fs_multi_v(rh264_ctx *ctx, rh264_batch *b, const BatchShape &S)
{
    const long tiles = (S.lanes[V] + 255) / 256;
    unsigned grid = (unsigned)tiles;
    if (S.persist > 0) {
        const long cap = (long)b->ncu * 4 * S.persist;
        grid = (unsigned)(tiles < cap ? tiles : cap);
    }
    hipLaunchKernelGGL((k_fs_multi<BW, BH, V>), dim3(grid), dim3(256), 0, ctx->stream, S.d_segs + (size_t)V * S.nseg,
                       S.nseg, S.lanes[V], b->range, b->pad, b->lr_fp, b->keys, b->counter, S.persist);
}

// mode 0: integer search, refinement and reference choice; 1: the integer search only;
// 2: refinement and reference choice from the fused search's keys.
template <int BW, int BH>
static int batch_shape_run(rh264_ctx *ctx, rh264_batch *b, const BatchShape &S, int mode)
{
    const uint64_t *keys = mode == 2 ? b->keys + S.koff : b->keys;
    if (mode != 2) {
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((S.blocks + 255) / 256)), dim3(256), 0, ctx->stream, b->keys,
                           (size_t)S.blocks, ~0ull);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, b->counter, (size_t)16, 0u);
        if (S.variant == 1) launch_fs_multi_v<BW, BH, 1>(ctx, b, S);
        else if (S.variant == 2) launch_fs_multi_v<BW, BH, 2>(ctx, b, S);
        else launch_fs_multi_v<BW, BH, 0>(ctx, b, S);
        LAUNCH_CHECK("k_fs_multi");
    }
    if (mode == 1) return 0;
    constexpr int GPW = RefShape<BW, BH>::GPW;
    if (RefShape<BW, BH>::NSB >= RH264_LDS_MIN_NSB && ctx->refine_lds)
        hipLaunchKernelGGL((k_refine_multi<BW, BH, RefShape<BW, BH>::NSB >= RH264_LDS_MIN_NSB>),
                           dim3((unsigned)((S.blocks + GPW - 1) / GPW)), dim3(256), 0, ctx->stream, S.d_segs, S.nseg,
                           S.blocks, b->range, b->pad, b->pat, b->lr, keys, b->res);
    else
        hipLaunchKernelGGL((k_refine_multi<BW, BH, false>), dim3((unsigned)((S.blocks + GPW - 1) / GPW)), dim3(256), 0,
                           ctx->stream, S.d_segs, S.nseg, S.blocks, b->range, b->pad, b->pat, b->lr, keys, b->res);
    LAUNCH_CHECK("k_refine_multi");
    hipLaunchKernelGGL(k_select_ref, dim3((unsigned)((S.sel_blocks + 255) / 256)), dim3(256), 0, ctx->stream, S.d_sels,
                       S.nsel, S.sel_blocks, b->res);
    LAUNCH_CHECK("k_select_ref");
    return 0;
}

static int batch_shape_dispatch(rh264_ctx *ctx, rh264_batch *b, const BatchShape &S, int mode)
{
    switch (S.bw * 100 + S.bh) {
    case 1616: return batch_shape_run<16, 16>(ctx, b, S, mode);
    case 1608: return batch_shape_run<16, 8>(ctx, b, S, mode);
    case 816: return batch_shape_run<8, 16>(ctx, b, S, mode);
    case 808: return batch_shape_run<8, 8>(ctx, b, S, mode);
    case 804: return batch_shape_run<8, 4>(ctx, b, S, mode);
    case 408: return batch_shape_run<4, 8>(ctx, b, S, mode);
    case 404: return batch_shape_run<4, 4>(ctx, b, S, mode);
    default: return RH264_DECLINED;
    }
}

static int batch_fused_search(rh264_ctx *ctx, rh264_batch *b)
{
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((b->fu_keys + 255) / 256)), dim3(256), 0, ctx->stream, b->keys,
                       b->fu_keys, ~0ull);
    FuKeys k;
    for (int f = 0; f < FU_NS; f++) k.k[f] = nullptr;
    for (auto &S : b->shapes)
        if (S.fused >= 0) k.k[S.fused] = b->keys + S.koff;
    const unsigned grid = (unsigned)((b->fu_units + 15) / 16); // 16 units of 16 lanes per workgroup
    if (b->fu_mvp)
        hipLaunchKernelGGL(k_fs_fused<true>, dim3(grid), dim3(256), 0, ctx->stream, b->d_fu, b->fu_nseg, b->fu_units,
                           b->range, b->lr_fp, k);
    else
        hipLaunchKernelGGL(k_fs_fused<false>, dim3(grid), dim3(256), 0, ctx->stream, b->d_fu, b->fu_nseg, b->fu_units,
                           b->range, b->lr_fp, k);
    LAUNCH_CHECK("k_fs_fused");
    return 0;
}

extern "C" {

int rh264_batch_run(rh264_ctx *ctx, rh264_batch *b)
{
    if (!b) return fail("batch_run: null batch");
    if (b->d_fu) {
        int rc = batch_fused_search(ctx, b);
        if (rc) return rc;
        for (auto &S : b->shapes)
            if (S.fused >= 0 && (rc = batch_shape_dispatch(ctx, b, S, 2))) return rc;
    }
    for (auto &S : b->shapes) {
        if (S.fused >= 0) continue;
        int rc = batch_shape_dispatch(ctx, b, S, 0);
        if (rc) return rc;
    }
    return 0;
}

int rh264_batch_search(rh264_ctx *ctx, rh264_batch *b)
{
    if (!b) return fail("batch_search: null batch");
    if (b->d_fu) {
        int rc = batch_fused_search(ctx, b);
        if (rc) return rc;
    }
    for (auto &S : b->shapes) {
        if (S.fused >= 0) continue;
        int rc = batch_shape_dispatch(ctx, b, S, 1);
        if (rc) return rc;
    }
    return 0;
}

int rh264_batch_autotune(rh264_ctx *ctx, rh264_batch *b, int reps)
{
    if (!b) return fail("batch_autotune: null batch");
    if (b->has_graph) return fail("batch_autotune: the batch is already captured in a graph");
    if (reps < 1) reps = 1;
    const int persists[4] = {0, 1, 2, 4};
    for (auto &S : b->shapes) {
        if (S.fused >= 0) continue; // one fused launch, nothing to choose
        float best_ms = 1e30f;
        int bv = 0, bp = 0;
        for (int v = 0; v < 3; v++)
            for (int pi = 0; pi < 4; pi++) {
                S.variant = v;
                S.persist = persists[pi];
                int rc = batch_shape_dispatch(ctx, b, S, 1); // warm
                if (rc) return rc;
                HIP_TRY(hipEventRecord(ctx->ev[14], ctx->stream));
                for (int k = 0; k < reps; k++) {
                    rc = batch_shape_dispatch(ctx, b, S, 1);
                    if (rc) return rc;
                }
                HIP_TRY(hipEventRecord(ctx->ev[15], ctx->stream));
                HIP_TRY(hipEventSynchronize(ctx->ev[15]));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[14], ctx->ev[15]));
                if (ms < best_ms) {
                    best_ms = ms;
                    bv = v;
                    bp = persists[pi];
                }
            }
        S.variant = bv;
        S.persist = bp;
    }
    return 0;
}

int rh264_batch_choice(rh264_batch *b, int shape_index, int *variant, int *persist)
{
    if (!b || shape_index < 0 || shape_index >= (int)b->shapes.size()) return fail("batch_choice: bad index");
    const BatchShape &S = b->shapes[shape_index];
    *variant = S.fused >= 0 ? 3 : S.variant;
    *persist = S.fused >= 0 ? 0 : S.persist;
    return 0;
}

int rh264_batch_capture(rh264_ctx *ctx, rh264_batch *b)
{
    if (!b) return fail("batch_capture: null batch");
    if (b->has_graph) {
        HIP_TRY(hipGraphExecDestroy(b->exec));
        HIP_TRY(hipGraphDestroy(b->graph));
        b->has_graph = false;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = rh264_batch_run(ctx, b);
    hipGraph_t g;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (e == hipSuccess) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail("hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail("hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    b->graph = g;
    b->has_graph = true;
    return 0;
}

int rh264_batch_replay(rh264_ctx *ctx, rh264_batch *b)
{
    if (!b || !b->has_graph) return fail("batch_replay: no captured graph");
    HIP_TRY(hipGraphLaunch(b->exec, ctx->stream));
    return 0;
}


// ---------------------------------------------------------------------------
// learned-proposal search: host side
// ---------------------------------------------------------------------------

struct rh264_mvnet;
int rh264_mvnet_forward(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *cur, const uint8_t *ref, size_t stride,
                        int n, int w, int h, float *logits, int32_t *top8);

static int ensure_buf(rh264_ctx *ctx, void **buf, size_t *cap, size_t bytes)
{
    if (*cap >= bytes) return 0;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (*buf) HIP_TRY(hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
    HIP_TRY(hipMalloc(buf, bytes));
    *cap = bytes;
    return 0;
}

int rh264_learned_search(rh264_ctx *ctx, const rh264_search_params *p, rh264_mvnet *net, int nframes,
                         const uint8_t *cur, const uint8_t *ref, size_t frame_stride, const uint8_t *planes,
                         size_t planes_stride, int local_range, int32_t fallback_cost, int32_t *mv, int32_t *cost,
                         int32_t *fallback)
{
    int rc = check_params(p);
    if (rc) return rc;
    if (p->bw != 16 || p->bh != 16 || local_range < 0 || local_range > 8) return RH264_DECLINED;
    if (nframes <= 0) return 0;
    if (!net || !mv || !cost || !fallback) return fail("learned_search: null argument");
    rc = check_strides(frame_stride, planes_stride);
    if (rc) return rc;
    const long nb = (long)(p->cw / 16) * (p->ch / 16) * nframes;
    // scratch: top8 | centres | start | list | count | keys8 (| keys of the fall-back search)
    const size_t o_top = 0, o_cen = o_top + (size_t)nb * LS_NCAND * 4, o_start = o_cen + (size_t)nb * LS_NCAND * 8;
    const size_t o_list = o_start + (size_t)nb * 8, o_count = o_list + (size_t)nb * 4 + 256;
    const size_t o_keys = (o_count + 256 + 255) & ~(size_t)255, total = o_keys + (size_t)nb * (LS_NCAND + 1) * 8;
    rc = ensure_buf(ctx, &ctx->ls_buf, &ctx->ls_cap, total);
    if (rc) return rc;
    uint8_t *base = (uint8_t *)ctx->ls_buf;
    int32_t *top8 = (int32_t *)(base + o_top), *centres = (int32_t *)(base + o_cen), *start = (int32_t *)(base + o_start);
    int32_t *list = (int32_t *)(base + o_list);
    unsigned *count = (unsigned *)(base + o_count);
    uint64_t *keys8 = (uint64_t *)(base + o_keys), *keysf = keys8 + (size_t)nb * LS_NCAND;

    // phases: input preparation, network, candidate search, fall-back search, refinement
    hipEvent_t *ev = prof_slot(ctx, 5);
    if (ev) HIP_TRY(hipEventRecord(ev[0], ctx->stream));
    // the network's input preparation also clears the fall-back list's counter
    rc = mvnet_forward_ev(ctx, net, cur, ref, frame_stride, nframes, p->cw, p->ch, nullptr, top8,
                          ev && ctx->mvnet_fused ? ev[1] : nullptr, count);
    if (rc) return rc;
    if (ev && !ctx->mvnet_fused) HIP_TRY(hipEventRecord(ev[1], ctx->stream));
    if (ev) HIP_TRY(hipEventRecord(ev[2], ctx->stream));
    if (local_range >= 1 && local_range <= 4 && ctx->cand_quad == 2) {
        // one lane per candidate: quad-SAD column groups plus v_sad_u8 for the rest; the
        // kernel clears the fall-back blocks' argmin slots itself
#define CM_LAUNCH(RL)                                                                                              \
    do {                                                                                                           \
        const dim3 grid((unsigned)((nb + 256 / LS_NCAND - 1) / (256 / LS_NCAND)));                                 \
        if (ctx->cand_lds)                                                                                         \
            hipLaunchKernelGGL((k_cand_search_m<RL, true>), grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride,  \
                               planes, planes_stride, top8, (const int32_t *)nullptr, nb, fallback_cost, start,     \
                               fallback, count, list, keysf);                                                       \
        else                                                                                                       \
            hipLaunchKernelGGL((k_cand_search_m<RL, false>), grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride, \
                               planes, planes_stride, top8, (const int32_t *)nullptr, nb, fallback_cost, start,     \
                               fallback, count, list, keysf);                                                       \
    } while (0)
        switch (local_range) {
        case 1: CM_LAUNCH(1); break;
        case 2: CM_LAUNCH(2); break;
        case 3: CM_LAUNCH(3); break;
        default: CM_LAUNCH(4); break;
        }
#undef CM_LAUNCH
        LAUNCH_CHECK("k_cand_search_m");
    } else if (local_range <= 3 && ctx->cand_quad) {
        // one pass over all candidates with quad-SADs (same result as the kernels below)
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx->stream, keysf,
                           (size_t)nb, ~0ull);
#define CQ_LAUNCH(RL)                                                                                             \
    hipLaunchKernelGGL(k_cand_search_q<RL>, dim3((unsigned)((nb + 256 / (LS_NCAND * ((2 * RL + 4) / 4)) - 1) /        \
                                                          (256 / (LS_NCAND * ((2 * RL + 4) / 4))))),                  \
                       dim3(256), 0, ctx->stream, *p, cur, frame_stride, planes, planes_stride, top8,                 \
                       (const int32_t *)nullptr, nb, fallback_cost, start, fallback, count, list)
        switch (local_range) {
        case 0: CQ_LAUNCH(0); break;
        case 1: CQ_LAUNCH(1); break;
        case 2: CQ_LAUNCH(2); break;
        default: CQ_LAUNCH(3); break;
        }
#undef CQ_LAUNCH
        LAUNCH_CHECK("k_cand_search_q");
    } else if (local_range <= 4) {
        // one pass over all candidates (the per-candidate searches below give the same result)
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx->stream, keysf,
                           (size_t)nb, ~0ull);
        const int lpb = LS_NCAND * (2 * local_range + 1), bpw = 256 / lpb;
        const dim3 grid((unsigned)((nb + bpw - 1) / bpw));
        switch (local_range) {
        case 0: hipLaunchKernelGGL(k_cand_search<0>, grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride, planes,
                                   planes_stride, top8, (const int32_t *)nullptr, nb, fallback_cost, start, fallback,
                                   count, list); break;
        case 1: hipLaunchKernelGGL(k_cand_search<1>, grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride, planes,
                                   planes_stride, top8, (const int32_t *)nullptr, nb, fallback_cost, start, fallback,
                                   count, list); break;
        case 2: hipLaunchKernelGGL(k_cand_search<2>, grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride, planes,
                                   planes_stride, top8, (const int32_t *)nullptr, nb, fallback_cost, start, fallback,
                                   count, list); break;
        case 3: hipLaunchKernelGGL(k_cand_search<3>, grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride, planes,
                                   planes_stride, top8, (const int32_t *)nullptr, nb, fallback_cost, start, fallback,
                                   count, list); break;
        default: hipLaunchKernelGGL(k_cand_search<4>, grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride, planes,
                                    planes_stride, top8, (const int32_t *)nullptr, nb, fallback_cost, start, fallback,
                                    count, list); break;
        }
        LAUNCH_CHECK("k_cand_search");
    } else {
        hipLaunchKernelGGL(k_cand_centres, dim3((unsigned)((nb * LS_NCAND + 255) / 256)), dim3(256), 0, ctx->stream,
                           top8, nb, centres);
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nb * (LS_NCAND + 1) + 255) / 256)), dim3(256), 0,
                           ctx->stream, keys8, (size_t)nb * (LS_NCAND + 1), ~0ull);
        LAUNCH_CHECK("learned_search setup");
        rh264_search_params q = *p;
        q.range = local_range;
        for (int k = 0; k < LS_NCAND; k++) {
            rc = launch_full<16, 16>(ctx, &q, nframes, cur, frame_stride, planes, planes_stride,
                                     centres + 2 * (size_t)k * nb, nullptr, keys8 + (size_t)k * nb);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_cand_select, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx->stream, keys8,
                           centres, nb, local_range, fallback_cost, start, fallback, count, list);
    }
    if (ev) HIP_TRY(hipEventRecord(ev[3], ctx->stream));
    // the window refinement reads the fall-back blocks' starts from their keys itself;
    // the other refinements take them merged into `start`
    const bool wn = p->subpel_pat == 4 && (ctx->refine_mfma || p->hadamard == 8);
    // with the window refinement, the fall-back search runs on the side stream beside
    // the refinement of the other blocks; its own blocks are refined after the join
    const bool ovl = wn && ctx->fallback_overlap;
    hipStream_t fs = ctx->stream;
    if (ovl) {
        HIP_TRY(hipEventRecord(ctx->fork_ev, ctx->stream));
        HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
        fs = ctx->side;
    }
    // the fall-back list is short (a few % of the blocks): the v_sad_u8 variant with two
    // candidates per lane gives it twice the lanes of the quad-SAD one
    {
        const int V = ctx->fallback_v == 2 ? 3 : ctx->fallback_v;
        const long lanes = nb * (V == 3 ? fs_groups<3>(2 * p->range + 1)
                                        : V == 1 ? fs_groups<1>(2 * p->range + 1) : fs_groups<0>(2 * p->range + 1));
        // at most fallback_wgs workgroups (8 per CU by default), which loop when the list
        // is longer than they cover
        const dim3 grid((unsigned)std::min<long>((lanes + 255) / 256, ctx->fallback_wgs));
        if (V == 3)
            hipLaunchKernelGGL((k_full_search_list<16, 16, 3>), grid, dim3(256), 0, fs, *p, cur, frame_stride,
                               planes, planes_stride, list, count, (const int32_t *)nullptr, keysf);
        else if (V == 1)
            hipLaunchKernelGGL((k_full_search_list<16, 16, 1>), grid, dim3(256), 0, fs, *p, cur, frame_stride,
                               planes, planes_stride, list, count, (const int32_t *)nullptr, keysf);
        else
            hipLaunchKernelGGL((k_full_search_list<16, 16, 0>), grid, dim3(256), 0, fs, *p, cur, frame_stride,
                               planes, planes_stride, list, count, (const int32_t *)nullptr, keysf);
    }
    LAUNCH_CHECK("k_full_search_list");
    if (ovl) {
        if (ev) HIP_TRY(hipEventRecord(ev[4], ctx->side));
        HIP_TRY(hipEventRecord(ctx->join_ev, ctx->side));
        rc = launch_refine<16, 16>(ctx, p, nframes, cur, frame_stride, planes, planes_stride, nullptr, nullptr, start,
                                   nullptr, nullptr, nullptr, mv, cost, nullptr, fallback, keysf, 1);
        if (rc) return rc;
        HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0));
        const dim3 grid((unsigned)std::min<long>((nb + 3) / 4, ctx->fallback_wgs));
        if (p->hadamard == 8)
            hipLaunchKernelGGL(k_refine_wn16_list<true>, grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride, planes,
                               planes_stride, list, count, keysf, mv, cost);
        else
            hipLaunchKernelGGL(k_refine_wn16_list<false>, grid, dim3(256), 0, ctx->stream, *p, cur, frame_stride,
                               planes, planes_stride, list, count, keysf, mv, cost);
        LAUNCH_CHECK("k_refine_wn16_list");
        if (ev) HIP_TRY(hipEventRecord(ev[5], ctx->stream));
        return 0;
    }
    if (!wn) {
        hipLaunchKernelGGL(k_fallback_merge, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx->stream, list,
                           count, keysf, p->range, start);
        LAUNCH_CHECK("learned_search select");
    }
    if (ev) HIP_TRY(hipEventRecord(ev[4], ctx->stream));
    rc = launch_refine<16, 16>(ctx, p, nframes, cur, frame_stride, planes, planes_stride, nullptr, nullptr, start,
                               nullptr, nullptr, nullptr, mv, cost, nullptr, wn ? fallback : nullptr,
                               wn ? keysf : nullptr);
    if (rc) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[5], ctx->stream));
    return 0;
}

} // extern "C"
