// rh264_mvnet.hip -- the learned motion-vector proposal network: a small BF16 CNN
// whose 3x3 convolutions run on MFMA (v_mfma_f32_16x16x16_bf16, FP32
// accumulation), with hand-written forward, backward and Adam kernels.
//
// Network (no reference counterpart; the reference has no learned component):
//   input  : current and reference luma at 1/4 resolution ((sum of 4x4) + 8) >> 4,
//            normalised (v - 128) / 64, 2 channels
//   L1     : conv3x3 stride 2, 2 -> 32, bias, ReLU          (1/8 resolution)
//   L2..L4 : conv3x3 stride 1, 32 -> 32, bias, ReLU          (1/8 resolution)
//   pool   : 2x2 average                                     (one cell per 16x16 MB)
//   head   : 1x1, 32 -> 289 logits = MV bins of 4 pels, dx, dy in -32..32 pels
// Activations are NHWC BF16; parameters FP32 masters with BF16 copies.
//
// Flat parameter layout (FP32): W1[32][9][2], b1[32], W2[32][9][32], b2[32],
// W3[32][9][32], b3[32], W4[32][9][32], b4[32], Wh[289][32], bh[289].

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "rh264_internal.h"

namespace {

constexpr int C = 32;     // channels
constexpr int K = 289;    // MV bins (17 x 17)
constexpr int NB = 17;    // bins per axis
constexpr int TOPK = 8;

constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + C * 9 * 2;
constexpr int OFF_W2 = OFF_B1 + C, OFF_B2 = OFF_W2 + C * 9 * C;
constexpr int OFF_W3 = OFF_B2 + C, OFF_B3 = OFF_W3 + C * 9 * C;
constexpr int OFF_W4 = OFF_B3 + C, OFF_B4 = OFF_W4 + C * 9 * C;
constexpr int OFF_WH = OFF_B4 + C, OFF_BH = OFF_WH + K * C;
constexpr int NPARAM = OFF_BH + K;

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (v_cvt_pk_bf16_f32)
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Two floats as packed BF16 (lo, hi), one conversion instruction.
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi)
{
    const bf16x2_t v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}

__global__ void k_fill_u32_nv(unsigned *p)
{
    if (threadIdx.x == 0) *p = 0u;
}

// ---------------------------------------------------------------------------
// input preparation: 1/4-resolution current + reference, 2 channels
// ---------------------------------------------------------------------------
// zero (may be null): a counter of the caller's next kernels, cleared here to save
// them a launch
__global__ void k_prep(const uint8_t *__restrict__ cur, const uint8_t *__restrict__ ref, size_t stride, int n, int w,
                       int h, uint16_t *__restrict__ x0, unsigned *__restrict__ zero)
{
    const int w4 = w / 4, h4 = h / 4;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (zero && i == 0) *zero = 0u;
    if (i >= (long)n * w4 * h4) return;
    const int img = (int)(i / ((long)w4 * h4));
    const int rem = (int)(i - (long)img * w4 * h4);
    const int y = rem / w4, x = rem - y * w4;
    const uint8_t *pc = cur + (size_t)img * stride + (size_t)(4 * y) * w + 4 * x;
    const uint8_t *pr = ref + (size_t)img * stride + (size_t)(4 * y) * w + 4 * x;
    uint32_t sc = 0, sr = 0;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            sc += pc[r * w + c];
            sr += pr[r * w + c];
        }
    x0[2 * i] = f2bf(((float)((sc + 8) >> 4) - 128.0f) / 64.0f);
    x0[2 * i + 1] = f2bf(((float)((sr + 8) >> 4) - 128.0f) / 64.0f);
}

// ---------------------------------------------------------------------------
// L1: conv3x3 stride 2, 2 -> 32 (VALU; 18 MACs per output channel)
// ---------------------------------------------------------------------------
__global__ void k_conv1_fwd(const uint16_t *__restrict__ x0, const float *__restrict__ p, int n, int h4, int w4,
                            uint16_t *__restrict__ a1)
{
    __shared__ float sw[C * 18 + C];
    // BF16-rounded weights (the biases stay FP32), as in the fused kernel
    for (int i = threadIdx.x; i < C * 18 + C; i += blockDim.x) sw[i] = i < C * 18 ? bf2f(f2bf(p[OFF_W1 + i])) : p[OFF_W1 + i];
    __syncthreads();
    const int h8 = h4 / 2, w8 = w4 / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * h8 * w8) return;
    const int img = (int)(i / ((long)h8 * w8));
    const int rem = (int)(i - (long)img * h8 * w8);
    const int y = rem / w8, x = rem - y * w8;
    float in[18];
    for (int t = 0; t < 9; t++) {
        const int yy = 2 * y + t / 3 - 1, xx = 2 * x + t % 3 - 1;
        const bool ok = yy >= 0 && yy < h4 && xx >= 0 && xx < w4;
        const size_t o = (((size_t)img * h4 + (ok ? yy : 0)) * w4 + (ok ? xx : 0)) * 2;
        in[2 * t] = ok ? bf2f(x0[o]) : 0.0f;
        in[2 * t + 1] = ok ? bf2f(x0[o + 1]) : 0.0f;
    }
    uint16_t *o = a1 + (size_t)i * C;
    for (int co = 0; co < C; co++) {
        float acc = sw[C * 18 + co];
        for (int k = 0; k < 18; k++) acc += sw[co * 18 + k] * in[k];
        o[co] = f2bf(acc > 0.0f ? acc : 0.0f);
    }
}

// ---------------------------------------------------------------------------
// conv3x3 stride 1, 32 -> 32 on MFMA.  out[p][co] = act(sum_{tap,ci} w[tap][co][ci] *
// in[p + tap][ci] + bias[co]); one wave makes 16 consecutive pixels of a row x 32
// output channels: D(co x pixel) = W(co x k) * X(k x pixel), k = (tap, ci), 18
// K-steps of 16.  Operand layout of v_mfma_f32_16x16x16_bf16: lane l supplies
// A[l%16][4(l/16)+i], B[4(l/16)+i][l%16] and receives D[4(l/16)+i][l%16].
// relu: clamp at 0; mask (may be null): zero where mask[p][co] == 0 (the ReLU
// derivative of the layer that produced mask).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_conv3_mfma(const uint16_t *__restrict__ in, const uint16_t *__restrict__ wt,
                                                    const float *__restrict__ bias, int relu,
                                                    const uint16_t *__restrict__ mask, int n, int H, int W,
                                                    uint16_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int wtiles = (W + 15) / 16;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= (long)n * H * wtiles) return;
    const int img = (int)(wave / ((long)H * wtiles));
    const int rem = (int)(wave - (long)img * H * wtiles);
    const int y = rem / wtiles, x0 = (rem - y * wtiles) * 16;
    const int col = lane & 15, kq = lane >> 4; // kq: which 4 of the 16 k of a step
    float4_t acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        const int yy = y + tap / 3 - 1, xx = x0 + col + tap % 3 - 1;
        const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
        const uint16_t *ip = in + (((size_t)img * H + (ok ? yy : 0)) * W + (ok ? xx : 0)) * C;
#pragma unroll
        for (int half = 0; half < 2; half++) {
            const int ci = half * 16 + 4 * kq;
            short4_t b = ok ? *(const short4_t *)(ip + ci) : short4_t{0, 0, 0, 0};
            short4_t a0 = *(const short4_t *)(wt + ((size_t)tap * C + col) * C + ci);
            short4_t a1 = *(const short4_t *)(wt + ((size_t)tap * C + col + 16) * C + ci);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b, acc1, 0, 0, 0);
        }
    }
    const int px = x0 + col;
    if (px >= W) return;
    const size_t po = (((size_t)img * H + y) * W + px) * C;
    uint16_t r0[4], r1[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int co = 4 * kq + i;
        float v0 = acc0[i] + (bias ? bias[co] : 0.0f), v1 = acc1[i] + (bias ? bias[co + 16] : 0.0f);
        if (relu) {
            v0 = v0 > 0.0f ? v0 : 0.0f;
            v1 = v1 > 0.0f ? v1 : 0.0f;
        }
        if (mask) {
            if (mask[po + co] == 0) v0 = 0.0f;
            if (mask[po + co + 16] == 0) v1 = 0.0f;
        }
        r0[i] = f2bf(v0);
        r1[i] = f2bf(v1);
    }
    *(short4_t *)(out + po + 4 * kq) = short4_t{(short)r0[0], (short)r0[1], (short)r0[2], (short)r0[3]};
    *(short4_t *)(out + po + 16 + 4 * kq) = short4_t{(short)r1[0], (short)r1[1], (short)r1[2], (short)r1[3]};
}

// 2x2 average pool (fp32 out) and its backward (grad / 4, masked by the ReLU of a4).
__global__ void k_pool_fwd(const uint16_t *__restrict__ a, int n, int H, int W, float *__restrict__ f)
{
    const int Hm = H / 2, Wm = W / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * Hm * Wm * C) return;
    const int c = (int)(i % C);
    const long cell = i / C;
    const int img = (int)(cell / ((long)Hm * Wm));
    const int rem = (int)(cell - (long)img * Hm * Wm);
    const int y = rem / Wm, x = rem - y * Wm;
    const size_t b = (((size_t)img * H + 2 * y) * W + 2 * x) * C + c;
    f[i] = 0.25f * (bf2f(a[b]) + bf2f(a[b + C]) + bf2f(a[b + (size_t)W * C]) + bf2f(a[b + (size_t)W * C + C]));
}

__global__ void k_pool_bwd(const float *__restrict__ df, const uint16_t *__restrict__ a, int n, int H, int W,
                           uint16_t *__restrict__ dz)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * H * W * C) return;
    const int c = (int)(i % C);
    const long px = i / C;
    const int img = (int)(px / ((long)H * W));
    const int rem = (int)(px - (long)img * H * W);
    const int y = rem / W, x = rem - y * W;
    const int Hm = H / 2, Wm = W / 2;
    float g = 0.0f;
    if (y / 2 < Hm && x / 2 < Wm) g = 0.25f * df[(((size_t)img * Hm + y / 2) * Wm + x / 2) * C + c];
    dz[i] = f2bf(a[i] != 0 ? g : 0.0f);
}

// head: logits[cell][k] = sum_c bf16(Wh[k][c]) bf16(f[cell][c]) + bh[k]
__global__ void k_head_fwd(const float *__restrict__ f, const float *__restrict__ p, long ncell,
                           float *__restrict__ logits)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell * K) return;
    const long cell = i / K;
    const int k = (int)(i - cell * K);
    const float *w = p + OFF_WH + (size_t)k * C;
    const float *fc = f + (size_t)cell * C;
    float acc = 0.0f;
    for (int c = 0; c < C; c++) acc += bf2f(f2bf(w[c])) * bf2f(f2bf(fc[c])); // BF16 operands, as on MFMA
    logits[i] = acc + p[OFF_BH + k];
}

// softmax cross-entropy per cell (one wave per cell): dlogits = (softmax - onehot)
// * scale, loss += -log softmax[label] * scale.
__global__ void k_softmax_ce(const float *__restrict__ logits, const int32_t *__restrict__ labels, long ncell,
                             float scale, float *__restrict__ dlogits, float *__restrict__ loss)
{
    const long cell = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (cell >= ncell) return;
    const float *z = logits + (size_t)cell * K;
    float m = -1e30f;
    for (int k = lane; k < K; k += 64) m = fmaxf(m, z[k]);
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    float se = 0.0f;
    for (int k = lane; k < K; k += 64) se += expf(z[k] - m);
    for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off, 64);
    const int lab = labels[cell];
    for (int k = lane; k < K; k += 64) {
        const float pk = expf(z[k] - m) / se;
        dlogits[(size_t)cell * K + k] = (pk - (k == lab ? 1.0f : 0.0f)) * scale;
    }
    if (lane == 0) atomicAdd(loss, (logf(se) + m - z[lab]) * scale);
}

// head backward: df[cell][c] = sum_k Wh[k][c] dlogits[cell][k]
__global__ void k_head_bwd_data(const float *__restrict__ dl, const float *__restrict__ p, long ncell,
                                float *__restrict__ df)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell * C) return;
    const long cell = i / C;
    const int c = (int)(i - cell * C);
    float acc = 0.0f;
    for (int k = 0; k < K; k++) acc += p[OFF_WH + (size_t)k * C + c] * dl[(size_t)cell * K + k];
    df[i] = acc;
}

// head weight grads: one thread per (k, c) or bias k, summing over a chunk of cells
__global__ void k_head_bwd_w(const float *__restrict__ dl, const float *__restrict__ f, long ncell,
                             float *__restrict__ g)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x; // 0 .. K*C + K - 1
    if (j >= K * C + K) return;
    const long c0 = (long)blockIdx.y * 256, c1 = c0 + 256 < ncell ? c0 + 256 : ncell;
    float acc = 0.0f;
    if (j < K * C) {
        const int k = j / C, c = j - k * C;
        for (long cell = c0; cell < c1; cell++) acc += dl[(size_t)cell * K + k] * f[(size_t)cell * C + c];
    } else {
        const int k = j - K * C;
        for (long cell = c0; cell < c1; cell++) acc += dl[(size_t)cell * K + k];
    }
    atomicAdd(&g[OFF_WH + j], acc);
}

// 3x3 weight gradients (VALU, FP32): dw[co][tap][ci] += sum_p dz[p][co] * x[in(p, tap)][ci],
// db[co] += sum_p dz[p][co].  One block per chunk of output pixels stages dz and the
// nine input taps in LDS (32 pixels per block), then each thread owns weights
// j = tid, tid + 256, ...
template <int CIN, int STRIDE>
__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dz,
                                                    int n, int Hin, int Win, int Hout, int Wout,
                                                    float *__restrict__ gw, float *__restrict__ gb)
{
    constexpr int P = 32;
    __shared__ float s_dz[P][C];
    __shared__ float s_x[P][9][CIN];
    const long npx = (long)n * Hout * Wout;
    const long p0 = (long)blockIdx.x * P;
    for (int i = threadIdx.x; i < P * C; i += 256) {
        const int pp = i / C, c = i - pp * C;
        const long pi = p0 + pp;
        s_dz[pp][c] = pi < npx ? bf2f(dz[(size_t)pi * C + c]) : 0.0f;
    }
    for (int i = threadIdx.x; i < P * 9 * CIN; i += 256) {
        const int pp = i / (9 * CIN), rem = i - pp * 9 * CIN, t = rem / CIN, c = rem - t * CIN;
        const long pi = p0 + pp;
        float v = 0.0f;
        if (pi < npx) {
            const int img = (int)(pi / ((long)Hout * Wout));
            const int r2 = (int)(pi - (long)img * Hout * Wout);
            const int yo = r2 / Wout, xo = r2 - yo * Wout;
            const int yy = yo * STRIDE + t / 3 - 1, xx = xo * STRIDE + t % 3 - 1;
            if (yy >= 0 && yy < Hin && xx >= 0 && xx < Win) v = bf2f(x[(((size_t)img * Hin + yy) * Win + xx) * CIN + c]);
        }
        s_x[pp][t][c] = v;
    }
    __syncthreads();
    const int NW = C * 9 * CIN;
    const int woff = CIN == 2 ? OFF_W1 : 0; // caller offsets gw for the 32-channel layers
    for (int j = threadIdx.x; j < NW + C; j += 256) {
        float acc = 0.0f;
        if (j < NW) {
            const int co = j / (9 * CIN), rem = j - co * 9 * CIN, t = rem / CIN, c = rem - t * CIN;
            for (int pp = 0; pp < P; pp++) acc += s_dz[pp][co] * s_x[pp][t][c];
            atomicAdd(&gw[woff + j], acc);
        } else {
            const int co = j - NW;
            for (int pp = 0; pp < P; pp++) acc += s_dz[pp][co];
            atomicAdd(&gb[co], acc);
        }
    }
}

// Adam step over all parameters, then the BF16 copies: forward [tap][co][ci] and,
// for the data gradient, the flipped transpose [8-tap][ci][co].
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

constexpr int WF_HEAD = 3 * 9 * C * C; // BF16 head weights [K][C] follow the conv layers in wf
constexpr int WF_COUNT = WF_HEAD + K * C;

__global__ void k_refresh_bf16(const float *__restrict__ p, uint16_t *__restrict__ wf, uint16_t *__restrict__ wb)
{
    // 3 layers x [32 co][9 tap][32 ci] -> wf[layer][tap][co][ci], wb[layer][8-tap][ci][co];
    // then the head [289][32] -> wf[WF_HEAD + k * 32 + c]
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= WF_COUNT) return;
    if (i >= WF_HEAD) {
        wf[i] = f2bf(p[OFF_WH + (i - WF_HEAD)]);
        return;
    }
    const int l = i / (C * 9 * C), j = i - l * C * 9 * C;
    const int co = j / (9 * C), rem = j - co * 9 * C, t = rem / C, ci = rem - t * C;
    const float w = p[OFF_W2 + l * (C * 9 * C + C) + j];
    const uint16_t b = f2bf(w);
    wf[(size_t)l * 9 * C * C + ((size_t)t * C + co) * C + ci] = b;
    wb[(size_t)l * 9 * C * C + ((size_t)(8 - t) * C + ci) * C + co] = b;
}

// Register-order BF16 weights of the register-resident inference kernel: one record
// of 64 lanes x 8 BF16 per MFMA A operand, read as one 16-byte load per lane.
// Records: L1 (h = 0, 1), then layer l, tap t, half h at 2 + (l * 9 + t) * 2 + h, then
// the head's bin tiles bt = 0..18.  Lane (q, col), slot s holds output channel (or bin)
// 16h + col against input channel ci(q, s) = s < 4 ? 4q + s : 16 + 4q + s - 4, the
// channel that slot s of lane q carries in a 16x16 MFMA result (rows 4q..4q+3 of the
// two 16-channel halves); L1's K slots are (tap 4q + s / 2, channel s % 2), taps >= 9
// zero.
constexpr int WR_REC = 64 * 8;                 // BF16 per record
constexpr int WR_LAYER = 2, WR_HEADREC = WR_LAYER + 3 * 9 * 2;
constexpr int WR_COUNT = (WR_HEADREC + (K + 15) / 16) * WR_REC;

__global__ void k_refresh_rr(const float *__restrict__ p, uint16_t *__restrict__ wr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= WR_COUNT) return;
    const int s = i & 7, lane = (i >> 3) & 63, rec = i / WR_REC;
    const int q = lane >> 4, col = lane & 15;
    const int ci = s < 4 ? 4 * q + s : 16 + 4 * q + (s - 4);
    float v;
    if (rec < WR_LAYER) {
        const int co = 16 * rec + col, t = 4 * q + (s >> 1);
        v = t < 9 ? p[OFF_W1 + co * 18 + 2 * t + (s & 1)] : 0.0f;
    } else if (rec < WR_HEADREC) {
        const int r = rec - WR_LAYER, l = r / 18, t = (r >> 1) % 9, co = 16 * (r & 1) + col;
        v = p[OFF_W2 + l * (C * 9 * C + C) + (co * 9 + t) * C + ci];
    } else {
        const int bin = 16 * (rec - WR_HEADREC) + col;
        v = bin < K ? p[OFF_WH + bin * C + ci] : 0.0f;
    }
    wr[i] = f2bf(v);
}

// Ranking of the proposals: each logit z of bin b gets the 32-bit key
// (order(z) & ~511) | (511 - b), order() = the order-preserving unsigned image of
// the FP32 bits.  The top-8 are the 8 largest keys, largest first: by logit with
// its 9 lowest order bits dropped, ties to the lower bin.  One wave per cell.
__device__ __forceinline__ uint32_t rank_key(float z, int bin)
{
    uint32_t u = __float_as_uint(z);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return (u & ~511u) | (uint32_t)(511 - bin);
}

__global__ void k_topk(const float *__restrict__ logits, long ncell, int32_t *__restrict__ top)
{
    const long cell = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (cell >= ncell) return;
    const float *z = logits + (size_t)cell * K;
    uint32_t v[5];
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const int k = lane + 64 * j;
        v[j] = k < K ? rank_key(z[k], k) : 0u;
    }
    for (int r = 0; r < TOPK; r++) {
        uint32_t m = v[0];
#pragma unroll
        for (int j = 1; j < 5; j++) m = max(m, v[j]);
        for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
        if (lane == 0) top[(size_t)cell * TOPK + r] = 511 - (int)(m & 511u);
#pragma unroll
        for (int j = 0; j < 5; j++)
            if (v[j] == m) v[j] = 0u;
    }
}

// ---------------------------------------------------------------------------
// fused inference: one workgroup per tile of 8x8 cells (16x16 pixels at 1/8
// resolution).  The tile's x0 window, a1..a4 (with the halo each following 3x3
// layer needs: 22, 20, 18, 16 pixels square) and the pooled features stay in
// LDS; the convolutions and the head run on v_mfma_f32_16x16x32_bf16 (for a 3x3
// layer one K step = one tap, all 32 input channels), the top-8 selection in
// registers.  Same arithmetic as the layer-by-layer kernels: BF16 activations,
// weights and features, FP32 accumulation, zero padding at every layer's border.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int FW = 4;                                             // waves per fused workgroup (8 measured slower)
constexpr int FT = 8;                                             // cells per tile side
constexpr int S4 = 2 * FT, S3 = S4 + 2, S2 = S4 + 4, S1 = S4 + 6; // a4..a1 window sides
constexpr int SX = 2 * S1 + 1;                                    // x0 window side (1/4 res)
constexpr int PSB = 80;                                           // LDS bytes per pixel (64 + 16 pad)
constexpr int LDS_A = S1 * S1 * PSB, LDS_B = S2 * S2 * PSB;       // a1/a3 and a2/a4
constexpr int LDS_X = FT * FT * PSB;                              // pooled features (BF16, PSB per cell)
constexpr int HSB = 80;                                           // LDS bytes per head-weight row (64 + 16 pad)
static_assert(K * HSB + K * 4 <= LDS_A, "the head's weights and bias fit the a1/a3 buffer");
static_assert(SX * SX * 4 <= LDS_B, "the x0 window fits the a2/a4 buffer");

__device__ __forceinline__ float4_t mfma_bf16_k32(u32x4 a, u32x4 b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}

// Epilogue of one 16-pixel tile of a conv layer: bias, ReLU, zero outside the
// image, BF16, four channels (8 bytes) per lane and half.
__device__ __forceinline__ void conv_store(uint8_t *out, int pc, bool valid, bool inside, const float4_t acc[2],
                                           const float bias[2][4], int q)
{
    if (!valid) return;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        uint32_t w[2];
#pragma unroll
        for (int i = 0; i < 4; i += 2) {
            float v0 = acc[h][i] + bias[h][i], v1 = acc[h][i + 1] + bias[h][i + 1];
            v0 = inside && v0 > 0.0f ? v0 : 0.0f;
            v1 = inside && v1 > 0.0f ? v1 : 0.0f;
            w[i / 2] = f2bf2(v0, v1);
        }
        *(uint2 *)(out + pc * PSB + (16 * h + 4 * q) * 2) = make_uint2(w[0], w[1]);
    }
}

__device__ __forceinline__ void load_bias(const float *__restrict__ bias, int q, float b[2][4])
{
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int i = 0; i < 4; i++) b[h][i] = bias[16 * h + 4 * q + i];
}

// 3x3 conv, 32 -> 32, from window `in` (side SIN) to window `out` (side SOUT =
// SIN - 2) whose origin lies `m` pixels up-left of the tile origin (oy, ox).  A
// wave takes two 16-pixel tiles at a time (four independent accumulators).
// The A operands (weights) of one 3x3 32 -> 32 layer for this lane.
__device__ __forceinline__ void load_w32(const uint16_t *__restrict__ wt, int lane, u32x4 (&wa)[9][2])
{
    const int q = lane >> 4, col = lane & 15;
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int h = 0; h < 2; h++) wa[t][h] = *(const u32x4 *)(wt + ((size_t)t * C + 16 * h + col) * C + 8 * q);
}

// wa, b: the layer's weights and bias, loaded by the caller ahead of the barrier
// before the layer so that their latency overlaps the other waves' work.
#ifndef RH264_CONV_U
#define RH264_CONV_U 2
#endif
template <int SIN, int U = RH264_CONV_U>
__device__ __forceinline__ void conv32_window(const uint8_t *in, uint8_t *out, const u32x4 (&wa)[9][2],
                                              const float (&b)[2][4], int oy, int ox, int m, int h8, int w8,
                                              int lane, int wave)
{
    // a wave takes U 16-pixel tiles at a time (2U independent accumulators)
    constexpr int SOUT = SIN - 2, NPX = SOUT * SOUT, NT = (NPX + 15) / 16;
    const int q = lane >> 4, col = lane & 15;
    for (int tile = wave; tile < NT; tile += FW * U) {
        int pc[U], y[U], x[U];
        bool valid[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int px = (tile + FW * u) * 16 + col;
            valid[u] = px < NPX;
            pc[u] = valid[u] ? px : NPX - 1;
            y[u] = pc[u] / SOUT;
            x[u] = pc[u] - y[u] * SOUT;
        }
        float4_t acc[U][2];
#pragma unroll
        for (int u = 0; u < U; u++) acc[u][0] = acc[u][1] = float4_t{0, 0, 0, 0};
        // B operands row by row of the 3x3 kernel: the next kernel row's loads are in
        // flight while the current row's products run
        const uint8_t *base[U];
#pragma unroll
        for (int u = 0; u < U; u++) base[u] = in + (y[u] * SIN + x[u]) * PSB + 16 * q;
        u32x4 bv[2][3][U];
#pragma unroll
        for (int tx = 0; tx < 3; tx++)
#pragma unroll
            for (int u = 0; u < U; u++) bv[0][tx][u] = *(const u32x4 *)(base[u] + tx * PSB);
#pragma unroll
        for (int ty = 0; ty < 3; ty++) {
            if (ty < 2) {
#pragma unroll
                for (int tx = 0; tx < 3; tx++)
#pragma unroll
                    for (int u = 0; u < U; u++)
                        bv[(ty + 1) & 1][tx][u] = *(const u32x4 *)(base[u] + ((ty + 1) * SIN + tx) * PSB);
            }
#pragma unroll
            for (int tx = 0; tx < 3; tx++)
#pragma unroll
                for (int u = 0; u < U; u++) {
                    acc[u][0] = mfma_bf16_k32(wa[3 * ty + tx][0], bv[ty & 1][tx][u], acc[u][0]);
                    acc[u][1] = mfma_bf16_k32(wa[3 * ty + tx][1], bv[ty & 1][tx][u], acc[u][1]);
                }
        }
        // keep that order: loads of rows 0 and 1, products of row 0, loads of row 2,
        // products of rows 1 and 2 (mask 0x100 = LDS reads, 0x008 = MFMA)
        __builtin_amdgcn_sched_group_barrier(0x100, 3 * U, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 3 * U, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 6 * U, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 3 * U, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 6 * U, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 6 * U, 0);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int gy = oy - m + y[u], gx = ox - m + x[u];
            conv_store(out, pc[u], valid[u], gy >= 0 && gy < h8 && gx >= 0 && gx < w8, acc[u], b, q);
        }
    }
}

// Sorted (descending) ranking keys: compare-and-swap, and the merge of this lane's
// top-8 with lane ^ xr's (the larger of A[i] and B[7-i] is a bitonic sequence
// holding the top 8, sorted by a bitonic merge).
__device__ __forceinline__ void key_cas(uint32_t &a, uint32_t &b)
{
    const uint32_t hi = max(a, b), lo = min(a, b);
    a = hi;
    b = lo;
}

__device__ __forceinline__ void bitonic8_desc(uint32_t t[8])
{
#pragma unroll
    for (int d = 4; d >= 1; d >>= 1)
#pragma unroll
        for (int j = 0; j < 8; j++)
            if ((j & d) == 0) key_cas(t[j], t[j + d]);
}

__device__ __forceinline__ void topk_merge_lanes(uint32_t t[8], int xr)
{
    uint32_t o[8];
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = (uint32_t)__shfl_xor((int)t[j], xr, 64);
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = max(t[j], o[7 - j]);
    bitonic8_desc(t);
}

__global__ __launch_bounds__(FW * 64) void k_fwd_fused(const uint16_t *__restrict__ x0, const float *__restrict__ p,
                                                   const uint16_t *__restrict__ wf, int h4, int w4, int tiles_x,
                                                   int tiles_y, float *__restrict__ logits,
                                                   int32_t *__restrict__ top8)
{
    __shared__ __align__(16) uint8_t s_a[LDS_A];
    __shared__ __align__(16) uint8_t s_b[LDS_B];
    __shared__ __align__(16) uint8_t s_f[LDS_X];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, col = lane & 15;
    const int tpi = tiles_x * tiles_y;
    const int img = blockIdx.x / tpi, tr = blockIdx.x - img * tpi;
    const int ty = tr / tiles_x, tx = tr - ty * tiles_x;
    const int h8 = h4 / 2, w8 = w4 / 2, hc = h8 / 2, wc = w8 / 2;
    const int oy = ty * S4, ox = tx * S4; // tile origin at 1/8 resolution

    // each 32 -> 32 layer's weights are loaded ahead of the barrier before the layer
    // (the first one's here, under the window staging and layer 1)
    u32x4 wa[9][2];
    float bl[2][4];
    load_w32(wf, lane, wa);
    load_bias(p + OFF_B2, q, bl);

    // x0 window (in the a2/a4 buffer): 1/4-res rows/cols 2(o - 3) - 1 .. + SX - 1,
    // zero outside the image
    uint32_t *s_x = (uint32_t *)s_b;
    {
        const uint32_t *xw = (const uint32_t *)x0 + (size_t)img * h4 * w4;
        const int y0 = 2 * (oy - 3) - 1, x0c = 2 * (ox - 3) - 1;
        for (int i = tid; i < SX * SX; i += FW * 64) {
            const int r = i / SX, c = i - r * SX, gy = y0 + r, gx = x0c + c;
            s_x[i] = gy >= 0 && gy < h4 && gx >= 0 && gx < w4 ? xw[(size_t)gy * w4 + gx] : 0u;
        }
    }
    __syncthreads();

    // L1: 3x3 stride 2, 2 -> 32; K = 9 taps x 2 channels = 18 of the 32 (rest zero)
    {
        u32x4 wa[2];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = 8 * q + 2 * j;
                const float *w = p + OFF_W1 + (16 * h + col) * 18 + k;
                wa[h][j] = k < 18 ? (uint32_t)f2bf(w[0]) | ((uint32_t)f2bf(w[1]) << 16) : 0u;
            }
        float b[2][4];
        load_bias(p + OFF_B1, q, b);
        constexpr int NPX = S1 * S1, NT = (NPX + 15) / 16;
        for (int tile = wave; tile < NT; tile += FW) {
            const int px = tile * 16 + col;
            const bool valid = px < NPX;
            const int pc = valid ? px : NPX - 1;
            const int y = pc / S1, x = pc - y * S1;
            u32x4 bv;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int t = 4 * q + j; // tap of k = 8q + 2j (both channels in one dword)
                bv[j] = t < 9 ? s_x[(2 * y + t / 3) * SX + 2 * x + t % 3] : 0u;
            }
            float4_t acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            acc[0] = mfma_bf16_k32(wa[0], bv, acc[0]);
            acc[1] = mfma_bf16_k32(wa[1], bv, acc[1]);
            const int gy = oy - 3 + y, gx = ox - 3 + x;
            conv_store(s_a, pc, valid, gy >= 0 && gy < h8 && gx >= 0 && gx < w8, acc, b, q);
        }
    }
    __syncthreads();
    conv32_window<S1>(s_a, s_b, wa, bl, oy, ox, 2, h8, w8, lane, wave);
    load_w32(wf + 9 * C * C, lane, wa);
    load_bias(p + OFF_B3, q, bl);
    __syncthreads();
    conv32_window<S2>(s_b, s_a, wa, bl, oy, ox, 1, h8, w8, lane, wave);
    load_w32(wf + 2 * 9 * C * C, lane, wa);
    load_bias(p + OFF_B4, q, bl);
    __syncthreads();
    conv32_window<S3>(s_a, s_b, wa, bl, oy, ox, 0, h8, w8, lane, wave);
    __syncthreads();

    // s_a is free once every wave has finished L4: stage the head's BF16 weights there
    // (rows of HSB bytes, padded so the 16 rows one MFMA reads fall in distinct banks)
    // with its FP32 bias, so the head loop reads LDS instead of waiting on global loads
    {
        const u32x4 *src = (const u32x4 *)(wf + WF_HEAD);
        for (int i = tid; i < K * 4; i += FW * 64) {
            const int row = i >> 2, part = i & 3;
            *(u32x4 *)(s_a + row * HSB + 16 * part) = src[i];
        }
        float *s_bh = (float *)(s_a + K * HSB);
        for (int i = tid; i < K; i += FW * 64) s_bh[i] = p[OFF_BH + i];
    }
    // 2x2 average pool -> f[cell][c] as BF16 (PSB bytes per cell); two channels per thread
    for (int i = tid; i < FT * FT * C / 2; i += FW * 64) {
        const int cell = i / (C / 2), c = 2 * (i - cell * (C / 2)), cy = cell / FT, cx = cell - cy * FT;
        const uint8_t *a = s_b + ((2 * cy) * S4 + 2 * cx) * PSB + 2 * c;
        uint32_t packed = 0;
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const float v00 = bf2f(*(const uint16_t *)(a + 2 * e)), v01 = bf2f(*(const uint16_t *)(a + PSB + 2 * e));
            const float v10 = bf2f(*(const uint16_t *)(a + S4 * PSB + 2 * e));
            const float v11 = bf2f(*(const uint16_t *)(a + S4 * PSB + PSB + 2 * e));
            packed |= (uint32_t)f2bf(0.25f * (v00 + v01 + v10 + v11)) << (16 * e);
        }
        *(uint32_t *)(s_f + cell * PSB + 2 * c) = packed;
    }
    __syncthreads();

    // head + top-8: wave w takes cells 16 (w % 4) .. + 15 of the tile (MFMA column =
    // cell) and half w / 4 of the bin tiles; the two halves' top-8 lists are merged
    // through LDS
    constexpr int NHALF = FW / 4; // wave groups sharing the bin tiles of the same 64 cells
    const int half = wave >> 2;
    const int cell = 16 * (wave & 3) + col, cy = cell / FT, cx = cell - cy * FT;
    const int gcy = ty * FT + cy, gcx = tx * FT + cx;
    const bool cvalid = gcy < hc && gcx < wc;
    const size_t gcell = ((size_t)img * hc + (cvalid ? gcy : 0)) * wc + (cvalid ? gcx : 0);
    const u32x4 fv = *(const u32x4 *)(s_f + cell * PSB + 16 * q); // B[k = 8q + i][cell]
    const float *s_bh = (const float *)(s_a + K * HSB);
    uint32_t top[8];
#pragma unroll
    for (int j = 0; j < 8; j++) top[j] = 0u;
    constexpr int NBT = (K + 15) / 16, NBH = NHALF == 2 ? (NBT + 1) / 2 : NBT;
    const int bt0 = half ? NBH : 0, bt1 = half ? NBT : NBH;
#pragma unroll 2
    for (int bt = bt0; bt < bt1; bt++) {
        const int arow = 16 * bt + col;
        const u32x4 wv = arow < K ? *(const u32x4 *)(s_a + arow * HSB + 16 * q) : u32x4{0, 0, 0, 0};
        float4_t acc = {0, 0, 0, 0};
        acc = mfma_bf16_k32(wv, fv, acc);
        uint32_t nk[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int bin = 16 * bt + 4 * q + i;
            const bool ok = bin < K;
            const float z = acc[i] + (ok ? s_bh[bin] : 0.0f);
            if (logits && cvalid && ok) logits[gcell * K + bin] = z;
            nk[i] = ok ? rank_key(z, bin) : 0u;
        }
        // sort the four new keys, keep the top 8 of the union
        key_cas(nk[0], nk[1]);
        key_cas(nk[2], nk[3]);
        key_cas(nk[0], nk[2]);
        key_cas(nk[1], nk[3]);
        key_cas(nk[1], nk[2]);
#pragma unroll
        for (int i = 0; i < 4; i++) top[4 + i] = max(top[4 + i], nk[3 - i]);
        bitonic8_desc(top);
    }
    topk_merge_lanes(top, 16);
    topk_merge_lanes(top, 32);
    // with two wave groups the upper hands its lists to the lower (s_b is free after the
    // pooling)
    uint32_t *s_top = (uint32_t *)s_b;
    if (NHALF == 2 && half && q == 0) {
#pragma unroll
        for (int j = 0; j < 8; j++) s_top[cell * 8 + j] = top[j];
    }
    if (NHALF == 2) __syncthreads();
    if (half) return;
    if (NHALF == 2) {
        uint32_t o[8];
#pragma unroll
        for (int j = 0; j < 8; j++) o[j] = s_top[cell * 8 + j];
#pragma unroll
        for (int j = 0; j < 8; j++) top[j] = max(top[j], o[7 - j]);
        bitonic8_desc(top);
    }
    if (top8 && cvalid && q == 0) {
        int b8[8];
#pragma unroll
        for (int j = 0; j < 8; j++) b8[j] = 511 - (int)(top[j] & 511u);
        *(int4 *)(top8 + gcell * TOPK) = make_int4(b8[0], b8[1], b8[2], b8[3]);
        *(int4 *)(top8 + gcell * TOPK + 4) = make_int4(b8[4], b8[5], b8[6], b8[7]);
    }
}

// ---------------------------------------------------------------------------
// Register-resident fused inference (the default).  A workgroup of RNV waves takes a
// tile of 32 x RH pixels at 1/8 resolution; wave w holds rows RS w .. RS w + RS - 1 of
// it in registers, each row as two MFMA B operands: E = the even pixels 2c, O = the odd
// pixels 2c + 1 (c = lane column).  Lane (q, c) of an operand carries channels
// {4q..4q+3, 16+4q..16+4q+3} of its pixel -- exactly what the two 16x16x32 products of
// that operand (output channels 0-15 and 16-31) return to the lane, so a layer's result
// feeds the next layer with no data movement (the weights take the channel permutation
// instead, k_refresh_rr).  Of a pixel's x-1 / x+1 neighbours, three of the four sit in
// the same lane; the other two operands are one DPP row shift each (the tile's side
// edges take zero).  A 3x3 layer walks its rows top-down with the input rows r-1, r,
// r+1; the waves exchange their first and last rows through LDS once per layer, so
// nothing inside a tile is computed twice; the tile's outer 4 pixels (3 layers of halo
// + cell alignment) are, by the neighbouring tile.  The 2x2 pool is E + O of two rows,
// inside the lane; the head takes the tile's cells 16 at a time from LDS, and its top-8
// selection merges only the bin tiles holding a logit at or above a lower bound of each
// cell's 8th largest (two logits of each of a cell's four lanes reach it).  Same
// arithmetic as the other inference kernels: BF16 activations and weights, FP32
// accumulation, zero padding at every layer's border.
// ---------------------------------------------------------------------------
constexpr int RNV = 4, RS = 14;                   // waves (strips stacked down the tile), rows per strip
constexpr int RW = 32, RH = RS * RNV;             // computed tile (1/8 res)
constexpr int RUW = RW - 8, RUH = RH - 8;         // useful output: 24 x 48 pixels = 12 x 24 cells
constexpr int RCW = RUW / 2, RCH = RUH / 2;       // cells per tile row / column
constexpr int RCELLS = RCW * RCH;
constexpr int RXW = 2 * RW + 1, RXH = 2 * RH + 1; // x0 window (1/4 res)
constexpr int RREC = 4 * 1024;                    // one strip's end rows: first E, first O, last E, last O
constexpr int RCP = 80;                           // LDS bytes per pooled cell (64 + 16 pad)
constexpr int RNBT = (K + 15) / 16;               // head bin tiles
constexpr int RLDS_A = ((RXH * RXW * 4 > RCELLS * RCP ? RXH * RXW * 4 : RCELLS * RCP) + 15) & ~15;
constexpr int RBIAS = 3 * RS * C * 4;              // layers 2-4 biases, one copy per strip row
constexpr int RWL = 12 * 1024;                      // one layer's kernel rows 1 and 3 (taps 0-2, 6-8), both halves
constexpr int RBH = RNBT * 16 * 4;                  // the head's biases (zero past bin 288)
// region B: end rows (parity 0), a zero row, the biases, the weights of layers 2 and 4,
// the head's biases; region A holds the x0 window, then (after layer 1) the end rows of
// parity 1 with layer 3's weights behind them, then the pooled cells
constexpr int RLDS_B = RNV * RREC + 1024 + RBIAS + 2 * RWL + RBH;
static_assert(RNV * RREC + RWL <= RLDS_A, "region A holds the end rows of parity 1 and layer 3's weights");
static_assert(RLDS_A % 16 == 0, "region B 16-byte aligned");
static_assert(RNV * RREC <= RLDS_A, "the end rows of parity 1 fit region A");
static_assert(RCELLS % 16 == 0 && RS % 2 == 0, "whole head fragments per tile, cells inside a strip");

__device__ __forceinline__ u32x4 dpp_from_left0(u32x4 x) // lane c <- c - 1, lane 0 <- 0
{
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (uint32_t)__builtin_amdgcn_mov_dpp((int)x[k], 0x111, 0xF, 0xF, true);
    return o;
}

__device__ __forceinline__ u32x4 dpp_from_right0(u32x4 x) // lane c <- c + 1, lane 15 <- 0
{
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (uint32_t)__builtin_amdgcn_mov_dpp((int)x[k], 0x101, 0xF, 0xF, true);
    return o;
}

// a value the compiler can neither see through nor move across the other empty asm
// statements, barriers and scheduling barriers
__device__ __forceinline__ uint32_t opaque_u32(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// bias already in the accumulators; ReLU and the image mask as one median (lim = +inf
// inside the image, 0 outside); BF16 in the B-operand slot order.  The packed words are
// made opaque where they are produced: otherwise the compiler sinks the packing into the
// next layer (the only user of most rows), past the barrier, and holds two FP32 values
// per word through the layer, which spills.
__device__ __forceinline__ u32x4 rr_pack(float4_t a0, float4_t a1, float lim)
{
    u32x4 o;
    o[0] = opaque_u32(f2bf2(__builtin_amdgcn_fmed3f(a0[0], 0.0f, lim), __builtin_amdgcn_fmed3f(a0[1], 0.0f, lim)));
    o[1] = opaque_u32(f2bf2(__builtin_amdgcn_fmed3f(a0[2], 0.0f, lim), __builtin_amdgcn_fmed3f(a0[3], 0.0f, lim)));
    o[2] = opaque_u32(f2bf2(__builtin_amdgcn_fmed3f(a1[0], 0.0f, lim), __builtin_amdgcn_fmed3f(a1[1], 0.0f, lim)));
    o[3] = opaque_u32(f2bf2(__builtin_amdgcn_fmed3f(a1[2], 0.0f, lim), __builtin_amdgcn_fmed3f(a1[3], 0.0f, lim)));
    return o;
}

// the 3 taps (weights w[0..2][h]) of one input row into the even (ODD = false) or odd
// output pixels of a row: even pixel 2c reads odd pixel 2c - 1 (lane c - 1), even 2c,
// odd 2c + 1; odd pixel 2c + 1 reads even 2c, odd 2c + 1, even 2c + 2 (lane c + 1)
template <bool ODD>
__device__ __forceinline__ void rr_row_taps(const u32x4 (&w)[3][2], u32x4 e, u32x4 o, float4_t (&a)[2])
{
    const u32x4 x0 = ODD ? e : dpp_from_left0(o), x1 = ODD ? o : e, x2 = ODD ? dpp_from_right0(e) : o;
#pragma unroll
    for (int h = 0; h < 2; h++) a[h] = mfma_bf16_k32(w[0][h], x0, a[h]);
#pragma unroll
    for (int h = 0; h < 2; h++) a[h] = mfma_bf16_k32(w[1][h], x1, a[h]);
#pragma unroll
    for (int h = 0; h < 2; h++) a[h] = mfma_bf16_k32(w[2][h], x2, a[h]);
}

// wr index of entry i (0..767) of layer l's LDS copy of kernel rows 1 and 3: record j =
// i / 64 (j = 0..5: taps 0-2, j = 6..11: taps 6-8, the two halves interleaved), lane i % 64
__device__ __forceinline__ int rr_wrow_rec(int l, int i)
{
    const int j = i >> 6, t = j < 6 ? j >> 1 : 6 + ((j - 6) >> 1);
    return (WR_LAYER + (l * 9 + t) * 2 + (j & 1)) * 64 + (i & 63);
}


// The top-8 ranking keys (largest first) of this lane's 76 bins of one cell.  A key v
// goes into the sorted list t by t'[0] = max(t[0], v), t'[k] = med3(t[k-1], t[k], v)
// (eight independent operations).  rank_key's value in three: the order map as
// z ^ (sign extension | 2^31), the bin's low bits inserted.
__device__ __forceinline__ void rr_head_top8(const float4_t (&acc)[RNBT], int q, uint32_t (&top)[8])
{
#pragma unroll
    for (int j = 0; j < 8; j++) top[j] = 0u;
    const uint32_t lb = 511u - 4u * (uint32_t)q;
#pragma unroll
    for (int bt = 0; bt < RNBT; bt++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (16 * bt + 15 >= K && i > 0) continue; // the last tile holds one bin (288, lane q 0)
            const int z = __float_as_int(acc[bt][i]);
            const uint32_t o = (uint32_t)z ^ ((uint32_t)(z >> 31) | 0x80000000u);
            uint32_t v = (o & ~511u) | (lb - (uint32_t)(16 * bt + i));
            if (16 * bt + 15 >= K && q > 0) v = 0u; // past the last bin
            uint32_t t2[8];
            t2[0] = max(top[0], v);
#pragma unroll
            for (int k = 1; k < 8; k++) t2[k] = max(min(top[k - 1], top[k]), min(max(top[k - 1], top[k]), v));
#pragma unroll
            for (int k = 0; k < 8; k++) top[k] = t2[k];
        }
    }
    topk_merge_lanes(top, 16);
    topk_merge_lanes(top, 32);
}

__global__ __launch_bounds__(RNV * 64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_fwd_rr(
    const uint32_t *__restrict__ x0, const float *__restrict__ p, const u32x4 *__restrict__ wr, int h4, int w4,
    int tiles_x, int tiles_y, float *__restrict__ logits, int32_t *__restrict__ top8)
{
    __shared__ __align__(16) uint8_t s_mem[RLDS_A + RLDS_B];
    uint8_t *s_A = s_mem, *s_B = s_mem + RLDS_A;
    constexpr int NT = RNV * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, col = lane & 15;
    const int tpi = tiles_x * tiles_y;
    const int img = blockIdx.x / tpi, tr = blockIdx.x - img * tpi;
    const int ty = tr / tiles_x, tx = tr - ty * tiles_x;
    const int h8 = h4 / 2, w8 = w4 / 2, hc = h8 / 2, wc = w8 / 2;
    const int X0 = tx * RUW - 4, Y0 = ty * RUH - 4; // tile origin (1/8 res)
    const int sy = Y0 + RS * wave;                  // this strip's first row
    const int pxe = X0 + 2 * col;                   // this lane's even pixel column
    const float inf = __builtin_huge_valf();
    const float limE = pxe >= 0 && pxe < w8 ? inf : 0.0f, limO = pxe + 1 >= 0 && pxe + 1 < w8 ? inf : 0.0f;
    uint8_t *zrow = s_B + RNV * RREC;
    float *s_bias = (float *)(zrow + 1024);
    uint8_t *s_wB = (uint8_t *)(s_bias + 3 * RS * C), *s_wA = s_A + RNV * RREC;
    float *s_bh = (float *)(s_wB + 2 * RWL);

    // x0 window: 1/4-res rows 2 Y0 - 1 .. + RXH - 1, cols 2 X0 - 1 .. + RXW - 1; zero
    // outside the image; the zero row of region B
    uint32_t *s_x = (uint32_t *)s_A;
    u32x4 w3v[12 * 64 / NT]; // layer 3's kernel rows 1 and 3, stored once the x0 window is dead
    {
        const uint32_t *xw = x0 + (size_t)img * h4 * w4;
        const int gy0 = 2 * Y0 - 1, gx0 = 2 * X0 - 1;
        // loads from clamped addresses, no branches: all of a thread's loads are in
        // flight together, the words outside the image zeroed afterwards
        constexpr int NX = (RXH * RXW + NT - 1) / NT;
        uint32_t v[NX];
#pragma unroll
        for (int k = 0; k < NX; k++) {
            const int i = tid + k * NT, r = i / RXW, c = i - r * RXW;
            const int gy = min(max(gy0 + r, 0), h4 - 1), gx = min(max(gx0 + c, 0), w4 - 1);
            v[k] = xw[(size_t)gy * w4 + gx];
        }
#pragma unroll
        for (int k = 0; k < NX; k++) {
            const int i = tid + k * NT, r = i / RXW, c = i - r * RXW, gy = gy0 + r, gx = gx0 + c;
            if (i < RXH * RXW) s_x[i] = gy >= 0 && gy < h4 && gx >= 0 && gx < w4 ? v[k] : 0u;
        }
        for (int i = tid; i < 1024 / 16; i += NT) *(u32x4 *)(zrow + 16 * i) = u32x4{0, 0, 0, 0};
        // the conv layers' biases, the head's biases, and layers 2 and 4's kernel rows 1
        // and 3 (read from LDS by every output row, so they hold no registers between rows);
        // every load of a thread is issued before the first store
        constexpr int NB = (3 * RS * C + NT - 1) / NT, NH = (RNBT * 16 + NT - 1) / NT, NW = 2 * 12 * 64 / NT;
        float vb[NB], vh[NH];
        u32x4 vw[NW];
#pragma unroll
        for (int k = 0; k < NB; k++) {
            // one copy per strip row: each row reads its own, so the compiler does not keep
            // them in registers across the layer (it has none spare)
            const int i = min(tid + k * NT, 3 * RS * C - 1);
            vb[k] = p[OFF_B2 + (i / (RS * C)) * (C * 9 * C + C) + i % C];
        }
#pragma unroll
        for (int k = 0; k < NH; k++) vh[k] = p[OFF_BH + min(tid + k * NT, K - 1)];
#pragma unroll
        for (int k = 0; k < NW; k++) {
            const int i = tid + k * NT; // 0..767 layer 2, 768..1535 layer 4
            vw[k] = wr[rr_wrow_rec(i < 768 ? 0 : 2, i % 768)];
        }
#pragma unroll
        for (int k = 0; k < 12 * 64 / NT; k++) w3v[k] = wr[rr_wrow_rec(1, tid + k * NT)];
#pragma unroll
        for (int k = 0; k < NB; k++)
            if (tid + k * NT < 3 * RS * C) s_bias[tid + k * NT] = vb[k];
#pragma unroll
        for (int k = 0; k < NH; k++)
            if (tid + k * NT < RNBT * 16) s_bh[tid + k * NT] = tid + k * NT < K ? vh[k] : 0.0f;
#pragma unroll
        for (int k = 0; k < NW; k++) *(u32x4 *)(s_wB + 16 * (tid + k * NT)) = vw[k];
    }
    __syncthreads();

    u32x4 XE[RS], XO[RS];
    // L1: 3x3 stride 2, 2 -> 32; K slots 8q + 2j + ch = (tap 4q + j, channel ch)
    {
        const u32x4 w1a = wr[lane], w1b = wr[64 + lane];
        float4_t b0, b1;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            b0[i] = p[OFF_B1 + 4 * q + i];
            b1[i] = p[OFF_B1 + 16 + 4 * q + i];
        }
        int toff[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int t = 4 * q + j;
            toff[j] = t < 9 ? (t / 3) * RXW + t % 3 : 0;
        }
        // pixel x of the window row 2y: x0 columns 2x .. 2x + 2
        const uint32_t *xb = s_x + 2 * (RS * wave) * RXW + 4 * col;
#pragma unroll
        for (int r = 0; r < RS; r++) {
            u32x4 be, bo;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // taps >= 9 read tap 0's word: their weights are zero and x0 is finite
                be[j] = xb[2 * r * RXW + toff[j]];
                bo[j] = xb[2 * r * RXW + 2 + toff[j]];
            }
            const float ry = (unsigned)(sy + r) < (unsigned)h8 ? inf : 0.0f;
            const u32x4 oe = rr_pack(mfma_bf16_k32(w1a, be, b0), mfma_bf16_k32(w1b, be, b1), fminf(ry, limE));
            const u32x4 oo = rr_pack(mfma_bf16_k32(w1a, bo, b0), mfma_bf16_k32(w1b, bo, b1), fminf(ry, limO));
            if (r == 0 || r == RS - 1) { // the strip's end rows: into its record (parity 0)
                *(u32x4 *)(s_B + RREC * wave + (r ? 2048 : 0) + 16 * lane) = oe;
                *(u32x4 *)(s_B + RREC * wave + (r ? 3072 : 1024) + 16 * lane) = oo;
            } else {
                XE[r] = oe;
                XO[r] = oo;
            }
        }
    }

    // A strip's first and last rows live in its record of the current parity, not in
    // registers: first E, first O, last E, last O (lane-major, 1 KB each).  The neighbours
    // read them from there, and so does the strip itself.  L1 wrote its records (parity 0)
    // above; layer l writes its first and last output rows into parity (l + 1) & 1 as soon
    // as it has them (every wave is past layer l - 1, that parity's last reader).
    const bool hasT = wave > 0, hasB = wave < RNV - 1;
    uint8_t *s_cell = s_A; // pooled cells: (ly, lx) of the tile at RCP bytes, written by layer 4
    const int lx = col - 2;
    const bool cx_ok = lx >= 0 && lx < RCW;
    __syncthreads();
    // every wave is past layer 1: the x0 window is dead, layer 3's weights go behind region
    // A's end rows (visible to layer 3 through the barrier after layer 2)
#pragma unroll
    for (int k = 0; k < 12 * 64 / NT; k++) *(u32x4 *)(s_wA + 16 * (tid + k * NT)) = w3v[k];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        u32x4 wc[3][2]; // the centre kernel row (taps 3-5); rows 1 and 3 come from LDS per output row
#pragma unroll
        for (int t = 0; t < 3; t++)
#pragma unroll
            for (int h = 0; h < 2; h++) wc[t][h] = wr[(WR_LAYER + (l * 9 + 3 + t) * 2 + h) * 64 + lane];
        const int wl = (l == 1 ? s_wA : s_wB + (l / 2) * RWL) - s_mem; // this layer's kernel rows 1 and 3
        const uint8_t *hb = (l & 1) ? s_A : s_B;       // this layer's input end rows
        uint8_t *ob = ((l + 1) & 1 ? s_A : s_B) + RREC * wave; // this strip's output end rows
        const uint8_t *own = hb + RREC * wave;
        const uint8_t *above = hasT ? hb + RREC * (wave - 1) + 2048 : zrow;
        u32x4 pE = *(const u32x4 *)(above + 16 * lane), pO = *(const u32x4 *)(above + (hasT ? 1024 : 0) + 16 * lane);
        u32x4 cE = *(const u32x4 *)(own + 16 * lane), cO = *(const u32x4 *)(own + 1024 + 16 * lane);
        u32x4 evE, evO; // layer 4: the even row of the pair being pooled
#pragma unroll
        for (int r = 0; r < RS; r++) {
            u32x4 nE, nO;
            if (r + 2 < RS) {
                nE = XE[r + 1];
                nO = XO[r + 1];
            } else if (r + 2 == RS) {
                nE = *(const u32x4 *)(own + 2048 + 16 * lane);
                nO = *(const u32x4 *)(own + 3072 + 16 * lane);
            } else {
                const uint8_t *below = hasB ? hb + RREC * (wave + 1) : zrow;
                nE = *(const u32x4 *)(below + 16 * lane);
                nO = *(const u32x4 *)(below + (hasB ? 1024 : 0) + 16 * lane);
            }
            const float *br = s_bias + (l * RS + r) * C + 4 * q;
            const float ry = (unsigned)(sy + r) < (unsigned)h8 ? inf : 0.0f;
            // the first and third kernel rows' weights, read per output row through an
            // address the compiler cannot prove equal to the other rows' (so it does not keep
            // them live), each applied to the even and the odd outputs before the next is read
            const uint8_t *wrow = s_mem + opaque_u32(wl + 16 * lane);
            float4_t aE[2] = {*(const float4_t *)br, *(const float4_t *)(br + 16)};
            float4_t aO[2] = {aE[0], aE[1]};
            {
                u32x4 w1[3][2];
#pragma unroll
                for (int t = 0; t < 3; t++)
#pragma unroll
                    for (int h = 0; h < 2; h++) w1[t][h] = *(const u32x4 *)(wrow + (2 * t + h) * 1024);
                rr_row_taps<false>(w1, pE, pO, aE);
                rr_row_taps<true>(w1, pE, pO, aO);
            }
            rr_row_taps<false>(wc, cE, cO, aE);
            rr_row_taps<true>(wc, cE, cO, aO);
            {
                u32x4 w3[3][2];
#pragma unroll
                for (int t = 0; t < 3; t++)
#pragma unroll
                    for (int h = 0; h < 2; h++) w3[t][h] = *(const u32x4 *)(wrow + (6 + 2 * t + h) * 1024);
                rr_row_taps<false>(w3, nE, nO, aE);
                rr_row_taps<true>(w3, nE, nO, aO);
            }
            const u32x4 outE = rr_pack(aE[0], aE[1], fminf(ry, limE)), outO = rr_pack(aO[0], aO[1], fminf(ry, limO));
            pE = cE;
            pO = cO;
            cE = nE;
            cO = nO;
            if (l < 2) {
                if (r == 0 || r == RS - 1) {
                    *(u32x4 *)(ob + (r ? 2048 : 0) + 16 * lane) = outE;
                    *(u32x4 *)(ob + (r ? 3072 : 1024) + 16 * lane) = outO;
                } else {
                    XE[r] = outE;
                    XO[r] = outO;
                }
            } else if ((r & 1) == 0) {
                evE = outE;
                evO = outO;
            } else {
                // layer 4: the 2x2 average pool of rows r - 1, r, straight into the cell
                // buffer (every wave is past layer 3, the last reader of region A)
                const int trow = RS * wave + r - 1, ly = (trow - 4) >> 1;
                if (trow >= 4 && ly < RCH) { // wave-uniform
                    u32x4 o;
#pragma unroll
                    for (int d = 0; d < 4; d++) {
                        const uint32_t e0 = evE[d], o0 = evO[d], e1 = outE[d], o1 = outO[d];
                        const float lo = ((__uint_as_float(e0 << 16) + __uint_as_float(o0 << 16)) +
                                          __uint_as_float(e1 << 16)) +
                                         __uint_as_float(o1 << 16);
                        const float hi = ((__uint_as_float(e0 & 0xffff0000u) + __uint_as_float(o0 & 0xffff0000u)) +
                                          __uint_as_float(e1 & 0xffff0000u)) +
                                         __uint_as_float(o1 & 0xffff0000u);
                        o[d] = f2bf2(0.25f * lo, 0.25f * hi);
                    }
                    if (cx_ok) *(u32x4 *)(s_cell + (ly * RCW + lx) * RCP + 16 * q) = o;
                }
            }
            // rows are scheduled one at a time: moving a later row's loads up holds
            // registers the layer does not have
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
    }
    u32x4 hw[RNBT]; // the head's weights
#pragma unroll
    for (int bt = 0; bt < RNBT; bt++) hw[bt] = wr[(WR_HEADREC + bt) * 64 + lane];

    // head + top-8, 16 cells per fragment (MFMA column = cell)
#pragma unroll 1
    for (int f = wave; f < RCELLS / 16; f += RNV) {
        const int c = 16 * f + col, ly = c / RCW, lx = c - ly * RCW;
        const int gcy = ty * RCH + ly, gcx = tx * RCW + lx;
        if (ty * RCH + (16 * f) / RCW >= hc) break; // every later fragment lies below the image
        const bool cvalid = gcy < hc && gcx < wc;
        const size_t gcell = ((size_t)img * hc + (cvalid ? gcy : 0)) * wc + (cvalid ? gcx : 0);
        u32x4 fv = *(const u32x4 *)(s_cell + c * RCP + 16 * q);
        if (!cvalid) fv = u32x4{0, 0, 0, 0};
        float4_t acc[RNBT];
#pragma unroll
        for (int bt = 0; bt < RNBT; bt++) acc[bt] = mfma_bf16_k32(hw[bt], fv, *(const float4_t *)(s_bh + 16 * bt + 4 * q));
        if (logits && cvalid) {
#pragma unroll
            for (int bt = 0; bt < RNBT; bt++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int bin = 16 * bt + 4 * q + i;
                    if (bin < K) logits[gcell * K + bin] = acc[bt][i];
                }
        }
        uint32_t top[8];
        rr_head_top8(acc, q, top);
        if (top8 && cvalid && q == 0) {
            int b8[8];
#pragma unroll
            for (int j = 0; j < 8; j++) b8[j] = 511 - (int)(top[j] & 511u);
            *(int4 *)(top8 + gcell * TOPK) = make_int4(b8[0], b8[1], b8[2], b8[3]);
            *(int4 *)(top8 + gcell * TOPK + 4) = make_int4(b8[4], b8[5], b8[6], b8[7]);
        }
    }
}

__global__ void k_fill_f32(float *p, size_t n, float v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned nblocks(long n, int b) { return (unsigned)((n + b - 1) / b); }

} // namespace

struct rh264_mvnet {
    int max_n, max_w, max_h;
    float *p, *g, *m, *v;   // FP32 parameters, gradients, Adam moments
    uint16_t *wf, *wb;      // BF16 conv weights (forward / data-gradient layouts)
    uint16_t *wr;           // BF16 weights in register order (k_fwd_rr)
    uint16_t *x0, *a1, *a2, *a3, *a4, *d4, *d3, *d2, *d1;
    float *f, *df, *logits, *dlogits, *loss;
    long step;
};

extern "C" {

size_t rh264_mvnet_param_count(void) { return NPARAM; }

int rh264_mvnet_destroy(rh264_mvnet *net)
{
    if (!net) return 0;
    void *bufs[] = {net->p, net->g, net->m, net->v, net->wf, net->wb, net->wr, net->x0, net->a1, net->a2, net->a3, net->a4,
                    net->d4, net->d3, net->d2, net->d1, net->f, net->df, net->logits, net->dlogits, net->loss};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete net;
    return 0;
}

int rh264_mvnet_create(rh264_ctx *ctx, int max_n, int max_w, int max_h, rh264_mvnet **out)
{
    if (!out) return fail("mvnet_create: null out");
    *out = nullptr;
    if (max_n < 1 || max_w < 16 || max_h < 16 || max_w % 16 || max_h % 16) return RH264_DECLINED;
    rh264_mvnet *net = new rh264_mvnet();
    memset(net, 0, sizeof(*net));
    net->max_n = max_n;
    net->max_w = max_w;
    net->max_h = max_h;
    const size_t n = max_n, px8 = (size_t)(max_w / 8) * (max_h / 8), px4 = (size_t)(max_w / 4) * (max_h / 4);
    const size_t cells = (size_t)(max_w / 16) * (max_h / 16);
    struct {
        void **ptr;
        size_t bytes;
    } allocs[] = {
        {(void **)&net->p, NPARAM * 4}, {(void **)&net->g, NPARAM * 4}, {(void **)&net->m, NPARAM * 4},
        {(void **)&net->v, NPARAM * 4}, {(void **)&net->wf, WF_COUNT * 2}, {(void **)&net->wb, 3 * 9 * C * C * 2},
        {(void **)&net->wr, WR_COUNT * 2},
        {(void **)&net->x0, n * px4 * 2 * 2}, {(void **)&net->a1, n * px8 * C * 2}, {(void **)&net->a2, n * px8 * C * 2},
        {(void **)&net->a3, n * px8 * C * 2}, {(void **)&net->a4, n * px8 * C * 2}, {(void **)&net->d4, n * px8 * C * 2},
        {(void **)&net->d3, n * px8 * C * 2}, {(void **)&net->d2, n * px8 * C * 2}, {(void **)&net->d1, n * px8 * C * 2},
        {(void **)&net->f, n * cells * C * 4}, {(void **)&net->df, n * cells * C * 4},
        {(void **)&net->logits, n * cells * K * 4}, {(void **)&net->dlogits, n * cells * K * 4},
        {(void **)&net->loss, 64},
    };
    for (auto &a : allocs) {
        hipError_t e = hipMalloc(a.ptr, a.bytes);
        if (e != hipSuccess) {
            rh264_mvnet_destroy(net);
            return fail("mvnet_create: %s", hipGetErrorString(e));
        }
    }
    hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(NPARAM, 256)), dim3(256), 0, ctx->stream, net->m, (size_t)NPARAM, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(NPARAM, 256)), dim3(256), 0, ctx->stream, net->v, (size_t)NPARAM, 0.0f);
    *out = net;
    return 0;
}

int rh264_mvnet_set_params(rh264_ctx *ctx, rh264_mvnet *net, const float *host, size_t count)
{
    if (count != (size_t)NPARAM) return fail("mvnet_set_params: %zu parameters, expected %d", count, NPARAM);
    HIP_TRY(hipMemcpyAsync(net->p, host, NPARAM * 4, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_refresh_bf16, dim3(nblocks(WF_COUNT, 256)), dim3(256), 0, ctx->stream, net->p, net->wf,
                       net->wb);
    hipLaunchKernelGGL(k_refresh_rr, dim3(nblocks(WR_COUNT, 256)), dim3(256), 0, ctx->stream, net->p, net->wr);
    LAUNCH_CHECK("k_refresh_bf16");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int rh264_mvnet_get_params(rh264_ctx *ctx, rh264_mvnet *net, float *host, size_t count)
{
    if (count != (size_t)NPARAM) return fail("mvnet_get_params: %zu parameters, expected %d", count, NPARAM);
    HIP_TRY(hipMemcpyAsync(host, net->p, NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int rh264_mvnet_get_grads(rh264_ctx *ctx, rh264_mvnet *net, float *host, size_t count)
{
    if (count != (size_t)NPARAM) return fail("mvnet_get_grads: %zu parameters, expected %d", count, NPARAM);
    HIP_TRY(hipMemcpyAsync(host, net->g, NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

static int mvnet_geom(rh264_mvnet *net, int n, int w, int h)
{
    if (n < 1 || n > net->max_n || w % 16 || h % 16 || w < 16 || h < 16 || (size_t)w * h > (size_t)net->max_w * net->max_h)
        return fail("mvnet: batch %d of %dx%d exceeds the network's buffers", n, w, h);
    return 0;
}

static int mvnet_forward_impl(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *cur, const uint8_t *ref, size_t stride,
                              int n, int w, int h)
{
    const int h4 = h / 4, w4 = w / 4, h8 = h / 8, w8 = w / 8;
    const long cells = (long)n * (h / 16) * (w / 16);
    hipLaunchKernelGGL(k_prep, dim3(nblocks((long)n * h4 * w4, 256)), dim3(256), 0, ctx->stream, cur, ref, stride, n, w,
                       h, net->x0, (unsigned *)nullptr);
    hipLaunchKernelGGL(k_conv1_fwd, dim3(nblocks((long)n * h8 * w8, 256)), dim3(256), 0, ctx->stream, net->x0, net->p,
                       n, h4, w4, net->a1);
    const long waves = (long)n * h8 * ((w8 + 15) / 16);
    uint16_t *acts[4] = {net->a1, net->a2, net->a3, net->a4};
    for (int l = 0; l < 3; l++)
        hipLaunchKernelGGL(k_conv3_mfma, dim3(nblocks(waves, 4)), dim3(256), 0, ctx->stream, acts[l],
                           net->wf + (size_t)l * 9 * C * C, net->p + OFF_B2 + l * (C * 9 * C + C), 1,
                           (const uint16_t *)nullptr, n, h8, w8, acts[l + 1]);
    hipLaunchKernelGGL(k_pool_fwd, dim3(nblocks(cells * C, 256)), dim3(256), 0, ctx->stream, net->a4, n, h8, w8, net->f);
    hipLaunchKernelGGL(k_head_fwd, dim3(nblocks(cells * K, 256)), dim3(256), 0, ctx->stream, net->f, net->p, cells,
                       net->logits);
    LAUNCH_CHECK("mvnet forward");
    return 0;
}

int mvnet_forward_ev(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *cur, const uint8_t *ref, size_t stride, int n,
                     int w, int h, float *logits, int32_t *top8, hipEvent_t after_prep, unsigned *zero)
{
    int rc = mvnet_geom(net, n, w, h);
    if (rc) return rc;
    if ((uintptr_t)top8 & 15) return fail("mvnet_forward: top8 must be 16-byte aligned");
    const long cells = (long)n * (h / 16) * (w / 16);
    if (ctx->mvnet_fused == 2) {
        const int h4 = h / 4, w4 = w / 4;
        const int tiles_x = (w / 16 + RCW - 1) / RCW, tiles_y = (h / 16 + RCH - 1) / RCH;
        hipLaunchKernelGGL(k_prep, dim3(nblocks((long)n * h4 * w4, 256)), dim3(256), 0, ctx->stream, cur, ref, stride,
                           n, w, h, net->x0, zero);
        if (after_prep) HIP_TRY(hipEventRecord(after_prep, ctx->stream));
        hipLaunchKernelGGL(k_fwd_rr, dim3((unsigned)((long)n * tiles_x * tiles_y)), dim3(RNV * 64), 0, ctx->stream,
                           (const uint32_t *)net->x0, net->p, (const u32x4 *)net->wr, h4, w4, tiles_x, tiles_y, logits,
                           top8);
        LAUNCH_CHECK("k_fwd_rr");
        return 0;
    }
    if (ctx->mvnet_fused) {
        const int h4 = h / 4, w4 = w / 4;
        const int tiles_x = (w / 16 + FT - 1) / FT, tiles_y = (h / 16 + FT - 1) / FT;
        hipLaunchKernelGGL(k_prep, dim3(nblocks((long)n * h4 * w4, 256)), dim3(256), 0, ctx->stream, cur, ref, stride,
                           n, w, h, net->x0, zero);
        if (after_prep) HIP_TRY(hipEventRecord(after_prep, ctx->stream));
        hipLaunchKernelGGL(k_fwd_fused, dim3((unsigned)((long)n * tiles_x * tiles_y)), dim3(FW * 64), 0, ctx->stream,
                           net->x0, net->p, net->wf, h4, w4, tiles_x, tiles_y, logits, top8);
        LAUNCH_CHECK("k_fwd_fused");
        return 0;
    }
    rc = mvnet_forward_impl(ctx, net, cur, ref, stride, n, w, h);
    if (rc) return rc;
    if (zero) {
        hipLaunchKernelGGL(k_fill_u32_nv, dim3(1), dim3(64), 0, ctx->stream, zero);
        LAUNCH_CHECK("k_fill_u32_nv");
    }
    if (logits)
        HIP_TRY(hipMemcpyAsync(logits, net->logits, (size_t)cells * K * 4, hipMemcpyDeviceToDevice, ctx->stream));
    if (top8) {
        hipLaunchKernelGGL(k_topk, dim3(nblocks(cells, 4)), dim3(256), 0, ctx->stream, net->logits, cells, top8);
        LAUNCH_CHECK("k_topk");
    }
    return 0;
}

int rh264_mvnet_forward(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *cur, const uint8_t *ref, size_t stride,
                        int n, int w, int h, float *logits, int32_t *top8)
{
    return mvnet_forward_ev(ctx, net, cur, ref, stride, n, w, h, logits, top8, nullptr, nullptr);
}

// Forward + backward on a batch; gradients (mean cross-entropy over cells) left in
// the network's gradient buffer; loss (mean) returned through loss_out (host).
int rh264_mvnet_backward(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *cur, const uint8_t *ref, size_t stride,
                         int n, int w, int h, const int32_t *labels, float *loss_out)
{
    int rc = mvnet_geom(net, n, w, h);
    if (rc) return rc;
    rc = mvnet_forward_impl(ctx, net, cur, ref, stride, n, w, h);
    if (rc) return rc;
    const int h4 = h / 4, w4 = w / 4, h8 = h / 8, w8 = w / 8;
    const long cells = (long)n * (h / 16) * (w / 16);
    const float scale = 1.0f / (float)cells;
    hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(NPARAM, 256)), dim3(256), 0, ctx->stream, net->g, (size_t)NPARAM, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, ctx->stream, net->loss, (size_t)1, 0.0f);
    hipLaunchKernelGGL(k_softmax_ce, dim3(nblocks(cells, 4)), dim3(256), 0, ctx->stream, net->logits, labels, cells,
                       scale, net->dlogits, net->loss);
    hipLaunchKernelGGL(k_head_bwd_w, dim3(nblocks(K * C + K, 256), nblocks(cells, 256)), dim3(256), 0, ctx->stream,
                       net->dlogits, net->f, cells, net->g);
    hipLaunchKernelGGL(k_head_bwd_data, dim3(nblocks(cells * C, 256)), dim3(256), 0, ctx->stream, net->dlogits, net->p,
                       cells, net->df);
    const long px8 = (long)n * h8 * w8;
    hipLaunchKernelGGL(k_pool_bwd, dim3(nblocks(px8 * C, 256)), dim3(256), 0, ctx->stream, net->df, net->a4, n, h8, w8,
                       net->d4);
    // layers 4, 3, 2: weight gradient from (input activation, dz); data gradient
    // dz_prev = conv(dz, flipped transposed w) masked by ReLU'(prev activation)
    uint16_t *acts[4] = {net->a1, net->a2, net->a3, net->a4};
    uint16_t *dzs[4] = {net->d1, net->d2, net->d3, net->d4};
    const long waves = (long)n * h8 * ((w8 + 15) / 16);
    for (int l = 2; l >= 0; l--) { // conv layer l+2 maps acts[l] -> acts[l+1], dz = dzs[l+1]
        const int off_w = OFF_W2 + l * (C * 9 * C + C), off_b = off_w + C * 9 * C;
        hipLaunchKernelGGL((k_conv_wgrad<C, 1>), dim3(nblocks(px8, 32)), dim3(256), 0, ctx->stream, acts[l], dzs[l + 1],
                           n, h8, w8, h8, w8, net->g + off_w, net->g + off_b);
        hipLaunchKernelGGL(k_conv3_mfma, dim3(nblocks(waves, 4)), dim3(256), 0, ctx->stream, dzs[l + 1],
                           net->wb + (size_t)l * 9 * C * C, (const float *)nullptr, 0, acts[l], n, h8, w8, dzs[l]);
    }
    hipLaunchKernelGGL((k_conv_wgrad<2, 2>), dim3(nblocks(px8, 32)), dim3(256), 0, ctx->stream, net->x0, net->d1, n, h4,
                       w4, h8, w8, net->g, net->g + OFF_B1);
    LAUNCH_CHECK("mvnet backward");
    if (loss_out) {
        HIP_TRY(hipMemcpyAsync(loss_out, net->loss, 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return 0;
}

int rh264_mvnet_adam(rh264_ctx *ctx, rh264_mvnet *net, float lr)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    net->step++;
    const float c1 = 1.0f - powf(b1, (float)net->step), c2 = 1.0f - powf(b2, (float)net->step);
    hipLaunchKernelGGL(k_adam, dim3(nblocks(NPARAM, 256)), dim3(256), 0, ctx->stream, net->p, net->g, net->m, net->v,
                       lr, b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(k_refresh_bf16, dim3(nblocks(WF_COUNT, 256)), dim3(256), 0, ctx->stream, net->p, net->wf,
                       net->wb);
    hipLaunchKernelGGL(k_refresh_rr, dim3(nblocks(WR_COUNT, 256)), dim3(256), 0, ctx->stream, net->p, net->wr);
    LAUNCH_CHECK("mvnet adam");
    return 0;
}

} // extern "C"
