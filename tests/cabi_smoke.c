/* A C caller of the provider, bound the way the Rust accel crate's extern "C" block
 * would bind it: the header, the shared library, plain pointers.  It checks one
 * per-block entry (rh264_sad_16x16, rh264_satd_8x8) against sums computed here, and
 * one batched entry (rh264_build_planes + rh264_full_search) on a frame whose
 * content moved by a known displacement.  Exit status 0 = all agree. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rh264_me.h"

static int sad_host(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h)
{
    int s = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) s += abs((int)a[y * sa + x] - (int)b[y * sb + x]);
    return s;
}

int main(void)
{
    enum { W = 64, H = 48, PAD = 32 };
    uint8_t *ref = malloc(W * H), *cur = malloc(W * H);
    unsigned s = 12345u;
    for (int i = 0; i < W * H; i++) {
        s = s * 1103515245u + 12345u;
        ref[i] = (uint8_t)(s >> 16);
    }
    /* cur(y, x) = ref(y + 2, x - 3): the content moved by (+3, -2) pels */
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int yy = y + 2 < H ? y + 2 : H - 1, xx = x - 3 >= 0 ? x - 3 : 0;
            cur[y * W + x] = ref[yy * W + xx];
        }
    int32_t v = -1;
    int rc = rh264_sad_16x16(cur + 16 * W + 16, W, ref + 5 * W + 9, W, &v);
    int want = sad_host(cur + 16 * W + 16, W, ref + 5 * W + 9, W, 16, 16);
    printf("rh264_sad_16x16 rc=%d value=%d expected=%d\n", rc, v, want);
    if (rc != 0 || v != want) return 1;
    rc = rh264_satd_8x8(cur, W, cur, W, &v);
    printf("rh264_satd_8x8 (identical blocks) rc=%d value=%d\n", rc, v);
    if (rc != 0 || v != 0) return 1;
    uint32_t q[4];
    size_t o[4] = {0, 1, W, W + 1};
    rc = rh264_sad_x4(cur, W, ref, o, W, 4, 4, q);
    printf("rh264_sad_x4 4x4 rc=%d (declined = 1)\n", rc);
    if (rc != RH264_DECLINED) return 1;

    rh264_ctx *ctx = NULL;
    if (rh264_init(0, &ctx)) { printf("init: %s\n", rh264_last_error()); return 1; }
    size_t pb = rh264_planes_bytes(W, H, PAD);
    void *dref, *dcur, *dpl, *dmv, *dcost;
    if (rh264_alloc(ctx, W * H, &dref) || rh264_alloc(ctx, W * H, &dcur) || rh264_alloc(ctx, pb, &dpl) ||
        rh264_alloc(ctx, 8 * 12, &dmv) || rh264_alloc(ctx, 4 * 12, &dcost)) {
        printf("alloc: %s\n", rh264_last_error());
        return 1;
    }
    rh264_h2d(ctx, dref, ref, W * H);
    rh264_h2d(ctx, dcur, cur, W * H);
    rh264_search_params p;
    memset(&p, 0, sizeof p);
    p.cw = W; p.ch = H; p.pad = PAD; p.bw = 16; p.bh = 16; p.range = 8; p.subpel_pat = 0; p.hadamard = 4;
    rh264_lambda_table(0.0, p.lr);
    rh264_lambda_table(0.0, p.lr_fp);
    rc = rh264_build_planes(ctx, dref, W * H, 1, W, H, PAD, dpl);
    if (!rc) rc = rh264_full_search(ctx, &p, 1, dcur, W * H, dpl, pb, NULL, NULL, dmv, dcost);
    if (rc) { printf("full_search: %s\n", rh264_last_error()); return 1; }
    int32_t mv[24], cost[12];
    rh264_d2h(ctx, mv, dmv, sizeof mv);
    rh264_d2h(ctx, cost, dcost, sizeof cost);
    /* an interior block (1, 1) finds (-3, +2) pels = (-12, +8) quarter pels at cost 0 */
    int b = 1 * (W / 16) + 1;
    printf("rh264_full_search block (1,1): mv (%d, %d) cost %d\n", mv[2 * b], mv[2 * b + 1], cost[b]);
    if (mv[2 * b] != -12 || mv[2 * b + 1] != 8 || cost[b] != 0) return 1;
    rh264_destroy(ctx);
    rh264_accel_release();
    printf("cabi_smoke ok\n");
    return 0;
}
