// rh264_internal.h -- declarations shared by the provider's HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <vector>

#include "rh264_me.h"

int rh264_fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
#define fail rh264_fail

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail("%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

#define LAUNCH_CHECK(name)                                                                    \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess) return fail("launch of %s failed: %s", name, hipGetErrorString(e_)); \
    } while (0)

constexpr int RH264_PROF_EV = 6; // phase boundaries per profiled search

struct rh264_ctx {
    int device;
    hipStream_t stream;
    hipStream_t side;             // second stream (default priority): the learned search's fall-back
    hipEvent_t fork_ev, join_ev;  // stream -> side after the candidates, side -> stream after the fall-back
    hipEvent_t ev[16];
    uint64_t *keys;               // per-block argmin keys of the full search
    uint64_t *noalloc_keys;       // this device's static key array (rh264_plain_search_noalloc)
    size_t keys_cap;              // in elements
    int32_t *tmp_mv;              // integer-stage MVs when the caller does not want them
    size_t tmp_mv_cap;            // in int32 elements
    int fs_kernel;                // integer search: 0 = quad-SAD, 1 = v_sad_u8 x2, 2 = v_sad_u8 x4
    int refine_lds;               // refinement stages the +-3 pel window in LDS (blocks of >= 8 sub-blocks)
    int refine_mfma;              // 16x16 refinement: Hadamards as matrix-core products, one wave per block
    int cand_quad;                // learned search: candidate windows with quad-SADs (local range <= 3)
    int cand_lds;                 // learned search: the lane-per-candidate kernel reads the source rows from LDS
    int fallback_v;               // learned search: fall-back full search, 0 = quad-SAD, 1 = v_sad_u8 x2
    int fallback_wgs;             // learned search: the fall-back full search's grid cap (workgroups)
    int fallback_overlap;         // learned search: the fall-back search runs beside the other blocks' refinement
    int planes_rw;                // plane build: 1 = register-window kernel, 0 = LDS tile kernel
    int mvnet_fused;              // network inference: 2 = register-resident fused kernel, 1 = the LDS-tile fused
                                  // kernel, 0 = layer by layer
    std::vector<hipEvent_t> prof; // RH264_PROF_EV events per profiled search
    std::vector<int> prof_np;     // phases recorded in each profiled search
    int prof_cap, prof_n;         // capacity and count of profiled searches
    void *ls_buf;                 // scratch of the learned-proposal search
    size_t ls_cap;
    void *la_buf;                 // padded half-res planes of the multi-reference lookahead
    size_t la_cap;
    void *lal_buf;                // proposals, vectors and fall-back list of its learned-seeded form
    size_t lal_cap;
    void *ms_buf;                 // predictors, work list and counters of the fixed-point decision
    size_t ms_cap;
    int la_scaled;                // learned-seeded lookahead: 1 = distance-1 proposals scaled by the distance,
                                  // 0 = the network run for every reference distance
    int decide_mode;              // encoder-order decision: 0 = one launch per anti-diagonal, 1 = fixed-point
                                  // passes, 2 = the fixed point for 1-2 references, else the anti-diagonals
};

// Network inference (rh264_mvnet_forward) recording `after_prep` (if not null) on the
// context's stream between the input preparation and the fused network kernel, and
// clearing the counter `zero` (if not null) on the stream before the network's end.
extern "C" int mvnet_forward_ev(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *cur, const uint8_t *ref,
                                size_t stride, int n, int w, int h, float *logits, int32_t *top8,
                                hipEvent_t after_prep, unsigned *zero);
