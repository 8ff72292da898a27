/*
 * rh264_me.h -- C ABI of the MI355X (gfx950) motion-search provider for
 * rusty_h264.
 *
 * This is the drop-in boundary for the encoder's motion-search hot path.  In
 * the reference every kernel on that path sits in the one crate that may use
 * `unsafe`, crates/rusty_h264-accel (src/lib.rs:30-58, src/x86_asm.rs:13-21);
 * its callers are crates/rusty_h264-common/src/inter.rs and
 * crates/rusty_h264-encoder/src/{mb16.rs,mbtree.rs}.  A Rust `extern "C"`
 * block in that crate binds the functions below.
 *
 * Conventions (mirroring the accel crate's):
 *   - return 0 = done, 1 = declined (unsupported shape/geometry: the caller runs
 *     its own path, as the reference does when an accel kernel returns None /
 *     false), < 0 = error (rh264_last_error() describes it).
 *   - all frame/plane/result pointers are DEVICE pointers (use rh264_alloc /
 *     rh264_h2d / rh264_d2h, or any HIP allocation on the context's device);
 *     no torch types; sizes are plain integers.
 *   - every call is asynchronous on the context's stream; rh264_sync() waits.
 *   - a context is used by one host thread at a time; separate contexts may be
 *     used concurrently (the reference's kernels are called from GOP-parallel
 *     threads, crates/rusty_h264-encoder/src/lib.rs:746-800).
 *   - motion vectors are quarter-pel int32 pairs (x, y); costs are int32.
 *
 * Plane-set layout: one buffer of rh264_planes_bytes(cw, ch, pad) bytes holding
 * the four planes of inter.rs:707-722 in the order F (edge-replicated full-pel
 * copy), H (horizontal half-pel "b"), V (vertical half-pel "h"), C (centre
 * "j"), each pw*ph bytes with stride pw = cw + 2*pad, ph = ch + 2*pad, followed
 * by a small zeroed tail so vector loads near the end stay inside the buffer.
 */
#ifndef RH264_ME_H
#define RH264_ME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RH264_OK 0
#define RH264_DECLINED 1
#define RH264_LR_N 128 /* entries of a lambda-rate table */

typedef struct rh264_ctx rh264_ctx;

/* ---- context and device-memory plumbing (no reference counterpart: the
 *      reference is in-process CPU code with no device) -------------------- */
int rh264_init(int device, rh264_ctx **out);
int rh264_destroy(rh264_ctx *ctx);
const char *rh264_last_error(void);
int rh264_device_name(rh264_ctx *ctx, char *buf, int buflen);
int rh264_alloc(rh264_ctx *ctx, size_t bytes, void **dptr);
int rh264_free(rh264_ctx *ctx, void *dptr);
int rh264_h2d(rh264_ctx *ctx, void *dst, const void *src, size_t bytes);
int rh264_d2h(rh264_ctx *ctx, void *dst, const void *src, size_t bytes);
int rh264_memset(rh264_ctx *ctx, void *dst, int value, size_t bytes);
int rh264_sync(rh264_ctx *ctx);
/* Device-side timing on the context's stream (HIP events). */
int rh264_event_record(rh264_ctx *ctx, int slot);
int rh264_event_elapsed_ms(rh264_ctx *ctx, int slot_a, int slot_b, float *ms);

/* ---- block metrics, batched -----------------------------------------------
 * Replaces the per-block accel family (crates/rusty_h264-accel/src/satd_sad.rs:
 * 309-363 `satd_16x16|16x8|8x16|8x8|4x4`, `sad_16x16|16x8|8x16`;
 * crates/rusty_h264-accel/src/satd_avg.rs:350-602 `satd_x4p`, `satd_x4`,
 * `sad_x4`, `satd_avg_x4`, `satd_avg`).  A GPU cannot serve one 17 ns block
 * per call, so each call evaluates n blocks: block i compares src + src_off[i]
 * (stride ss) with ref + ref_off[i] (stride rs) -- or, for RH264_METRIC_SATD_AVG,
 * with the rounded average (a+b+1)>>1 of ref + ref_off[i] and ref + ref_off2[i].
 * kind: RH264_METRIC_SAD        sum |a-b|                    (sad_scalar)
 *       RH264_METRIC_SATD_WELS  sum_4x4 ((sum|H d|+1)>>1)    (satd_region_scalar)
 *       RH264_METRIC_SATD       sum_4x4 sum|H d|             (satd_x4 / satd_px domain)
 *       RH264_METRIC_SATD_AVG   sum|H (src - avg(a,b))|      (satd_avg)
 *       RH264_METRIC_SATD8      sum_8x8 sum|H8 d H8| / 2     (no reference
 *         counterpart: the 8x8 Hadamard, H8 the order-8 Sylvester matrix; every
 *         coefficient has the parity of sum d, so the halving is exact; w, h
 *         multiples of 8)
 * Shapes: (w,h) in {16x16,16x8,8x16,8x8,8x4,4x8,4x4}; others decline. */
#define RH264_METRIC_SAD 0
#define RH264_METRIC_SATD_WELS 1
#define RH264_METRIC_SATD 2
#define RH264_METRIC_SATD_AVG 3
#define RH264_METRIC_SATD8 4
int rh264_block_metric(rh264_ctx *ctx, int kind, int w, int h, int n, const uint8_t *src, int ss,
                       const int64_t *src_off, const uint8_t *ref, int rs, const int64_t *ref_off,
                       const int64_t *ref_off2, int32_t *out);

/* ---- block metrics, per block (the accel crate's own call shape) ---------
 * The functions of crates/rusty_h264-accel with its names, argument order (slices
 * become pointer + stride) and value domains, for drop-in parity with the existing
 * call sites (crates/rusty_h264-common/src/inter.rs, crates/rusty_h264-encoder/src/
 * mb16.rs; the call list in accel/src/lib.rs:30-58 and x86_asm.rs:13-21).  HOST
 * pointers; each call is synchronous: it stages its blocks on the calling thread's
 * default context (device rh264_accel_set_device, default 0), runs one launch of
 * the kernels above and returns the values.  Return 0 with *out written, 1 where
 * the reference returns None / false (a shape outside satd_avg.rs:342-344's set, a
 * plane hpel_fused leaves to the caller), < 0 on error.  A GPU serves a batch, not
 * a 17 ns block: these are for parity and for callers that have one block; the
 * batched entries are the fast path. */
/* satd_16x16|16x8|8x16|8x8|4x4 (satd_sad.rs:309-340): sum_4x4 ((sum|H d| + 1) >> 1);
 * sad_16x16|16x8|8x16 (satd_sad.rs:342-363): sum |a - b| */
int rh264_satd_16x16(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out);
int rh264_satd_16x8(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out);
int rh264_satd_8x16(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out);
int rh264_satd_8x8(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out);
int rh264_satd_4x4(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out);
int rh264_sad_16x16(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out);
int rh264_sad_16x8(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out);
int rh264_sad_8x16(const uint8_t *a, size_t sa, const uint8_t *b, size_t sb, int32_t *out);
/* sad_x4 (satd_avg.rs:462-492) and satd_x4 (388-414): one w x h source block against
 * base + o[i] (stride rs); SAD, or the unhalved sum|H d|.  Shapes 16x16/16x8/8x16/8x8. */
int rh264_sad_x4(const uint8_t *src, size_t ss, const uint8_t *base, const size_t o[4], size_t rs, size_t w, size_t h,
                 uint32_t out[4]);
int rh264_satd_x4(const uint8_t *src, size_t ss, const uint8_t *base, const size_t o[4], size_t rs, size_t w,
                  size_t h, uint32_t out[4]);
/* satd_x4p (satd_avg.rs:350-381): four independent operands r[i] + o[i], stride rs */
int rh264_satd_x4p(const uint8_t *src, size_t ss, const uint8_t *const r[4], const size_t o[4], size_t rs, size_t w,
                   size_t h, uint32_t out[4]);
/* satd_avg_x4 (satd_avg.rs:536-571): sum|H (src - (a+b+1)>>1)| for the pairs
 * (a[i] + oa[i], b[i] + ob[i]) */
int rh264_satd_avg_x4(const uint8_t *src, size_t ss, const uint8_t *const a[4], const size_t oa[4],
                      const uint8_t *const b[4], const size_t ob[4], size_t rs, size_t w, size_t h, uint32_t out[4]);
/* satd_avg (satd_avg.rs:579-602) */
int rh264_satd_avg(const uint8_t *src, size_t src_stride, const uint8_t *a, const uint8_t *b, size_t stride, size_t w,
                   size_t h, uint32_t *out);
/* hpel_fused (hpel.rs:34-43): H, V, C (each pw x ph, stride pw) of the edge-padded
 * full-pel plane f, every tap clamped into f; declines (the caller builds them its
 * own way, as when the reference returns false) unless pw >= 24 and pw, ph are
 * multiples of 16 (pad 32 planes of 16-multiple frames always are). */
int rh264_hpel_fused(const uint8_t *f, size_t pw, size_t ph, uint8_t *h, uint8_t *v, uint8_t *c);
/* MeCtx (mectx.rs:27-166): new validates the geometry once and uploads the four
 * planes and the source block (1 = None: stride != pw, a shape outside
 * 16x16/16x8/8x16/8x8, a short slice, a plane too small); eval prices one quarter-pel
 * candidate, the encoder-domain SATD, 1 (None) outside the validated window
 * (mectx.rs:114-118); eval_n prices n candidates in two launches, ok[i] = 0 where
 * eval would return None. */
typedef struct rh264_mectx rh264_mectx;
int rh264_mectx_new(const uint8_t *src, size_t src_len, size_t cw, const uint8_t *f, const uint8_t *h,
                    const uint8_t *v, const uint8_t *c, size_t plane_len, size_t stride, size_t pad, size_t pw,
                    size_t ph, size_t lx, size_t ly, size_t w, size_t bh, rh264_mectx **out);
int rh264_mectx_eval(rh264_mectx *m, int32_t mvx, int32_t mvy, uint32_t *out);
int rh264_mectx_eval_n(rh264_mectx *m, int n, const int32_t *mvs, uint32_t *out, uint8_t *ok);
int rh264_mectx_free(rh264_mectx *m);
/* The device of the calling threads' default contexts (before their first call),
 * and the release of the calling thread's (no reference counterpart). */
int rh264_accel_set_device(int device);
int rh264_accel_release(void);

/* ---- half-pel plane sets -------------------------------------------------
 * Replaces build_hpel_planes (crates/rusty_h264-common/src/inter.rs:816-843:
 * pad_plane 786-814 + the fused builder crates/rusty_h264-accel/src/hpel.rs:34-203
 * / its twins inter.rs:868-987), for nframes frames in one launch.  Frame i
 * (cw x ch luma, stride cw) is at frames + i*frame_stride; its plane set is
 * written at planes + i*rh264_planes_bytes(cw, ch, pad).  cw, ch multiples of
 * 16; pad a multiple of 4 in [4, 128] (the reference clamps RFF_HPEL_PAD to
 * 0..128, default 32). */
size_t rh264_planes_bytes(int cw, int ch, int pad);
int rh264_build_planes(rh264_ctx *ctx, const uint8_t *frames, size_t frame_stride, int nframes,
                       int cw, int ch, int pad, uint8_t *planes);

/* ---- motion search --------------------------------------------------------
 * Search parameters (the knobs the reference reads from FrameEncoder fields and
 * RFF_* variables, made explicit).  lr_fp / lr are lambda-rate tables:
 * lr[b] = (int64)(lambda * b) truncated exactly as mb16.rs:2563 computes
 * `(lam_r * rate as f64) as i64` (rate = mvbits(dx) + mvbits(dy), mb16.rs:
 * 2474-2475); lr_fp prices the SAD full-pel phase (B2, lambda * RFF_ME_SADL,
 * mb16.rs:2582-2590, 756-763), lr the SATD phase.  Use rh264_lambda_table().
 * Entries lie in 0..2^26 (else the searches fail with an error). */
typedef struct {
    int32_t cw, ch, pad;   /* frame size (multiples of 16) and plane pad         */
    int32_t bw, bh;        /* partition: 16x16, 16x8, 8x16, 8x8, 8x4, 4x8, 4x4   */
    int32_t range;         /* full-search radius R in full pels, 1..64           */
    int32_t subpel_pat;    /* 0 ring8+iterate, 1 ring4+iterate, 2 ring8 single,
                              3 ring4 single, -1 no rings (SATD reprice only),
                              4 quarter-pel window: the 25 positions within +-2
                              quarter pels of the start, cheapest wins, ties to
                              the start, then raster order (dy, then dx); no
                              reference counterpart, the learned search's
                              refinement                                          */
    int32_t lr_fp[RH264_LR_N];
    int32_t lr[RH264_LR_N];
    int32_t hadamard;      /* SATD transform of the sub-pel stage: 4 (or 0) =
                              the reference's 4x4 sums; 8 = the 8x8 Hadamard,
                              sum over 8x8 blocks of sum|H8 d H8| / 2 (no
                              reference counterpart; 16x16 blocks only, other
                              shapes decline)                                     */
} rh264_search_params;

void rh264_lambda_table(double lambda, int32_t *lr /* RH264_LR_N */);

/* Plain exhaustive integer search (build-defined; the reference's exhaustive
 * grids are mb16.rs:2778-2889 and 2897-2955): for every bw x bh block of each
 * of nframes frames, cost(mv) = SAD(mv) + lr_fp[rate(mv - mvp)] over
 * mv = centre + 4*(dx,dy), dx,dy in [-R,R], raster order (dy outer, dx inner),
 * first minimum kept (strict <, mb16.rs:2843-2855).  Frame i: cur + i*cur_stride,
 * plane set planes + i*planes_stride.  centres / mvps: per-block (x,y) qpel
 * pairs (centres multiples of 4), frame-major, or NULL for (0,0).  Outputs per
 * block: int_mv (x,y), int_cost.  Reference positions outside the padded planes
 * read as infinite edge replication (mc_luma's clamp, inter.rs:120-124). */
int rh264_full_search(rh264_ctx *ctx, const rh264_search_params *p, int nframes,
                      const uint8_t *cur, size_t cur_stride, const uint8_t *planes,
                      size_t planes_stride, const int32_t *centres, const int32_t *mvps,
                      int32_t *int_mv, int32_t *int_cost);

/* Sub-pel refinement (mb16.rs:2956-3291, cascade ring walk, steps 2 then 1,
 * ring8 order (s,0),(-s,0),(0,s),(0,-s),(s,s),(-s,-s),(s,-s),(-s,s)) of every
 * block from start_mv, cost(mv) = SATD(mv) + lr[rate(mv - mvp)] with SATD the
 * encoder-domain sum|H d| of mb16.rs:6027-6074 over the qpel prediction of
 * inter.rs:1015-1105 (MeCtx::eval, crates/rusty_h264-accel/src/mectx.rs:112-166).
 * Outputs per block: mv, cost, and (optional, may be NULL) evals = number of
 * candidate evaluations. */
int rh264_subpel_refine(rh264_ctx *ctx, const rh264_search_params *p, int nframes,
                        const uint8_t *cur, size_t cur_stride, const uint8_t *planes,
                        size_t planes_stride, const int32_t *start_mv, const int32_t *mvps,
                        int32_t *mv, int32_t *cost, int32_t *evals);

/* Plain search pipeline = rh264_full_search then rh264_subpel_refine from its
 * result, two launches for the whole batch.  int_mv / int_cost may be NULL. */
int rh264_plain_search(rh264_ctx *ctx, const rh264_search_params *p, int nframes,
                       const uint8_t *cur, size_t cur_stride, const uint8_t *planes,
                       size_t planes_stride, const int32_t *centres, const int32_t *mvps,
                       int32_t *int_mv, int32_t *int_cost, int32_t *mv, int32_t *cost);

/* rh264_plain_search that allocates nothing (no counterpart in the reference: it is
 * for callers that cannot allocate device memory, such as code under HIP graph
 * capture or a serving loop).  Same arguments, results and profiling phases as
 * rh264_plain_search; its argmin keys live in a static device array of
 * RH264_NOALLOC_KEYS keys (2 MiB) that each device holds until the process ends,
 * cleared by a fill kernel at the start of every call.  Every call on a device
 * shares that device's array, so calls that overlap on one device (two contexts or
 * two host threads) race.  Errors when nframes x blocks per frame exceeds
 * RH264_NOALLOC_KEYS. */
#define RH264_NOALLOC_KEYS (1u << 18)
int rh264_plain_search_noalloc(rh264_ctx *ctx, const rh264_search_params *p, int nframes, const uint8_t *cur,
                               size_t cur_stride, const uint8_t *planes, size_t planes_stride, const int32_t *centres,
                               const int32_t *mvps, int32_t *int_mv, int32_t *int_cost, int32_t *mv, int32_t *cost);

/* Stage timing of rh264_plain_search and rh264_learned_search (no reference
 * counterpart; the reference's stage profiler is
 * crates/rusty_h264-common/src/prof.rs).  rh264_set_profiling(capacity > 0) makes
 * the next `capacity` searches record HIP events at their phase boundaries on the
 * context's stream, without any host synchronisation.  Phases: plain = integer
 * search, refinement; learned = network input preparation, network, candidate
 * search (the local searches and their selection), fall-back full search (with its
 * merge), refinement.  rh264_profile_get(i) returns search i's time before its last
 * phase and that phase's time; rh264_profile_phases(i) writes up to `cap` phase
 * durations (ms) and their count; rh264_profile_times(i) writes each phase's end
 * time after the search's start (ms) and their count.  capacity 0 turns profiling
 * off and resets the count. */
int rh264_set_profiling(rh264_ctx *ctx, int capacity);
int rh264_profile_count(rh264_ctx *ctx);
int rh264_profile_get(rh264_ctx *ctx, int i, float *search_ms, float *refine_ms);
int rh264_profile_phases(rh264_ctx *ctx, int i, float *ms, int cap, int *nphases);
int rh264_profile_times(rh264_ctx *ctx, int i, float *ms, int cap, int *nevents);

/* Kernel-variant selection (no reference counterpart; set by the autotuner).
 * Knobs: "fs_kernel" 0 = quad-SAD (v_qsad_pk_u16_u8) accumulator ring
 * (default), 1 = v_sad_u8 ring with 2 candidates per lane, 2 = v_sad_u8 ring
 * with 4 candidates per lane; "refine_lds" 1 (default) = the refinement of
 * 16x16/16x8/8x16 blocks stages each block's +-3 pel plane window in LDS, 0 =
 * it reads the plane sets in HBM; "refine_mfma" 1 (default) = the refinement of
 * 16x16 blocks computes its Hadamards on the matrix cores (one wave per block),
 * 0 = the lane-per-4x4 kernel above; "cand_quad" 2 (default) = the learned search's
 * candidate windows (local range 1..4) take one lane per candidate, quad-SADs for
 * column groups of four and v_sad_u8 for the rest, 1 = quad-SAD column groups
 * (local range <= 3), 0 = v_sad_u8 lanes; "cand_lds" 1 (default) = that
 * lane-per-candidate kernel reads each block's source rows from LDS where it uses
 * them, 0 = it holds them in registers;
 * "fallback_v" 1 (default) = its fall-back full search uses the v_sad_u8 x2
 * variant, 2 = v_sad_u8 with one candidate per lane, 0 = quad-SADs; "fallback_wgs"
 * (1..65536, default 2048) = that search's grid cap in workgroups (they loop over a
 * longer list); "fallback_overlap" 1 = with the +-2 quarter-pel window refinement,
 * the fall-back search runs on a second stream of the context beside the refinement
 * of the other blocks and its blocks are refined after it (the context's stream
 * waits for it), 0 (default) = it runs on the context's stream before one
 * refinement of every block (both kernels are throughput-bound: side by side they
 * take longer than in turn); "planes_rw" 1 (default) = the plane build slides a
 * register window down 256x32 tiles, 0 = the LDS-tile kernel; "mvnet_fused" 2
 * (default) = network inference in one fused kernel per tile of 12x24 MBs holding the
 * activations in registers, 1 = the fused kernel with the activations in LDS (tiles of
 * 8x8 MBs), 0 = layer by layer; "decide_mode" of
 * rh264_motion_search_wavefront: 0 = one launch per anti-diagonal, 1 = fixed-point
 * passes (each re-searches the macroblocks whose predictors changed; the stream
 * is synchronised every 8 passes to test for the fixed point), 2 (default) = 1
 * for one or two references, else 0; "la_scaled" of
 * rh264_mbtree_costs_multi_learned (see there; it changes the proposals, so the
 * results).  Every other search variant returns identical results; the two
 * network paths agree up to FP32 summation order. */
int rh264_tune(rh264_ctx *ctx, const char *knob, int value);

/* Quarter-pel prediction of blocks (mc_luma, inter.rs:1193-1262, served from
 * the plane set as hpel_block inter.rs:1107-1162 does, any MV): block i of
 * size bw x bh at (x[i], y[i]) with mv[2i..2i+1] is written to out + i*bw*bh. */
int rh264_predict_blocks(rh264_ctx *ctx, const uint8_t *planes, int cw, int ch, int pad, int bw,
                         int bh, int n, const int32_t *xy, const int32_t *mv, uint8_t *out);

/* Reference-mode motion search of every p->bw x p->bh block (motion_search,
 * mb16.rs:2433-3302, with every knob explicit; the same steps as the oracle's
 * orc_motion_search_ex).  Per block: seeds (0,0) then the block's predictors
 * preds[(frame*nblk + b)*npreds + i] (quarter-pel pairs; strict <), preds[0] pricing
 * the rate; the integer snap ((v + 2) div_euclid 4) * 4 of a fractional winner;
 * the diamond rungs dc->steps (quarter pels, each walked until it stops improving):
 * with dc->fc, on a full-pel centre whose +-step box lies inside the frame, for
 * 16x16/16x8/8x16/8x8 blocks, the fixed-centre argmin of the 4-point ring
 * (mb16.rs:2646-2706), otherwise one cascade pass; with dc->sadfp the seed and
 * diamond phases price SAD and the winner and seed are repriced (mb16.rs:2739-2742);
 * with dc->rescue the me_wide grid (mb16.rs:2778-2889); then the sub-pel stage
 * p->subpel_pat; with dc->snap the seed is kept if it is cheaper (mb16.rs:3297-3300).
 * Rate and prices (mb16.rs:2453-2565): mv_cost 0 = Exp-Golomb lengths priced by
 * p->lr (cost closure) and p->lr_fp (B2 phase); mv_cost 1 = per-|d| bits
 * mv_tab[min(|d|, 4095)] (the smooth or measured tables of mb16.rs:505-589, in
 * quarter bits) priced (int64)(l * rate) with l = lambda / 4 in the cost closure,
 * lambda_fp in the B2 phase, and lambda in the reference's batched evaluations (the
 * 16-wide fixed-centre x4 batch outside B2, mb16.rs:2680-2683; the 16x16 rescue
 * grid inside the frame, mb16.rs:2834; the rescue's distortion estimate).
 * npreds 0..8, nsteps 0..8; shapes 16x16, 16x8, 8x16, 8x8, 8x4, 4x8, 4x4 (the 8x8
 * transform: 16x16 only, others decline).  events (may be NULL): per block, bit 0 =
 * the rescue grid ran, bit 1 = it cut the cost to <= 15/16 (the payoff the
 * reference's frame-level learning window counts, mb16.rs:2870-2886). */
typedef struct {
    int32_t steps[8];      /* diamond rungs in quarter pels, e.g. 16, 8, 4 (Fast: 16, 4) */
    int32_t nsteps;
    int32_t sadfp;         /* 1: B2, SAD + lr_fp for the seed and diamond phases */
    int32_t snap;          /* 1: integer snap of a fractional seed, seed kept if cheaper */
    int32_t fc;            /* 1: fixed-centre passes (me_fc, the reference's default with asm) */
    int32_t sad_only;      /* 1: the Fast preset's cost closure, SAD (mb16.rs:2512); no fc, no rescue */
    int32_t mv_cost;       /* 0: Exp-Golomb with p->lr / p->lr_fp; 1: mv_tab with the lambdas */
    int32_t rescue;        /* 1: the me_wide rescue grid */
    int32_t rescue_range;  /* me_range, full pels (default 24), 0..64 */
    int32_t rescue_thresh; /* me_rescue (default 3): distortion per pixel above which it runs */
    int32_t rescue_fast;   /* me_fast (default 1): skip the flat-block test */
    int32_t rescue_var;    /* me_wide_var (default 800): the flat test's variance bound */
    double lambda;         /* mv_cost 1: lambda_me */
    double lambda_fp;      /* mv_cost 1: the B2 lambda, lambda_me * RFF_ME_SADL */
    const uint16_t *mv_tab; /* mv_cost 1: 4096 entries, device memory */
    int32_t rescue_learn;  /* me_learn (default 40): rh264_motion_search_wavefront's learning window */
    int32_t rescue_payoff; /* me_payoff_pct (default 15), percent */
} rh264_diamond;
int rh264_motion_search(rh264_ctx *ctx, const rh264_search_params *p, const rh264_diamond *dc, int nframes,
                        const uint8_t *cur, size_t cur_stride, const uint8_t *planes, size_t planes_stride,
                        const int32_t *preds, int npreds, int32_t *mv, int32_t *cost, int32_t *events);
/* The smooth mvd rate table (mb16.rs:507-516, mv_cost_kind 1): tab[a] = round(4 *
 * (2 log2(a + 1) + 0.718 + (a != 0))), a = 0..4095, into host memory. */
void rh264_mv_cost_smooth(uint16_t *tab /* 4096 */);

/* The encoder's per-macroblock 16x16 decision in macroblock order, for P frames
 * whose every macroblock takes its P_L0_16x16 best_part result (mb16.rs:4921-4959
 * with no extra seeds, as called at mb16.rs:5491): for each reference r < nrefs,
 * the search of rh264_motion_search with the single predictor predict_mv
 * (inter.rs:36-73) against r of the left, above and above-right (else above-left)
 * neighbours' decided vectors and references (mv_neighbors, mb16.rs:1813-1836);
 * cost + (int64)(lambda_ref * ref_bits(r, nrefs)) (mb16.rs:5967-5980), the first
 * minimum in reference order.  planes holds nframes + nrefs - 1 plane sets of
 * consecutive frames, planes_stride bytes apart; frame f's reference r (0 = the
 * most recent) is plane set f + nrefs - 1 - r.  All frames are decided at once,
 * by one of two exact schedules (rh264_tune "decide_mode"): one launch per
 * anti-diagonal x + 2y = t, whose macroblocks' neighbours lie on earlier ones; or
 * fixed-point passes that re-search the macroblocks whose predictors changed
 * until none does (the encoder-order result is the only fixed point; this
 * schedule synchronises the stream).
 * Outputs per macroblock: ref, mv (pair) and cost, int32.  16x16 only; nrefs
 * 1..4.  The search knobs are rh264_motion_search's.  With dc->rescue the frame's
 * learning window applies (mb16.rs:2868-2886): the first dc->rescue_learn searches
 * of the frame (macroblock, then reference order) whose rescue grid ran are
 * tallied, with those that cut their cost to <= 15/16; a window that fills under
 * dc->rescue_payoff % of them turns the rescue off for every later search of the
 * frame.  A search's switch depends only on earlier searches' events, so the
 * fixed-point schedule (always used with the rescue) also re-searches the
 * macroblocks whose switch changed, from a per-frame scan of the events each pass. */
int rh264_motion_search_wavefront(rh264_ctx *ctx, const rh264_search_params *p, const rh264_diamond *dc,
                                  int nframes, int nrefs, double lambda_ref, const uint8_t *cur, size_t cur_stride,
                                  const uint8_t *planes, size_t planes_stride, int32_t *ref, int32_t *mv,
                                  int32_t *cost);

/* The P macroblock decision with partitions, in macroblock order (the default CAVLC
 * path, mb16.rs:5520-5567; mb16.rs:8849-9010 on the CABAC path without its sub-8x8
 * arm; the shape picked in the SATD currency: shape RD off, the all-8x8 P_8x8 arm,
 * every macroblock inter).  Per macroblock: the 16x16 best_part (mb16.rs:4921-4959)
 * as rh264_motion_search_wavefront, predictors predict_mv (inter.rs:36-73) of the
 * neighbours read on the 4x4 motion grid (mv_neighbors_block, mb16.rs:1871-1888: the
 * left macroblock's top-right quadrant, the bottom-left quadrant of the above and
 * above-right ones, else the above-left's bottom-right); when its cost exceeds
 * split_gate (rh264_split_gate(qp), mb16.rs:8866-8868), best_part of the two 16x8
 * halves, the two 8x16 halves and with p8x8 the four 8x8 quadrants, each with the
 * macroblock's neighbours and the 16x16 vector as the extra seed (mb16.rs:
 * 8893-8948); the first strict minimum of c16, ct + cb, cl + cr and
 * (int64)(lambda_ref * 4) + the quadrants' sum.  lambda_ref = the search lambda
 * (it prices ref_bits and the P_8x8 syntax).  The search knobs are
 * rh264_motion_search's (16x16 in p; the 8x8 transform declines); with dc->rescue
 * the frame's learning window of rh264_motion_search_wavefront counts every search in
 * call order: the 16x16 references, then each partition's (16x8 top, bottom, 8x16
 * left, right, 8x8 TL, TR, BL, BR) references.
 * planes as rh264_motion_search_wavefront.  Decided by fixed-point passes (the
 * stream is synchronised every 8 passes).  rec: 14 int32 per macroblock, {mode (0
 * 16x16, 1 16x8, 2 8x16, 3 8x8), cost, then ref, mvx, mvy of each quadrant TL, TR,
 * BL, BR}. */
int rh264_partition_decision(rh264_ctx *ctx, const rh264_search_params *p, const rh264_diamond *dc, int nframes,
                             int nrefs, double lambda_ref, int32_t split_gate, int p8x8, const uint8_t *cur,
                             size_t cur_stride, const uint8_t *planes, size_t planes_stride, int32_t *rec);
/* ((30 * (qstep16 + 160)) >> 3) * 2, qstep16 = {10, 11, 13, 14, 16, 18}[qp % 6] <<
 * (qp / 6) (mb16.rs:8866-8868); -1 outside 0..51 */
int32_t rh264_split_gate(int qp);

/* The luma predictions of decided P frames (the step after the decision): per
 * macroblock the P_Skip vector skip_mv (skip_mv, mb16.rs:1841-1853: (0,0) when the
 * left or above neighbour is unavailable or is a zero vector of reference 0, else
 * predict_mv(A, B, C, 0), inter.rs:36-73), pred = the motion-compensated 16x16
 * prediction at the decided (ref, mv), and skip_pred = the P_Skip prediction from
 * reference 0 at skip_mv (mc_luma, inter.rs:1193-1262, as mc_luma_cached serves it
 * from the plane sets, mb16.rs:4514-4541; any vector).  ref/mv as
 * rh264_motion_search_wavefront writes them (ref values outside 0..nrefs-1 are
 * clamped into it), planes in its layout.  pred and skip_pred: nframes frames of
 * cw x ch bytes, cw*ch apart; skip_mv: int32 pairs per macroblock. */
int rh264_decision_predict(rh264_ctx *ctx, int cw, int ch, int pad, int nframes, int nrefs, const uint8_t *planes,
                           size_t planes_stride, const int32_t *ref, const int32_t *mv, int32_t *skip_mv,
                           uint8_t *pred, uint8_t *skip_pred);

/* The mb-tree QP-offset policy of a GOP (gop_qp_offsets, crates/rusty_h264-encoder/
 * src/mbtree.rs:369-593) from its frame costs, on the HOST (host pointers; the f64
 * propagation is order-dependent and stays in the reference's order): f64 backward
 * propagation (frames descending, raster order, mbtree.rs:268-297, 440-455); the
 * predictability back-off (all zero when 1 - mean((intra-inter)/intra) over frames
 * 1.. is below res_min > 0); offsets -strength*log2((intra + propagated)/intra)
 * centred on the GOP mean; the differentiation latch (all zero when the RMS of the
 * centred offsets / strength is below sd_min); round half away from zero, clamp to
 * +-6; with grain set (the grain latch of mbtree.rs:401-420, from
 * rh264_grain_signature below) and nframes >= 2 every offset is 0 and so are the
 * stats.  intra/inter [nframes][mb_w*mb_h] with 1 <= intra, 0 <= inter <= intra (as
 * rh264_mbtree_costs writes them), mv pairs alike; offsets the same shape; stats
 * (may be NULL) = {residual fraction, spread, effective strength}.  The reference's
 * defaults: strength 0.9, res_min 0.03 (RFF_MBTREE_RESMIN), sd_min = the
 * configuration's mbtree_spread_min, 0.0 by default (config.rs:566). */
int rh264_mbtree_qp_offsets(int nframes, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                            const int32_t *mv, double strength, double res_min, double sd_min, int grain,
                            int32_t *offsets, double *stats);
/* The grain signature that feeds `grain` above (FrameSignals::grain_signature,
 * crates/rusty_h264-encoder/src/signals.rs:528-533, of the GOP's coded-size frames 1
 * and 0 at mbtree.rs:404-406), on the HOST (host pointers, cw x ch luma, multiples of
 * 16): *grain = 1 when the median macroblock variance (signals.rs:247-266) is below
 * 200, the lower-quartile zero-motion residual per pixel (grain_floor, 311-338) is
 * above 5 and the +-8 full-pel grid's mean SAD gain (b2_mgain, 61-129) is below
 * 0.1.  stats (may be NULL) = those three values. */
int rh264_grain_signature(const uint8_t *cur, const uint8_t *ref, int cw, int ch, int32_t *grain, double *stats);

/* ---- batched multi-stream search ------------------------------------------
 * Many streams of uneven size, 1..4 references each, several partition shapes,
 * searched together: for every shape, one integer-search launch over all
 * (stream, reference) pairs (work-queue persistent kernel or one tile per block,
 * chosen by rh264_batch_autotune), one refinement launch, and one launch that
 * picks each block's reference as best_part does (cost + (int64)(lambda_ref *
 * ref_bits(r, nrefs)), first minimum in reference order; mb16.rs:4921-4959,
 * 5967-5980).  Each (stream, reference, shape) is the plain search above (zero
 * centres).  A batch is prepared once (tables uploaded, scratch allocated), then
 * run any number of times, directly or as a captured HIP graph.
 *
 * When the batch holds the 4x4 shape and at least one other of 16x16, 16x8, 8x16,
 * 8x8, every stream is whole macroblocks (cw, ch multiples of 16) and every lr_fp
 * entry is at most 2^21, those shapes share one integer-search launch instead: it
 * computes each 4x4 block's SADs once per candidate and adds them into the larger
 * partitions' SADs, pricing every partition with its own predictor.  The records are
 * the same as with one search per shape.
 *
 * Per stream: planes = nrefs plane sets back to back (reference 0 first); mvps =
 * per-shape block predictors in shape order (rh264_batch_mvp_len int32 pairs
 * total) or NULL; out = per-shape records {ref, mvx, mvy, cost} int32, shapes in
 * order (rh264_batch_out_len int32 words). */
typedef struct {
    const uint8_t *cur;
    const uint8_t *planes;
    const int32_t *mvps;
    int32_t cw, ch, nrefs;
    int32_t *out;
} rh264_stream;
typedef struct rh264_batch rh264_batch;
size_t rh264_batch_out_len(int cw, int ch, const int32_t *shapes, int nshapes);
size_t rh264_batch_mvp_len(int cw, int ch, const int32_t *shapes, int nshapes);
int rh264_batch_create(rh264_ctx *ctx, const rh264_stream *streams, int nstreams, const int32_t *shapes,
                       int nshapes, int pad, int range, int subpel_pat, const int32_t *lr_fp,
                       const int32_t *lr, double lambda_ref, rh264_batch **out);
int rh264_batch_run(rh264_ctx *ctx, rh264_batch *b);
/* The integer searches of one run alone (no refinement, no reference choice; the
 * records are not written): for timing that phase.  No reference counterpart. */
int rh264_batch_search(rh264_ctx *ctx, rh264_batch *b);
/* Times every kernel variant x scheduling choice of each shape's integer search
 * (reps runs each) and keeps the fastest; results are identical for all.  Shapes
 * that search fused are left out: rh264_batch_choice reports variant 3 for them. */
int rh264_batch_autotune(rh264_ctx *ctx, rh264_batch *b, int reps);
int rh264_batch_choice(rh264_batch *b, int shape_index, int *variant, int *persist);
/* Capture one rh264_batch_run into a HIP graph; replay launches it. */
int rh264_batch_capture(rh264_ctx *ctx, rh264_batch *b);
int rh264_batch_replay(rh264_ctx *ctx, rh264_batch *b);
int rh264_batch_destroy(rh264_batch *b);

/* ---- learned motion-vector proposal (no reference counterpart) -------------
 * A small BF16 CNN (rusty_h264_amd/csrc/rh264_mvnet.hip: 1/4-res current+reference
 * luma, conv3x3/2 2->32, 3 x conv3x3 32->32 on MFMA, 2x2 pool, 1x1 head to 289
 * MV bins of 4 pels) proposes 8 integer-MV centres per 16x16 MB.  Parameters are
 * FP32 (rh264_mvnet_param_count of them, layout in that file).
 * rh264_mvnet_forward: logits [n][mb][289] (device, may be NULL) and/or top8
 *   [n][mb][8] bin indices (device, 16-byte aligned, may be NULL) for frames
 *   cur/ref (stride apart); by default one fused kernel per tile of 12x24 MBs
 *   (rh264_tune "mvnet_fused" 1 selects the LDS-tile fused kernel, 0 the
 *   layer-by-layer kernels of training).
 *   The four conv layers use BF16 weights and activations, FP32 accumulation.
 * rh264_mvnet_backward: forward + backward of the mean cross-entropy against
 *   labels [n][mb] (bin indices); gradients stay on the device (get_grads), the
 *   loss is returned on the host.  rh264_mvnet_adam: one Adam step. */
typedef struct rh264_mvnet rh264_mvnet;
size_t rh264_mvnet_param_count(void);
int rh264_mvnet_create(rh264_ctx *ctx, int max_n, int max_w, int max_h, rh264_mvnet **out);
int rh264_mvnet_destroy(rh264_mvnet *net);
int rh264_mvnet_set_params(rh264_ctx *ctx, rh264_mvnet *net, const float *host, size_t count);
int rh264_mvnet_get_params(rh264_ctx *ctx, rh264_mvnet *net, float *host, size_t count);
int rh264_mvnet_get_grads(rh264_ctx *ctx, rh264_mvnet *net, float *host, size_t count);
int rh264_mvnet_forward(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *cur, const uint8_t *ref,
                        size_t stride, int n, int w, int h, float *logits, int32_t *top8);
int rh264_mvnet_backward(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *cur, const uint8_t *ref,
                         size_t stride, int n, int w, int h, const int32_t *labels, float *loss_out);
int rh264_mvnet_adam(rh264_ctx *ctx, rh264_mvnet *net, float lr);

/* Learned-proposal search of every 16x16 block of nframes frames (frame i: cur +
 * i*frame_stride against ref + i*frame_stride, plane set planes + i*planes_stride):
 * the network's 8 centres are re-costed exactly (the integer search above with
 * radius local_range around each centre); the cheapest (first minimum in rank
 * order) starts the sub-pel refinement (p->subpel_pat: a ring walk, or 4, the
 * +-2 quarter-pel window, which is what the proposal pairs with); a block whose best integer
 * cost exceeds fallback_cost is searched exhaustively with radius p->range
 * instead (fallback[i] = 1).  Costs, rate tables and tie-breaks are the plain
 * search's, so every evaluated candidate is priced exactly. */
int rh264_learned_search(rh264_ctx *ctx, const rh264_search_params *p, rh264_mvnet *net, int nframes,
                         const uint8_t *cur, const uint8_t *ref, size_t frame_stride,
                         const uint8_t *planes, size_t planes_stride, int local_range,
                         int32_t fallback_cost, int32_t *mv, int32_t *cost, int32_t *fallback);

/* ---- mb-tree lookahead (crates/rusty_h264-encoder/src/mbtree.rs, HalfRes,
 *      the default LookaheadMode, config.rs:568) ------------------------------
 * rh264_mbtree_downsample: downsample2x (mbtree.rs:61-74) of nframes cw x ch
 *   luma frames (frame i at frames + i*frame_stride) into half, frame i at
 *   half + i*(cw/2)*(ch/2).
 * rh264_mbtree_costs: frame_costs (mbtree.rs:208-263) for a GOP of nframes
 *   half-res frames: frame 0 is the IDR (inter = intra, mv = 0); frame f > 0 is
 *   searched against frame f-1 with inter_cost's seeded cascade diamond
 *   (mbtree.rs:154-197, steps 8,4,2,1 full pel, seed = left neighbour's vector or,
 *   in column 0, the one above, halved), 8x8 blocks; intra_cost (78-93) is the AC
 *   SATD.  Outputs per (frame, MB): intra, inter = min(inter, intra), mv
 *   (full-res quarter pel, x2 as in mbtree.rs:249).  mb_w x mb_h macroblocks.
 * rh264_mbtree_propagate_i32: build-defined integer form of the backward
 *   propagation (mbtree.rs:268-297, 440-455): prop[n-1] = 0; for f = n-1..1,
 *   amount = ((intra + prop[f]) * (intra - inter)) / intra (int64), each of the
 *   up to four overlapped MBs of frame f-1 receives (amount * wx * wy) >> 8;
 *   sums are exact in int64 and saturated to INT32_MAX when stored in prop. */
int rh264_mbtree_downsample(rh264_ctx *ctx, const uint8_t *frames, size_t frame_stride, int nframes,
                            int cw, int ch, uint8_t *half);
int rh264_mbtree_costs(rh264_ctx *ctx, const uint8_t *half, int nframes, int mb_w, int mb_h,
                       int32_t *intra, int32_t *inter, int32_t *mv);
int rh264_mbtree_propagate_i32(rh264_ctx *ctx, int nframes, int mb_w, int mb_h, const int32_t *intra,
                               const int32_t *inter, const int32_t *mv, int32_t *prop);

/* Multi-reference mb-tree costs (build-defined, no reference counterpart: the 4K
 * lookahead configuration's "3 references, +-64" form; the reference's inter_cost,
 * mbtree.rs:154-197, walks a seeded diamond against one reference).  half as for
 * rh264_mbtree_costs: nframes frames of mb_w*8 x mb_h*8.  8x8 blocks, one per MB:
 *   intra = max(1, (sum of the |AC coefficients| of H8 x H8) >> 1);
 *   frame 0: inter = intra, mv = (0,0), ref = 0;
 *   frame f > 0, for each reference k = 1..min(nrefs, f) (frame f-k): the
 *     exhaustive full-pel SAD search over +-range (reads clamped into the frame,
 *     raster first minimum, dy outer), then the 8x8 SATD sum|H8 d H8|/2 at its
 *     winner; the cheapest k wins (ties to the smaller): inter = min(its SATD,
 *     intra), mv = 8 * winner (full-resolution quarter pel), ref = k.
 * nrefs 1..4, range 1..64; other values decline.  Scratch: padded copies of the
 * frames, held by the context. */
int rh264_mbtree_costs_multi(rh264_ctx *ctx, const uint8_t *half, int nframes, int mb_w, int mb_h, int nrefs,
                             int range, int32_t *intra, int32_t *inter, int32_t *mv, int32_t *ref);
/* The learned-seeded form of the above (build-defined; the 4K configuration's
 * "learned proposal seeding the inter search").  frames: the nframes full-resolution
 * cw x ch frames (stride frame_stride) that half was downsampled from.  The network
 * (net, created for at least nframes - 1 pairs of cw x ch) proposes 8 vectors per
 * 16x16 MB, i.e. per 8x8 half-res block, for reference k of frame f: with
 * rh264_tune "la_scaled" 1 (the default) its proposals for the pair (frame f,
 * frame f - 1) times k, one network run in all; with 0 its proposals for the pair
 * (frame f, frame f - k), one run per distance.  A proposal's centre (2 half-res
 * pels per bin, times k when scaled) is clamped to +-(P - 8 - local_range), P =
 * (max(range, 16 + local_range) + 11) & ~3.  Each is searched exhaustively over
 * +-local_range half-res pels (1..4), the first minimum in proposal order (then
 * raster order) winning; a best SAD above fallback_sad falls back to the exhaustive
 * +-range search.  Costs,
 * reference choice and outputs as rh264_mbtree_costs_multi; fallback[f][mb] = bit
 * k-1 set when reference k fell back (device, may not be NULL). */
int rh264_mbtree_costs_multi_learned(rh264_ctx *ctx, rh264_mvnet *net, const uint8_t *frames, size_t frame_stride,
                                     const uint8_t *half, int nframes, int cw, int ch, int nrefs, int range,
                                     int local_range, int32_t fallback_sad, int32_t *intra, int32_t *inter,
                                     int32_t *mv, int32_t *ref, int32_t *fallback);
/* rh264_mbtree_propagate_i32 with frame f's amount going to frame f - ref[f][mb]
 * (the ref output of rh264_mbtree_costs_multi); ref all 1 is the single-reference
 * propagation. */
int rh264_mbtree_propagate_multi_i32(rh264_ctx *ctx, int nframes, int mb_w, int mb_h, const int32_t *intra,
                                     const int32_t *inter, const int32_t *mv, const int32_t *ref, int32_t *prop);

/* Rate-control frame complexity, lookahead::complexity (lookahead.rs:38-101), of
 * nframes coded-size cw x ch luma frames (frame i at cur + i*frame_stride; cw, ch
 * multiples of 16): out[i] = max(total, 1) as a double, where total sums over
 * the macroblocks
 *  - ref == NULL (IDR): intra_activity (55-72), the AC part (coefficient 0
 *    excluded) of every 4x4 Hadamard of the pixels;
 *  - otherwise: inter_activity (76-101), the minimum over the full-pel candidates
 *    (0,0), (+-1,0), (0,+-1), (+-2,0), (0,+-2) of the residual's sum |H d| against
 *    frame i of ref (ref + i*frame_stride), read with coordinates clamped into the
 *    frame (mc_luma, inter.rs:120-124).
 * out: nframes doubles in device memory. */
int rh264_rc_complexity(rh264_ctx *ctx, const uint8_t *cur, const uint8_t *ref, size_t frame_stride, int nframes,
                        int cw, int ch, double *out);

/* ---- The bitstream of decided frames (host code, no GPU) ----
 * H.264 Annex-B bytes in the reference's CAVLC configuration (Constrained Baseline,
 * profile 66 with constraint_set1, CAVLC, one slice per picture, deblocking signalled
 * on; params.rs:31-172, slice.rs:49-107, nal.rs:78-117).  Each entry writes into out
 * (cap bytes), sets *len to the bytes the unit needs and returns 0, or returns < 0
 * when they do not fit (and for malformed arguments).
 * rh264_write_headers: SPS + PPS (Sps/Pps::write_rbsp, params.rs:56-101, 148-172):
 *   mb grid ceil(width/16) x ceil(height/16) with cropping, max_num_ref_frames and the
 *   PPS default active count num_ref_frames, pic_init_qp qp.
 * rh264_write_idr_pcm: the IDR picture (write_idr_slice_header, slice.rs:87-107) of
 *   I_PCM macroblocks: y is (16 mb_w) x (16 mb_h) bytes, u and v (8 mb_w) x (8 mb_h).
 * rh264_write_p_slice: a P picture (write_p_slice_header, slice.rs:51-85;
 *   frame_num and poc_lsb are written mod 16; nrefs active references, signalled when
 *   num_ref_frames > 1) from rh264_partition_decision's records (14 int32 per
 *   macroblock: mode 0..3, cost, ref/mvx/mvy of the TL, TR, BL, BR quadrants; the
 *   16x16 decision's (ref, mv) is mode 0 with four equal quadrants).  No residual is
 *   coded: a 16x16 reference-0 macroblock at its P_Skip vector (8.4.1.1) is skipped,
 *   every other is P_L0_16x16 / P_L0_L0_16x8 / P_L0_L0_8x16 / P_8x8 (four P_L0_8x8)
 *   with ref_idx (mb16.rs:5957-5964) and mvd against the clause 8.4.1.3 predictors,
 *   coded_block_pattern 0 (mb16.rs:4416-4435; cavlc.rs:136-138). */
typedef struct rh264_seq {
    int32_t width, height;  /* luma size; the macroblock grid rounds up */
    int32_t qp;             /* pic_init_qp; every slice has slice_qp_delta 0 */
    int32_t num_ref_frames; /* 1..16 */
    int32_t level_idc;      /* 30 is the reference's default (config.rs:463) */
} rh264_seq;
int rh264_write_headers(const rh264_seq *s, uint8_t *out, size_t cap, size_t *len);
int rh264_write_idr_pcm(const rh264_seq *s, const uint8_t *y, const uint8_t *u, const uint8_t *v, uint8_t *out,
                        size_t cap, size_t *len);
int rh264_write_p_slice(const rh264_seq *s, int frame_num, int poc_lsb, int nrefs, const int32_t *rec, uint8_t *out,
                        size_t cap, size_t *len);
/* rh264_write_p_slice with the luma residual: src is the current picture's luma and
 * pred its inter prediction at the records' vectors (both (16 mb_w) x (16 mb_h) host
 * bytes, stride 16 mb_w).  Each 4x4 block of src - pred goes through the forward core
 * transform and the inter quantiser at s->qp ((|W| MF + 2^qbits / 6) >> qbits,
 * transform.rs:92, 121, 172-174); a macroblock's coded_block_pattern has the bits of its
 * 8x8 quadrants with a nonzero level (inter me(v) mapping, cavlc.rs:94-98, 136-138; no
 * chroma residual), mb_qp_delta 0, and each coded block is residual_block_cavlc
 * (cavlc.rs: coeff_token with the nC context of its left and upper blocks, trailing-one
 * signs, levels with the adaptive suffix length, total_zeros, run_before; H.264
 * Tables 9-5, 9-7, 9-8, 9-10).  A 16x16 reference-0 macroblock at its P_Skip vector
 * with no nonzero level is skipped.  recon (same layout) receives the decoder's luma:
 * pred + the scaled, inverse-transformed residual (8.5.12), clipped to 0..255; the
 * slice header signals disable_deblocking_filter_idc 1 (the reference's writes 0 and
 * deblocks its reconstruction; this writer does not deblock), so recon is exactly a
 * decoder's output picture. */
int rh264_write_p_slice_residual(const rh264_seq *s, int frame_num, int poc_lsb, int nrefs, const int32_t *rec,
                                 const uint8_t *src, const uint8_t *pred, uint8_t *recon, uint8_t *out, size_t cap,
                                 size_t *len);

#ifdef __cplusplus
}
#endif
#endif
