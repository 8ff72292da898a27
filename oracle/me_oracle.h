/*
 * me_oracle.h -- CPU restatement of rusty_h264's encoder motion-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * provider in rusty_h264_amd/ and the timed CPU baseline in bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker / the baseline -- never as the thing measured or
 * shipped.  The product library never links it.
 *
 * Every function restates a reference function (file:line, paths relative to
 * the reference repository root); arithmetic and evaluation order follow the
 * reference's scalar twins exactly.  Functions marked "build-defined" have no
 * reference counterpart: they are the plain exhaustive search and the int32
 * lookahead propagation, pinned by their definitions (brute force, closed
 * forms) in tests/.
 *
 * Plane-set layout shared with the GPU provider: one buffer of 4*pw*ph bytes,
 * planes in the order F (edge-replicated full-pel), H (b), V (h), C (j), each
 * pw*ph bytes with row stride pw, pw = cw + 2*pad, ph = ch + 2*pad
 * (crates/rusty_h264-common/src/inter.rs:707-722).
 */
#ifndef RH264_ME_ORACLE_H
#define RH264_ME_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- block metrics: crates/rusty_h264-accel/src/satd_sad.rs ---------------- */
int32_t orc_sad(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h);
int32_t orc_satd4x4_wels(const uint8_t *a, int sa, const uint8_t *b, int sb);
int32_t orc_satd_wels(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h);
void orc_hadamard_4x4(const int32_t in[16], int32_t out[16]);
int64_t orc_satd_px(const uint8_t *src, int ss, const uint8_t *pred, int ps, int w, int h);
uint32_t orc_satd_avg(const uint8_t *src, int ss, const uint8_t *a, const uint8_t *b,
                      int stride, int w, int h);
int orc_x4_shape(int w, int h);

/* ---- luma interpolation: crates/rusty_h264-common/src/inter.rs ------------- */
int orc_luma_sample(const uint8_t *ref, int cw, int ch, int ix, int iy, int fx, int fy);
void orc_mc_luma(const uint8_t *ref, int cw, int ch, int x0, int y0, int bw, int bh,
                 int mvx, int mvy, uint8_t *out);
void orc_pad_plane(const uint8_t *src, int w, int h, int pad, uint8_t *out);
void orc_build_hpel_tiles(const uint8_t *f, int pw, int ph, uint8_t *h, uint8_t *v, uint8_t *c);
void orc_build_hpel_fused(const uint8_t *f, int pw, int ph, uint8_t *h, uint8_t *v, uint8_t *c);
/* planes: 4*pw*ph bytes (F,H,V,C); fused=0 uses the tile walk, 1 the fused pass */
void orc_build_hpel_planes(const uint8_t *y, int cw, int ch, int pad, int fused, uint8_t *planes);
/* hpel_block: 1 = accepted (out filled), 0 = declined (inter.rs:1107-1162) */
int orc_hpel_block(const uint8_t *planes, int cw, int ch, int pad, int x0, int y0, int bw, int bh,
                   int mvx, int mvy, uint8_t *out);
/* prediction read from the plane set with every coordinate clamped into the
 * padded planes; equals orc_mc_luma for any MV when pad >= 3 (build-defined
 * read rule the GPU uses; pinned against orc_mc_luma in tests). */
void orc_pred_block(const uint8_t *planes, int cw, int ch, int pad, int x0, int y0, int bw,
                    int bh, int mvx, int mvy, uint8_t *out);

/* ---- motion search ---------------------------------------------------------- */
/* Exp-Golomb se(d) length, mb16.rs:2474-2475 */
uint32_t orc_mvbits(int32_t d);
/* lambda-rate table: lr[b] = (int64)(lambda * b), b = 0..n-1 (mb16.rs:2563) */
void orc_lambda_table(double lambda, int n, int64_t *lr);

#define ORC_LR_N 128

typedef struct {
    int cw, ch, pad;   /* frame size (MB multiples) and plane pad              */
    int bw, bh;        /* partition shape                                      */
    int range;         /* full-pel search radius R (full search)               */
    int subpel_pat;    /* 0: ring8 iterate, 1: ring4 iterate, 2: ring8 single,
                          3: ring4 single, -1: no sub-pel (mb16.rs:2956-3291);
                          4: quarter-pel window (build-defined, below)          */
    const int64_t *lr_fp;  /* SAD-phase lambda table, ORC_LR_N entries        */
    const int64_t *lr;     /* SATD-phase lambda table, ORC_LR_N entries       */
    int hadamard;          /* sub-pel SATD transform: 4 (or 0) 4x4, 8 = 8x8     */
} orc_search_cfg;

/* Plain exhaustive integer search of one block (build-defined; raster
 * first-min as mb16.rs:2843-2855).  centre = (cx,cy) qpel multiple of 4,
 * mvp = rate predictor (qpel).  Returns best cost, writes mv (qpel). */
int64_t orc_full_search_block(const orc_search_cfg *cfg, const uint8_t *cur,
                              const uint8_t *planes, int lx, int ly, int cx, int cy,
                              int mvpx, int mvpy, int32_t *mv_out);
/* SATD-domain cost of one candidate (mb16.rs:2540-2565 with the sub-pel ladder
 * values of mb16.rs:1930-2007). */
int64_t orc_satd_cost(const orc_search_cfg *cfg, const uint8_t *cur, const uint8_t *planes,
                      int lx, int ly, int mvx, int mvy, int mvpx, int mvpy);
/* Sub-pel ring refinement (mb16.rs:2956-3291, cascade path).  Pattern 4 is the
 * build-defined quarter-pel window that refines the learned proposal: the 25
 * positions start + (dx, dy), dx, dy in -2..2 quarter pels, priced like the ring's
 * candidates; the cheapest wins, ties going to the start, then to the earlier
 * position in raster order (dy, then dx).  evals = 25. */
int64_t orc_subpel_refine_block(const orc_search_cfg *cfg, const uint8_t *cur,
                                const uint8_t *planes, int lx, int ly, int mvx, int mvy,
                                int mvpx, int mvpy, int32_t *mv_out, int32_t *evals);

/* Whole-frame plain search: every bw x bh block of the frame, independent.
 * centres/mvps: per-block int32 pairs (qpel) or NULL for zero.
 * outputs per block: int_mv[2], int_cost, mv[2], cost. nthreads>1 uses OpenMP. */
void orc_plain_search_frame(const orc_search_cfg *cfg, const uint8_t *cur, const uint8_t *planes,
                            const int32_t *centres, const int32_t *mvps, int32_t *int_mv,
                            int64_t *int_cost, int32_t *mv, int64_t *cost, int nthreads);

/* Reference-mode motion_search (mb16.rs:2433-3302) with every knob explicit; the
 * steps, their order and the reference's pricing of each phase are listed above
 * orc_motion_search_ex in me_oracle.c.  cfg supplies the frame, the block shape
 * (bw x bh, any of the seven) and subpel_pat; the lambdas come from here (the
 * table-free form (int64)(lambda * rate) of mb16.rs:2563). */
typedef struct {
    const int32_t *steps; int nsteps;   /* diamond rungs in qpel units, e.g. {16,8,4} */
    int sadfp;                           /* 1: SAD full-pel phase (B2) priced with lambda_fp */
    int snap;                            /* integer snap of the seed (me_snap)         */
    int fc;                              /* fixed-centre argmin-of-4 passes (me_fc)    */
    int sad_only;                        /* the Fast preset's SAD cost closure          */
    const uint16_t *mv_tab;              /* NULL: Exp-Golomb; else bits per |d| (4096)  */
    double lambda, lambda_fp;            /* lambda_me and the B2 lambda                 */
    int rescue;                          /* me_wide rescue enabled for this search      */
    int rescue_range, rescue_thresh;     /* me_range (24), me_rescue (3)                */
    int rescue_fast, rescue_var;         /* me_fast (1), me_wide_var (800)              */
    int rescue_learn, rescue_payoff;     /* the frame's learning window: me_learn (40),
                                            me_payoff_pct (15); orc_motion_search_frame */
} orc_diamond_cfg;
/* rate of one mvd component: Exp-Golomb length, or tab[min(|d|, 4095)] */
uint32_t orc_mv_rate(const uint16_t *tab, int32_t d);
int64_t orc_motion_search_ref(const orc_search_cfg *cfg, const orc_diamond_cfg *dc,
                              const uint8_t *cur, const uint8_t *planes, int lx, int ly,
                              const int32_t *preds, int npreds, int32_t *mv_out);
/* the same search with its rescue events (events[0] = the grid ran, events[1] = it
 * cut the cost to <= 15/16; may be NULL) and an evaluation trace: up to trace_cap
 * entries {phase, mvx, mvy, cost} (int64), *trace_n = the number priced. */
int64_t orc_motion_search_ex(const orc_search_cfg *cfg, const orc_diamond_cfg *dc, const uint8_t *cur,
                             const uint8_t *planes, int lx, int ly, const int32_t *preds, int npreds,
                             int32_t *mv_out, int32_t *events, int64_t *trace, int trace_cap, int *trace_n);

/* predict_mv (inter.rs:36-73): nb = neighbours A (left), B (above), C (above-right
 * or above-left) as {available, mvx, mvy, ref_idx}; out = the predictor. */
void orc_predict_mv(const int32_t nb[12], int cur_ref, int32_t out[2]);
/* A P frame decided macroblock by macroblock in raster order, every macroblock
 * taking its 16x16 best_part result (mb16.rs:4921-4959, no extra seeds): for each
 * reference r (planes[r], r = 0 the most recent) orc_motion_search_ref with the
 * single predictor predict_mv(.., r) of the decided neighbours (mv_neighbors,
 * mb16.rs:1813-1836), cost + (int64)(lambda_ref * ref_bits(r, nrefs)), the first
 * minimum in reference order.  ref, cost: one per macroblock; mv: two.  With
 * dc->rescue, the frame's me_wide learning window (mb16.rs:2868-2886): the first
 * rescue_learn searches (in macroblock, then reference order) whose rescue grid ran
 * are tallied, those that cut their cost to <= 15/16 counted; when the window fills
 * with fewer than rescue_payoff % of them, the rescue is off for every later search
 * of the frame. */
void orc_motion_search_frame(const orc_search_cfg *cfg, const orc_diamond_cfg *dc, const uint8_t *cur,
                             const uint8_t *const *planes, int nrefs, double lambda_ref, int32_t *ref,
                             int32_t *mv, int64_t *cost);

/* The P macroblock decision with partitions in raster order (mb16.rs:8849-9010 with
 * the shape picked in the SATD currency, the all-8x8 P_8x8 arm when p8x8, every
 * macroblock inter; steps in me_oracle.c): planes[r] = reference r's plane set; lme
 * = the search lambda that also prices ref_bits and the P_8x8 syntax; split_gate
 * from orc_split_gate.  rec: 14 int32 per macroblock, {mode, cost, (ref, mvx, mvy)
 * x 4 quadrants TL, TR, BL, BR}. */
void orc_partition_decision_frame(const orc_search_cfg *cfg, const orc_diamond_cfg *dc, const uint8_t *cur,
                                  const uint8_t *const *planes, int nrefs, double lme, int64_t split_gate, int p8x8,
                                  int32_t *rec);
int64_t orc_split_gate(int qp);

/* skip_mv (mb16.rs:1841-1853): (0,0) when A or B is unavailable or a zero vector of
 * reference 0, else predict_mv(A, B, C, 0); nb as orc_predict_mv's. */
void orc_skip_mv(const int32_t nb[12], int32_t out[2]);
/* The luma predictions of a decided P frame: per macroblock the P_Skip vector of the
 * decided neighbours, mc_luma (inter.rs:1193-1262) of the frame refs[ref] at mv into
 * pred and of refs[0] at the skip vector into skip_pred (cw x ch each). refs: the
 * unpadded reference frames, most recent first. */
void orc_decision_predict(int cw, int ch, int nrefs, const uint8_t *const *refs, const int32_t *ref,
                          const int32_t *mv, int32_t *skip_mv, uint8_t *pred, uint8_t *skip_pred);

/* ---- mb-tree lookahead: crates/rusty_h264-encoder/src/mbtree.rs ------------- */
void orc_downsample2x(const uint8_t *y, int cw, int ch, uint8_t *out);
int32_t orc_intra_cost(const uint8_t *sy, int cw, int bx0, int by0, int bs);
int64_t orc_mbtree_mc_satd(const uint8_t *sy, int cw, int ch, const uint8_t *ref, int bx0,
                           int by0, int bs, int mvx, int mvy);
int32_t orc_inter_cost(const uint8_t *sy, int cw, int ch, const uint8_t *ref, int bx0, int by0,
                       int bs, int seedx, int seedy, int max_step, int32_t *mv_out);
/* the same with an evaluation trace: up to cap entries {mvx, mvy, satd} (int64);
 * *n = the number of evaluations */
int32_t orc_inter_cost_ex(const uint8_t *sy, int cw, int ch, const uint8_t *ref, int bx0, int by0, int bs,
                          int seedx, int seedy, int max_step, int32_t *mv_out, int64_t *trace, int cap, int *n);
/* HalfRes frame costs: half/ref_half are (mb_w*8)x(mb_h*8); ref_half NULL = IDR.
 * out: intra[mb], inter[mb], mv[2*mb] (full-res qpel). mbtree.rs:208-263 */
void orc_frame_costs_half(const uint8_t *half, const uint8_t *ref_half, int mb_w, int mb_h,
                          int32_t *intra, int32_t *inter, int32_t *mv);
/* f64 backward propagation over n frames (mbtree.rs:268-297, 440-455).
 * costs laid out [n][mb]; prop_out [n][mb] f64. */
void orc_propagate_f64(int n, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                       const int32_t *mv, double *prop_out);
/* gop_qp_offsets' policy (mbtree.rs:369-593) over frame costs [n][mb]: the f64
 * propagation above, the back-off latch (res_min), -strength*log2(total/intra)
 * centred on the GOP mean, the spread latch (sd_min), round half away from zero,
 * clamp +-6; with grain set (the grain latch, mbtree.rs:401-420) and n >= 2 every offset is
 * 0.  stats (may be NULL): residual fraction, spread, effective strength. */
void orc_mbtree_qp_offsets(int n, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                           const int32_t *mv, double strength, double res_min, double sd_min, int grain, int32_t *off,
                           double *stats);
/* grain_signature (signals.rs:528-533) of coded-size frame cur against ref (the GOP's
 * frames 1 and 0, mbtree.rs:404-406): 1 = grain; stats = {median variance, residual
 * floor, motion gain} (signals.rs:247-266, 311-338, 61-129). */
int orc_grain_signature(const uint8_t *cur, const uint8_t *ref, int cw, int ch, double *stats);
/* int32 fixed-point propagation (build-defined): amount_q = floor((intra+prop)
 * * (intra-inter) / intra) in int64 then per-target weights (wx*wy) >> 8. */
void orc_propagate_i32(int n, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                       const int32_t *mv, int32_t *prop_out);

/* Build-defined 8x8 Hadamard SATD (no reference counterpart): sum over the 8x8
 * blocks of a w x h region (w, h multiples of 8) of sum |H8 d H8^T| / 2, H8 the
 * order-8 Sylvester matrix (every coefficient has the parity of sum d, so the sum
 * is even and the halving exact). */
void orc_hadamard_8x8(const int32_t in[64], int32_t out[64]);
int64_t orc_satd8_px(const uint8_t *src, int ss, const uint8_t *pred, int ps, int w, int h);

/* lookahead::complexity (lookahead.rs:38-101): rate-control score of one coded-size
 * frame, IDR when ref is NULL; max(total, 1). */
double orc_rc_complexity(const uint8_t *sy, const uint8_t *ref, int cw, int ch);

/* Build-defined multi-reference lookahead (the 4K configuration's 3 refs +-64):
 * half-res frames mb_w*8 x mb_h*8, 8x8 blocks.  intra = max(1, AC sum|H8 x H8| >> 1);
 * frame 0: inter = intra, mv 0, ref 0; frame f > 0: per reference k = 1..min(nrefs,
 * f) the exhaustive full-pel SAD search over +-range (clamped reads, raster first
 * minimum), the 8x8 SATD at its winner, the cheapest k (ties to the smaller);
 * inter = min(that, intra), mv = 8 * winner (full-res quarter pel), ref = k. */
int64_t orc_la_block_multi(const uint8_t *half, int cw, int ch, int f, int nrefs, int range, int bx0, int by0,
                           int32_t *mv_out, int32_t *ref_out);
/* Learned-seeded form: per reference k the 8 centres cen[(k-1)*16 + 2c .. +1] (half-res
 * full pels) are searched over +-local_range (first minimum in candidate order, then
 * raster order); a best SAD above fb_sad falls back to the exhaustive +-range search
 * (bit k-1 of *fb_mask); then the 8x8 SATD and reference choice as above. */
int64_t orc_la_block_learned(const uint8_t *half, int cw, int ch, int f, int nrefs, int range, int local_range,
                             int fb_sad, const int32_t *cen, int bx0, int by0, int32_t *mv_out, int32_t *ref_out,
                             int32_t *fb_mask);
void orc_mbtree_costs_multi(const uint8_t *half, int nframes, int mb_w, int mb_h, int nrefs, int range,
                            int32_t *intra, int32_t *inter, int32_t *mv, int32_t *ref, int nthreads);
/* int32 propagation with frame f's amount going to frame f - ref[f][mb]
 * (ref NULL: every block references f - 1, which is orc_propagate_i32). */
void orc_propagate_multi_i32(int n, int mb_w, int mb_h, const int32_t *intra, const int32_t *inter,
                             const int32_t *mv, const int32_t *ref, int32_t *prop);

#ifdef __cplusplus
}
#endif
#endif
