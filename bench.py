"""Motion-search throughput on one MI355X: 1080p plain vs learned search, Mpx/s.

Workload (one step): a 1920x1088 clip of 16 frames (15 P-frames, 1 reference =
the previous frame) already resident in HBM; build the F/H/V/C plane sets of the
15 references (one launch), then either
  plain:   exhaustive +-32 integer SAD search of every 16x16 block, cascade ring
           refinement to quarter pel with 4x4-Hadamard SATD;
  learned: the proposal network's 8 candidates per block re-costed exactly (+-2 pel
           each), exhaustive +-32 fall-back above the threshold, the +-2 quarter-pel
           window refinement;
lambda of QP 26.  Mpx/s = 1920*1088*15 / step time; the headline is the faster path
within 1% SATD-cost regret.

Prints one JSON line: metric/value/unit, timing, `roofline` (the headline path's
dominant kernel: its algorithmic work over its HIP-event-timed duration on the
library's stream, against the published peak), the paths with their own rooflines,
and `cpu_baseline` (the C restatement on the host cores, 1 thread and all of the
box's share).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, H, NFRAMES, R, PAD, QP, SEED = 1920, 1088, 16, 32, 32, 26, 2026
HBM_PEAK_GBS = 8000.0
# byte abs-diff-adds: v_sad_u8 does 4 per lane per cycle at full rate, v_qsad_pk_u16_u8 16 at
# quarter rate: 64 per SIMD per cycle x 1024 SIMDs x 2.4 GHz
VALU_PEAK_TOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0  # dense matrix-core peaks
FP16_PEAK_TFLOPS = 2500.0
C_NET, K_NET = 32, 289     # the proposal network's channels and MV bins
# HBM bytes per launch of this workload's kernels, from rocprofv3 --pmc FETCH_SIZE (KiB,
# x2: a plain 512 MiB copy kernel reads 262,160 KiB) and WRITE_SIZE (KiB, x1), as
# recorded in profiles/motion_search_1080p_pmc.md
TRAFFIC_SOURCE = "profiles/motion_search_1080p_pmc.md"
KERNEL_TRAFFIC = {
    "k_full_search<16,16>": (33054 * 2 + 1912) * 1024,
    "k_prep": (30624 * 2 + 7650) * 1024,
    "k_fwd_rr": (8864 * 2 + 3825) * 1024,
    "k_cand_search_m<2,true>": (34148 * 2 + 1531) * 1024,
    "k_full_search_list<16,16,1>": (12630 * 2 + 5090) * 1024,
    "k_refine_wn16<false>": (80170 * 2 + 1434) * 1024,
    "k_refine_mx16<0,false>": (80386 * 2 + 2152) * 1024,
    "k_build_planes_rw": (15438 * 2 + 134003) * 1024,
}
HBM_PEAK_TBS = HBM_PEAK_GBS / 1000.0
# learned-seeded 4K lookahead: +-2 half-res pels around each proposal; fall-back SAD
LA_LOCAL_RANGE = 2
LA_FALLBACK_SAD = 1024


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def box_threads():
    """The host threads this run may use: the box's CPU share (OMP_NUM_THREADS, 16
    on the GPU box), never more than the cores present."""
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(os.cpu_count() or 1, share))


def cpu_baseline(frames, lam, budget_s=6.0):
    """The oracle (reference-semantics C restatement, -O3 -march=x86-64-v3, psadbw SAD,
    OpenMP over blocks) on the host cores: the same work per P-frame as the GPU plain
    path, cycling over the clip's P-frames for about budget_s per arm, once on one
    thread (how the reference publishes its numbers) and once on the box's share."""
    from oracle.oracle import Oracle
    o = Oracle(fast=True)
    cfg = o.cfg(W, H, PAD, 16, 16, R, 0, lam)
    arms = {}
    for name, nt in (("1_thread", 1), ("all_cores", box_threads())):
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < budget_s or done == 0:
            f = done % (NFRAMES - 1)
            pl = o.planes(frames[f], W, H, PAD, fused=True)
            o.plain_search_frame(cfg, frames[f + 1], pl, nthreads=nt)
            done += 1
        dt = time.perf_counter() - t0
        arms[name] = {"value": round(W * H * done / dt / 1e6, 3), "threads": nt, "frames": done,
                      "seconds": round(dt, 2)}
    allc = arms["all_cores"]
    return {"value": allc["value"], "unit": "Mpx/s", "cores": allc["threads"], "kind": "port",
            "sample": f"{allc['frames']} P-frame searches of the 1080p clip (cycling over its 15 P-frames): plane "
                      f"build + +-32 full search + qpel ring refine, {allc['threads']} OpenMP threads, "
                      f"{allc['seconds']} s; 1 thread: {arms['1_thread']['frames']} frames in "
                      f"{arms['1_thread']['seconds']} s",
            "arms": arms, "cpu_model": _cpu_model(), "host_cores": os.cpu_count(),
            "build": "oracle/me_oracle.c -O3 -march=x86-64-v3 (the reference's target-cpu), psadbw SAD"}


def cpu_reference_search(frames, lam, budget_s=6.0):
    """The reference's own search on the CPU, beside the GPU decision path: every
    macroblock's 16x16 best_part in raster order (predict_mv predictor, integer snap,
    diamond 16/8/4 with the fixed-centre passes, ring-8 sub-pel walk; the oracle's
    orc_motion_search_frame), one frame per thread, frames cycling over the clip."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.oracle import Oracle
    o = Oracle(fast=True)
    cfg = o.cfg(W, H, PAD, 16, 16, R, 0, lam)
    planes, results = {}, {}

    def one(f):
        if f not in planes:
            planes[f] = o.planes(frames[f], W, H, PAD, fused=True)
        _, mv, cost = o.motion_search_frame(cfg, frames[f + 1], planes[f], fc=True)
        results.setdefault(f, (mv, cost))

    arms = {}
    for name, nt in (("1_thread", 1), ("all_cores", box_threads())):
        t0 = time.perf_counter()
        done = 0
        with ThreadPoolExecutor(nt) as ex:
            while time.perf_counter() - t0 < budget_s or done == 0:
                list(ex.map(one, [(done + k) % (NFRAMES - 1) for k in range(nt)]))
                done += nt
        dt = time.perf_counter() - t0
        arms[name] = {"mpx_s": round(W * H * done / dt / 1e6, 3), "threads": nt, "frames": done,
                      "seconds": round(dt, 2)}
    return arms, results


MAX_REGRET = 0.01  # the learned path counts only at <= 1% mean SATD-cost regret


def run_learned(dev, args, cur, refs, planes, p, plain_cost, pat=None):
    """The learned-proposal path on the same clip: plane build + rh264_learned_search
    (network, 8 local searches, selection, fall-back full search, then the +-2
    quarter-pel window around the best integer vector: the 25 positions, exact SATD).
    Regret = sum of final costs / sum of the plain search's final costs - 1."""
    from rusty_h264_amd import accel, mvnet
    if not os.path.exists(mvnet.WEIGHTS):
        return None
    nf = NFRAMES - 1
    nb = (W // 16) * (H // 16) * nf
    pbytes = accel.planes_bytes(W, H, PAD)
    fb_cost = mvnet.FALLBACK_COST if args.fallback_cost is None else args.fallback_cost
    net = mvnet.MVNet(dev, nf, W, H, mvnet.load_weights())
    # the proposal's refinement: the +-2 quarter-pel window around the best integer
    # vector (or, for comparison, the plain search's ring walk, pattern 0)
    pat = mvnet.REFINE_PAT if pat is None else pat
    pl = accel.search_params(W, H, PAD, 16, 16, R, pat, accel.lambda_me(QP))
    outs = (dev.alloc(8 * nb), dev.alloc(4 * nb), dev.alloc(4 * nb))
    lib = dev.lib

    def step():
        accel.build_planes(dev, refs, nf, W, H, PAD, out=planes)
        mvnet.learned_search(dev, net, pl, nf, cur, refs, W * H, planes, pbytes, mvnet.LOCAL_RANGE, fb_cost, out=outs)

    for _ in range(args.warmup):
        step()
    lib.rh264_set_profiling(dev.ctx, args.steps)
    dev.sync()
    dev.record(0)
    for _ in range(args.steps):
        step()
    dev.record(1)
    dev.sync()
    ms_step = dev.elapsed_ms(0, 1) / args.steps
    import ctypes as C
    tt = (C.c_float * 8)()
    nev = C.c_int()
    times = []
    for i in range(lib.rh264_profile_count(dev.ctx)):
        assert lib.rh264_profile_times(dev.ctx, i, tt, 8, C.byref(nev)) == 0, lib.rh264_last_error()
        times.append([tt[k] for k in range(nev.value)])
    lib.rh264_set_profiling(dev.ctx, 0)
    cost = dev.download(outs[1], (nb,), np.int32).astype(np.int64)
    fb = dev.download(outs[2], (nb,), np.int32)
    net.close()
    pc = plain_cost.astype(np.int64)
    # phase end times after the search's start: prep, network, candidates, fall-back, refine
    t = np.mean(np.array(times), axis=0)
    ph_mean = [t[0]] + [t[k] - t[k - 1] for k in range(1, 5)]
    search_ms = float(t[-1])
    # the learned step's phases against their rooflines (algorithmic work per step
    # over the phase's event-timed duration on the library's stream)
    nblk1 = (W // 16) * (H // 16)
    w4, h4 = W // 4, H // 4
    net_flop = nf * 2 * ((w4 // 2) * (h4 // 2) * C_NET * (9 * 2 + 3 * 9 * C_NET) + nblk1 * K_NET * C_NET)
    n_fb = float(fb.sum())
    cand_ops = nb * 8 * (2 * mvnet.LOCAL_RANGE + 1) ** 2 * 256
    fb_ops = n_fb * (2 * R + 1) ** 2 * 256
    # the window's unique reads per block: the source, and each of F/H/V/C over the
    # 17x17 that the 16 operands at integer offsets {-1, 0}^2 cover
    win_bytes = nb * (256 + 4 * 17 * 17)
    fbv = int(dict(kv.split("=", 1) for kv in args.tune).get("fallback_v", 1))
    fb_kernel = f"k_full_search_list<16,16,{ {0: 0, 1: 1, 2: 3}[fbv] }>"
    prep_bytes = nf * (2 * W * H + 2 * w4 * h4 * 2)  # both frames read, the 1/4-res BF16 pair written
    roof = []
    for name, kernel, work, ms, peak, unit, bound, wunit in (
            ("input preparation", "k_prep", prep_bytes, ph_mean[0], HBM_PEAK_TBS, "TB/s", "hbm", "bytes"),
            ("network", "k_fwd_rr", net_flop, ph_mean[1], BF16_PEAK_TFLOPS, "TFLOP/s", "mfma", "FLOP"),
            ("candidates", "k_cand_search_m<2,true>", cand_ops, ph_mean[2], VALU_PEAK_TOPS, "TOP/s", "valu",
             "abs-diff-adds"),
            ("fall-back", fb_kernel, fb_ops, ph_mean[3], VALU_PEAK_TOPS, "TOP/s", "valu", "abs-diff-adds"),
            ("refine", "k_refine_wn16<false>" if pat == 4 else f"k_refine_mx16<{pat},false>", win_bytes, ph_mean[4],
             HBM_PEAK_TBS, "TB/s", "hbm", "bytes")):
        if ms > 0:
            a = work / (ms / 1e3) / 1e12
            roof.append({"phase": name, "kernel": kernel, "bound": bound, "work_per_step": float(work),
                         "work_unit": wunit, "ms": round(float(ms), 4), "achieved": round(a, 3), "peak": peak,
                         "unit": unit, "frac": round(a / peak, 4), "traffic": KERNEL_TRAFFIC.get(kernel)})
    return {"mpx_s": round(W * H * nf / (ms_step / 1e3) / 1e6, 2), "ms_per_step": round(ms_step, 4),
            "roofline": roof,
            "regret": round(float(cost.sum() / pc.sum() - 1.0), 6),
            "blocks_worse": round(float((cost > pc).mean()), 5), "blocks_better": round(float((cost < pc).mean()), 5),
            "fallback_rate": round(float(fb.mean()), 5), "local_range": mvnet.LOCAL_RANGE, "fallback_cost": fb_cost,
            "refine": "+-2 quarter-pel window (subpel_pat 4)" if pat == 4 else f"ring walk (subpel_pat {pat})",
            "phase_ms": {r["phase"]: r["ms"] for r in roof}, "search_ms": round(search_ms, 4),
            "plane_build_ms": round(ms_step - search_ms, 4)}


BD_QPS = (22, 27, 32, 37)


def bjontegaard_rate(r1, p1, r2, p2):
    """BD-rate (%) of the curve (r2, p2) against (r1, p1): cubic fits of log10(rate)
    against PSNR, their mean difference over the common PSNR interval (Bjontegaard,
    VCEG-M33)."""
    c1, c2 = np.polyfit(p1, np.log10(r1), 3), np.polyfit(p2, np.log10(r2), 3)
    lo, hi = max(min(p1), min(p2)), min(max(p1), max(p2))
    i1, i2 = np.polyint(c1), np.polyint(c2)
    avg = ((np.polyval(i2, hi) - np.polyval(i2, lo)) - (np.polyval(i1, hi) - np.polyval(i1, lo))) / (hi - lo)
    return float((10 ** avg - 1) * 100)


def run_bd_rate(dev, frames, cur, refs, planes):
    """BD-rate of the learned path's vectors against the plain path's: the clip coded
    as an I_PCM IDR + 15 P pictures (16x16, one reference, luma residual, CAVLC) at
    QP 22/27/32/37, each path searching with that QP's lambda on the source frames and
    the pictures predicted from the previous reconstruction (accel.encode_p_frames).
    Rate = the P pictures' bytes, distortion = their mean luma PSNR."""
    from rusty_h264_amd import accel, mvnet
    if not os.path.exists(mvnet.WEIGHTS):
        return None
    nf = NFRAMES - 1
    nb = (W // 16) * (H // 16) * nf
    pbytes = accel.planes_bytes(W, H, PAD)
    accel.build_planes(dev, refs, nf, W, H, PAD, out=planes)
    net = mvnet.MVNet(dev, nf, W, H, mvnet.load_weights())
    outs = (dev.alloc(8 * nb), dev.alloc(4 * nb), dev.alloc(4 * nb))
    curves = {"plain": ([], []), "learned": ([], [])}
    per_qp = []
    t0 = time.perf_counter()
    for qp in BD_QPS:
        lam = accel.lambda_me(qp)
        ms = accel.MotionSearch(dev, accel.search_params(W, H, PAD, 16, 16, R, 0, lam))
        ms.run_plain(cur, planes, nf)
        mvs = {"plain": ms.results(nf)["mv"]}
        pl = accel.search_params(W, H, PAD, 16, 16, R, mvnet.REFINE_PAT, lam)
        mvnet.learned_search(dev, net, pl, nf, cur, refs, W * H, planes, pbytes, mvnet.LOCAL_RANGE,
                             mvnet.FALLBACK_COST, out=outs)
        mvs["learned"] = dev.download(outs[0], (nb, 2), np.int32)
        row = {"qp": qp}
        for name, mv in mvs.items():
            _, sizes, recon = accel.encode_p_frames(dev, frames, mv, qp, pad=PAD)
            psnr = [10 * np.log10(255.0 ** 2 / max(np.mean((recon[f].astype(np.float64) - frames[f]) ** 2), 1e-9))
                    for f in range(1, NFRAMES)]
            bits = 8 * sum(sizes[1:])
            curves[name][0].append(bits)
            curves[name][1].append(float(np.mean(psnr)))
            row[name] = {"p_bytes": bits // 8, "psnr_y": round(float(np.mean(psnr)), 4)}
        per_qp.append(row)
    net.close()
    for x in outs:
        x.free()
    bd = bjontegaard_rate(curves["plain"][0], curves["plain"][1], curves["learned"][0], curves["learned"][1])
    return {"bd_rate_pct": round(bd, 4), "qps": per_qp, "seconds": round(time.perf_counter() - t0, 1),
            "what": "learned vs plain vectors, each path at the QP's lambda; I_PCM IDR + 15 P pictures, 16x16 "
                    "P_L0 / P_Skip, 1 ref, luma residual (forward core transform, inter dead zone, CAVLC), "
                    "predicted from the previous reconstruction; rate = P-picture bytes, PSNR = mean luma PSNR "
                    "of the P pictures; negative = the learned vectors need fewer bits"}


def run_decision(dev, args, cur, planes, lam, nf):
    """The encoder-order 16x16 decision of the 15 P-frames (rh264_motion_search_wavefront:
    one reference, each macroblock's predictor predict_mv of its decided neighbours,
    diamond rungs 16/8/4 with the fixed-centre passes and the iterated ring-8 sub-pel
    walk), timed with HIP events."""
    from rusty_h264_amd import accel
    p = accel.search_params(W, H, PAD, 16, 16, R, 0, lam)
    out = accel.motion_search_wavefront(dev, p, nf, cur, planes, fc=True)
    for _ in range(max(1, args.warmup)):
        accel.motion_search_wavefront(dev, p, nf, cur, planes, out=out, fc=True)
    steps = max(1, min(args.steps, 10))
    dev.sync()
    dev.record(0)
    for _ in range(steps):
        accel.motion_search_wavefront(dev, p, nf, cur, planes, out=out, fc=True)
    dev.record(1)
    dev.sync()
    ms = dev.elapsed_ms(0, 1) / steps
    cost = dev.download(out[2], (nf * (W // 16) * (H // 16),), np.int32)
    gmv = dev.download(out[1], (nf, (W // 16) * (H // 16), 2), np.int32)
    # the step after: P_Skip vectors and both luma predictions of every macroblock
    nblk = nf * (W // 16) * (H // 16)
    smv, pred, spred = dev.alloc(8 * nblk), dev.alloc(W * H * nf), dev.alloc(W * H * nf)
    pstride = accel.planes_bytes(W, H, PAD)

    def predict():
        rc = dev.lib.rh264_decision_predict(dev.ctx, W, H, PAD, nf, 1, int(planes), pstride, out[0].ptr, out[1].ptr,
                                            smv.ptr, pred.ptr, spred.ptr)
        assert rc == 0, dev.lib.rh264_last_error()
    predict()
    dev.sync()
    dev.record(0)
    for _ in range(steps):
        predict()
    dev.record(1)
    dev.sync()
    ms_pred = dev.elapsed_ms(0, 1) / steps
    for x in tuple(out) + (smv, pred, spred):
        x.free()
    # the decision with partitions (16x16, 16x8, 8x16, P_8x8 in the SATD currency)
    gate = accel.split_gate(QP)
    rec = accel.partition_decision(dev, p, nf, cur, planes, lambda_ref=lam, gate=gate, fc=True)
    dev.sync()
    dev.record(0)
    for _ in range(steps):
        accel.partition_decision(dev, p, nf, cur, planes, lambda_ref=lam, gate=gate, fc=True, out=rec)
    dev.record(1)
    dev.sync()
    ms_part = dev.elapsed_ms(0, 1) / steps
    modes = np.bincount(dev.download(rec, (nf * (W // 16) * (H // 16), 14), np.int32)[:, 0], minlength=4)
    rec.free()
    part = {"mpx_s": round(W * H * nf / (ms_part / 1e3) / 1e6, 2), "ms_per_step": round(ms_part, 4),
            "split_gate": gate, "modes": {k: int(v) for k, v in zip(("16x16", "16x8", "8x16", "8x8"), modes)},
            "what": "encoder-order decision with partitions: 16x16 best_part, then above the split gate 16x8, "
                    "8x16 and P_8x8 best_parts seeded with the 16x16 vector, shape by SATD + rate"}
    return {"mpx_s": round(W * H * nf / (ms / 1e3) / 1e6, 2), "ms_per_step": round(ms, 4), "steps": steps,
            "mean_cost": float(cost.mean()), "predict_ms": round(ms_pred, 4), "partitions": part,
            "_gpu": (gmv, cost.reshape(nf, -1)),
            "what": "encoder-order decision: 16x16, 1 ref, predict_mv predictors, diamond 16/8/4 with the "
                    "fixed-centre passes (the reference's default), ring-8 walk; predict_ms: P_Skip vectors + MC "
                    "and P_Skip luma predictions of the decided frames"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 1000 for 1080p, about a second of GPU work per path; "
                         "20 for batched; 100 for lookahead4k)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed warm-up steps (default 10, 3 for batched)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal path")
    ap.add_argument("--no-decision", action="store_true", help="skip the encoder-order decision path")
    ap.add_argument("--no-bdrate", action="store_true",
                    help="skip the BD-rate of the learned path's vectors against the plain path's")
    ap.add_argument("--fallback-cost", type=int, default=None, help="learned path: fall-back threshold")
    ap.add_argument("--la-fb", type=int, default=LA_FALLBACK_SAD,
                    help="lookahead4k: 8x8 SAD above which a learned-seeded block falls back to the full search")
    ap.add_argument("--tune", action="append", default=[], help="knob=value passed to rh264_tune")
    ap.add_argument("--workload", default="1080p", choices=["1080p", "batched", "lookahead4k"],
                    help="1080p (the headline metric) or a secondary workload")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"1080p": 1000, "batched": 20, "lookahead4k": 100}[args.workload]
    if args.warmup is None:
        args.warmup = 3 if args.workload == "batched" else 10
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the motion search on a single MI355X "
              "(streams and GOPs shard across GPUs with no exchange between them). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "lookahead4k":
        return bench_lookahead4k(args)
    from rusty_h264_amd import accel, synth
    frames = synth.hd_clip(NFRAMES, seed=SEED, w=W, h=H)
    lam = accel.lambda_me(QP)
    nf = NFRAMES - 1
    with accel.Device(0) as dev:
        dname = dev.name()
        clip = dev.alloc(frames.nbytes)
        # host-to-device upload of the clip, outside the timed region (reported apart:
        # the C ABI takes device pointers, so no transfer is part of a step)
        h2d = []
        for _ in range(3):
            t0 = time.perf_counter()
            accel._check(dev.lib.rh264_h2d(dev.ctx, clip.ptr, frames.ctypes.data, frames.nbytes), "rh264_h2d")
            h2d.append((time.perf_counter() - t0) * 1e3)
        h2d_ms = float(np.median(h2d))
        cur = int(clip) + W * H  # frames 1..15
        refs = int(clip)         # frames 0..14
        pbytes = accel.planes_bytes(W, H, PAD)
        planes = dev.alloc(pbytes * nf)
        p = accel.search_params(W, H, PAD, 16, 16, R, 0, lam)
        ms = accel.MotionSearch(dev, p)
        lib = dev.lib
        for kv in args.tune:
            k, v = kv.split("=")
            dev.tune(k, int(v))

        def step():
            accel.build_planes(dev, refs, nf, W, H, PAD, out=planes)
            ms.run_plain(cur, planes, nf)

        for _ in range(args.warmup):
            step()
        dev.sync()
        # timed region: K steps bracketed by device syncs; HIP events on the stream
        # the kernels run on time the whole region and, per step, the search kernels
        lib.rh264_set_profiling(dev.ctx, args.steps)
        dev.sync()
        t0 = time.perf_counter()
        dev.record(0)
        for _ in range(args.steps):
            step()
        dev.record(1)
        dev.sync()
        wall = time.perf_counter() - t0
        ev_ms = dev.elapsed_ms(0, 1)
        import ctypes as C
        fs, rf = C.c_float(), C.c_float()
        full_ms, refine_ms = [], []
        for i in range(lib.rh264_profile_count(dev.ctx)):
            lib.rh264_profile_get(dev.ctx, i, C.byref(fs), C.byref(rf))
            full_ms.append(fs.value)
            refine_ms.append(rf.value)
        lib.rh264_set_profiling(dev.ctx, 0)
        res = ms.results(nf, with_evals=False)
        learned = learned_ring = None
        if not args.no_learned:
            learned = run_learned(dev, args, cur, refs, planes, p, res["cost"])
            # the same proposal refined by the plain search's ring walk, for comparison
            learned_ring = run_learned(dev, args, cur, refs, planes, p, res["cost"], pat=0)
        decision = None if args.no_decision else run_decision(dev, args, cur, planes, lam, nf)
        bd = None if (args.no_learned or args.no_bdrate) else run_bd_rate(dev, frames, cur, refs, planes)
        if learned is not None and bd is not None:
            learned["bd_rate"] = bd

    ms_per_step = ev_ms / args.steps
    mpx = W * H * nf / (ms_per_step / 1e3) / 1e6
    nblk = (W // 16) * (H // 16)
    ops = nf * nblk * (2 * R + 1) ** 2 * 256  # abs-diff-adds per full-search launch
    full_avg = float(np.mean(full_ms))
    achieved = ops / (full_avg / 1e3) / 1e12
    plain_roof = {"phase": "integer search", "kernel": "k_full_search<16,16>", "bound": "valu",
                  "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                  "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": KERNEL_TRAFFIC["k_full_search<16,16>"],
                  "traffic_unit": "bytes per launch", "traffic_source": TRAFFIC_SOURCE,
                  "algorithmic_bytes": nf * (W * H + (W + 2 * PAD) * (H + 2 * PAD)),
                  "work": f"{ops} abs-diff-adds per launch ((2R+1)^2*256 per 16x16 block)",
                  "kernel_ms": round(full_avg, 4), "refine_ms": round(float(np.mean(refine_ms)), 4)}
    out = {
        "metric": "1080p motion-search Mpx/s, GPU plain vs learned, at <=1% SATD-cost regret",
        "value": round(mpx, 2), "unit": "Mpx/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "1080p 1920x1088 clip, 16 frames (15 P-frames searched), 1 ref, 6-tap half-pel "
                               "planes, QP 26 lambda; plain: +-32 full-search SAD + 4x4-Hadamard SATD qpel ring "
                               "walk", "search": "plain", "block": "16x16", "range": R, "pad": PAD,
                   "device": dname},
        "roofline": dict(plain_roof),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "mean_cost": float(res["cost"].mean()),
        "h2d": {"ms": round(h2d_ms, 3), "bytes": int(frames.nbytes),
                "gb_s": round(frames.nbytes / (h2d_ms / 1e3) / 1e9, 2),
                "mpx_s_incl_h2d": None,
                "note": "upload of the 16-frame clip from pageable host memory, once per clip; never part of value"},
        "paths": {"plain": {"mpx_s": round(mpx, 2), "ms_per_step": round(ms_per_step, 4), "regret": 0.0,
                            "roofline": [plain_roof]},
                  "learned": learned, "learned_ring": learned_ring, "decision": decision},
    }
    if learned is not None and learned["regret"] <= MAX_REGRET and learned["mpx_s"] > mpx:
        # the headline is the faster path that meets the regret bar, and the top-level
        # roofline is its dominant kernel (the longest event-timed phase)
        out["value"], out["ms_per_step"] = learned["mpx_s"], learned["ms_per_step"]
        out["config"]["search"] = "learned"
        out["config"]["workload"] = ("1080p 1920x1088 clip, 16 frames (15 P-frames searched), 1 ref, 6-tap half-pel "
                                     "planes, QP 26 lambda; learned: BF16 proposal network (8 candidates per 16x16 "
                                     "block), each re-costed by an exact +-2 pel SAD search, +-32 full-search "
                                     "fall-back above the threshold, 4x4-Hadamard SATD +-2 quarter-pel window")
        dom = max(learned["roofline"], key=lambda r: r["ms"])
        out["roofline"] = dict(dom, traffic=KERNEL_TRAFFIC.get(dom["kernel"]), traffic_unit="bytes per launch",
                               traffic_source=TRAFFIC_SOURCE if dom["kernel"] in KERNEL_TRAFFIC else None,
                               note="the learned step's longest phase; every phase is in paths.learned.roofline, "
                                    "the plain kernel in paths.plain.roofline")
    out["h2d"]["mpx_s_incl_h2d"] = round(W * H * nf / ((out["ms_per_step"] + h2d_ms) / 1e3) / 1e6, 2)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(frames, lam)
        if decision is not None:
            arms, oracle_res = cpu_reference_search(frames, lam)
            gmv, gcost = decision["_gpu"]
            same = all(np.array_equal(gmv[f], m) and np.array_equal(gcost[f], c) for f, (m, c) in oracle_res.items())
            decision["cpu_reference_search"] = dict(arms, what="the same search on the host (the oracle), frames "
                                                               "spread over threads as the reference's GOP threads")
            decision["parity_vs_oracle"] = ("bit-exact" if same else "MISMATCH") + \
                f" (vectors and costs of P-frames {sorted(oracle_res)})"
    if decision is not None:
        decision.pop("_gpu", None)
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """64 concurrent streams (24 CIF, 24 1280x720, 16 1920x1088, distinct seeds), 1-4
    references each, partitions 16x16/16x8/8x16/8x8/4x4, +-32: per step, the newest
    reference plane set of every stream is built and one P-frame of every stream is
    searched (all references, all shapes) as one captured HIP graph."""
    from rusty_h264_amd import accel, synth
    sizes = [(352, 288)] * 24 + [(1280, 720)] * 24 + [(1920, 1088)] * 16
    shapes = [(16, 16), (16, 8), (8, 16), (8, 8), (4, 4)]
    lam = accel.lambda_me(QP)
    with accel.Device(0) as dev:
        streams, planes_jobs, px, ref_px = [], [], 0, 0
        for i, (w, h) in enumerate(sizes):
            nrefs = 1 + i % 4
            gen = synth.cif_clip if w == 352 else synth.hd_clip
            clip = gen(nrefs + 1, seed=1000 + i, w=w, h=h)
            frames = dev.upload(clip[::-1].copy())  # frame 0 = current, 1.. = references newest first
            pb = accel.planes_bytes(w, h, PAD)
            planes = dev.alloc(nrefs * pb)
            accel.build_planes(dev, int(frames) + w * h, nrefs, w, h, PAD, out=planes)
            streams.append({"cur": frames, "planes": planes, "cw": w, "ch": h, "nrefs": nrefs})
            planes_jobs.append((int(frames) + w * h, int(planes), w, h))
            px += w * h
            ref_px += nrefs * w * h
        bt = accel.Batch(dev, streams, shapes, PAD, R, 0, lam)
        choices = bt.autotune(reps=1)

        def step():
            for src, dst, w, h in planes_jobs:  # newest reference of each stream
                accel.build_planes(dev, src, 1, w, h, PAD, out=dst)
            bt.replay()

        bt.capture()
        for _ in range(args.warmup):
            step()
        dev.sync()
        t0 = time.perf_counter()
        dev.record(0)
        for _ in range(args.steps):
            step()
        dev.record(1)
        dev.sync()
        wall = time.perf_counter() - t0
        ms = dev.elapsed_ms(0, 1) / args.steps
        # the integer searches alone, eager launches on the library's stream
        bt.search()
        dev.sync()
        dev.record(2)
        for _ in range(args.steps):
            bt.search()
        dev.record(3)
        dev.sync()
        ms_search = dev.elapsed_ms(2, 3) / args.steps
        bt.close()
    fused = all(v == 3 for v, _ in choices)
    out = {"metric": "batched multi-stream motion-search Mpx/s", "value": round(px / (ms / 1e3) / 1e6, 2),
           "unit": "Mpx/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "64 streams (24 CIF, 24 1280x720, 16 1920x1088), 1-4 refs, shapes "
                                  "16x16/16x8/8x16/8x8/4x4, +-32 full search + qpel refine + ref choice, "
                                  "graph replay", "pixels_per_step": px, "autotune": choices},
           "wall_ms_per_step": round(wall * 1e3 / args.steps, 4)}
    # The integer search's work: every pixel of every (stream, reference) against every
    # candidate, once per shape searched on its own.  The fused search (k_fs_fused)
    # takes each 4x4 block's SADs once and adds them into the larger partitions, so its
    # abs-diff-adds are one shape's worth; five_shape_work is what five separate
    # searches would do.
    five = ref_px * len(shapes) * (2 * R + 1) ** 2
    ops = ref_px * (2 * R + 1) ** 2 if fused else five
    ach = ops / (ms_search / 1e3) / 1e12
    out["phases"] = {"integer_search_ms": round(ms_search, 4), "refine_choice_planes_ms": round(ms - ms_search, 4)}
    out["roofline"] = {"bound": "valu",
                       "scope": ("k_fs_fused: the five shapes' integer search in one pass, abs-diff-adds done / its "
                                 "live duration" if fused else "the five integer searches' abs-diff-adds / their "
                                 "live duration"),
                       "work": float(ops), "five_shape_work": float(five), "achieved": round(ach, 2),
                       "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": round(ach / VALU_PEAK_TOPS, 4)}
    print(json.dumps(out), flush=True)


def bench_lookahead4k(args):
    """4K lookahead over a 40-frame window of 3840x2160 frames.  One step = the
    whole window (three forms of the cost stage are timed; the headline is the
    fastest multi-reference form within the 1% regret bar, normally the
    learned-seeded one):
      * downsample to half resolution (1920x1080; 8x8 blocks, one per 16x16 MB);
      * costs, in the configuration's form (build-defined): against each of the 3
        previous frames a +-64 full-pel search -- exhaustive, or seeded by the
        network's 8 proposals per block (+-2 half-res pels each, exhaustive fall-back
        above an 8x8 SAD of 1024) -- the 8x8 Hadamard cost of each winner, the
        cheapest reference; or in the reference's form (reported beside them): the
        seeded cascade diamond against the previous frame with 4x4 Hadamard costs
        (mbtree.rs:61-297);
      * int32 backward propagation (to each block's chosen reference);
      * the rate-control complexity of every full-resolution frame
        (lookahead.rs:38-101; frame 0 as the IDR, the others against their
        predecessors).
    Mpx/s = 3840*2160*40 / step time.  Parity: sampled blocks of frames 1..3 and
    whole frames 0..2 against the oracle, and each form's propagation from the
    GPU's own costs."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle.oracle import Oracle
    from rusty_h264_amd import accel, synth
    w, h, n = 3840, 2160, 40
    nrefs, rng_la = 3, 64
    mb_w, mb_h = w // 16, h // 16
    nm = mb_w * mb_h
    t_gen = time.perf_counter()
    frames = synth.hd_clip(n, seed=SEED + 5, w=w, h=h)
    t_gen = time.perf_counter() - t_gen
    with accel.Device(0) as dev:
        dname = dev.name()
        fb = dev.upload(frames)
        la_scaled = 1  # the library's default: distance-1 proposals scaled by the distance
        for kv in args.tune:
            k, v = kv.split("=")
            dev.tune(k, int(v))
            la_scaled = int(v) if k == "la_scaled" else la_scaled
        mt = accel.MbTree(dev, w, h, n)
        mref = dev.alloc(4 * nm * n)
        mfb = dev.alloc(4 * nm * n)
        cx = dev.alloc(8 * n)
        lib = dev.lib
        from rusty_h264_amd import mvnet
        net = mvnet.MVNet(dev, n - 1, w, h, mvnet.load_weights()) if os.path.exists(mvnet.WEIGHTS) else None

        def step(multi, ev=None):
            def mark(k):
                if ev is not None:
                    dev.record(ev + k)
            mark(0)
            accel._check(lib.rh264_mbtree_downsample(dev.ctx, int(fb), w * h, n, w, h, mt.half.ptr), "downsample")
            mark(1)
            if multi == 2:
                accel._check(lib.rh264_mbtree_costs_multi_learned(
                    dev.ctx, net.h, int(fb), w * h, mt.half.ptr, n, w, h, nrefs, rng_la, LA_LOCAL_RANGE, args.la_fb,
                    mt.intra.ptr, mt.inter.ptr, mt.mv.ptr, mref.ptr, mfb.ptr), "costs")
            elif multi:
                accel._check(lib.rh264_mbtree_costs_multi(dev.ctx, mt.half.ptr, n, mb_w, mb_h, nrefs, rng_la,
                                                          mt.intra.ptr, mt.inter.ptr, mt.mv.ptr, mref.ptr), "costs")
            else:
                accel._check(lib.rh264_mbtree_costs(dev.ctx, mt.half.ptr, n, mb_w, mb_h, mt.intra.ptr, mt.inter.ptr,
                                                    mt.mv.ptr), "costs")
            mark(2)
            if multi:
                accel._check(lib.rh264_mbtree_propagate_multi_i32(dev.ctx, n, mb_w, mb_h, mt.intra.ptr, mt.inter.ptr,
                                                                  mt.mv.ptr, mref.ptr, mt.prop.ptr), "propagate")
            else:
                accel._check(lib.rh264_mbtree_propagate_i32(dev.ctx, n, mb_w, mb_h, mt.intra.ptr, mt.inter.ptr,
                                                            mt.mv.ptr, mt.prop.ptr), "propagate")
            mark(3)
            accel._check(lib.rh264_rc_complexity(dev.ctx, int(fb), None, w * h, 1, w, h, cx.ptr), "complexity")
            accel._check(lib.rh264_rc_complexity(dev.ctx, int(fb) + w * h, int(fb), w * h, n - 1, w, h, cx.ptr + 8),
                         "complexity")
            mark(4)

        def timed(multi):
            for _ in range(args.warmup):
                step(multi)
            dev.sync()
            t0 = time.perf_counter()
            dev.record(0)
            for _ in range(args.steps):
                step(multi)
            dev.record(1)
            dev.sync()
            wall = time.perf_counter() - t0
            ms = dev.elapsed_ms(0, 1) / args.steps
            step(multi, ev=2)  # one more, with phase events
            dev.sync()
            names = ("downsample", "costs", "propagate", "rc_complexity")
            phases = {k: round(dev.elapsed_ms(2 + i, 3 + i), 4) for i, k in enumerate(names)}
            res = mt.results()
            if multi:
                res["ref"] = dev.download(mref, (n, nm), np.int32)
            if multi == 2:
                res["fb"] = dev.download(mfb, (n, nm), np.int32)
            res["cplx"] = dev.download(cx, (n,), np.float64)
            return ms, wall, phases, res

        ms_r, wall_r, ph_r, res_r = timed(0)
        ms_m, wall_m, ph_m, res_m = timed(1)
        res_l = None
        if net is not None:
            ms_l, wall_l, ph_l, res_l = timed(2)
            # the proposals of frames 1..3, for the parity spot check: the pairs (j + k, j),
            # or with la_scaled the pairs (j + 1, j) times the distance k, clamped to the
            # library's bound P - 8 - local range
            cen = np.zeros((4, nrefs, nm, 8, 2), np.int32)
            cmax = ((max(rng_la, 16 + LA_LOCAL_RANGE) + 8 + 3) & ~3) - 8 - LA_LOCAL_RANGE
            for k in range(1, (1 if la_scaled else nrefs) + 1):
                t8 = dev.alloc(4 * (4 - k) * nm * 8)
                net.forward(int(fb) + k * w * h, int(fb), w * h, 4 - k, w, h, None, t8)
                top = dev.download(t8, (4 - k, nm, 8), np.int32)
                t8.free()
                for kk in ((1, 2, 3)[:nrefs] if la_scaled else (k,)):
                    cen[kk:, kk - 1, :, :, 0] = np.clip((top[kk - k:] % 17 - 8) * 2 * (kk if la_scaled else 1),
                                                        -cmax, cmax)
                    cen[kk:, kk - 1, :, :, 1] = np.clip((top[kk - k:] // 17 - 8) * 2 * (kk if la_scaled else 1),
                                                        -cmax, cmax)
            net.close()
    o = Oracle(fast=True)
    halves = [o.downsample2x(frames[f], w, h) for f in range(4)]
    # reference form: whole frames 0..2
    ok_r = all(np.array_equal(res_r["half"][f], halves[f]) for f in range(3))
    for f in range(3):
        a, b, m = o.frame_costs_half(halves[f], halves[f - 1] if f else None, mb_w, mb_h)
        ok_r = ok_r and np.array_equal(res_r["intra"][f], a) and np.array_equal(res_r["inter"][f], b) and \
            np.array_equal(res_r["mv"][f], m)
    ok_r = ok_r and np.array_equal(res_r["prop"], o.propagate_i32(res_r["intra"], res_r["inter"], res_r["mv"],
                                                                  mb_w, mb_h))
    ok_c = all(res_m["cplx"][f] == o.rc_complexity(frames[f], frames[f - 1] if f else None, w, h) for f in range(3))
    # multi-reference form: 64 sampled blocks of frames 1..3 (the exhaustive oracle is
    # slow at 4K), frame 0 whole, and the propagation
    hv = np.stack(halves)
    smp = np.random.default_rng(7).integers(0, nm, (3, 64))
    ok_m = True
    i0, _, _, _ = o.mbtree_costs_multi(hv[:1], mb_w, mb_h, nrefs, rng_la)
    ok_m = ok_m and np.array_equal(res_m["intra"][0], i0[0]) and np.array_equal(res_m["inter"][0], i0[0])
    for f in (1, 2, 3):
        for mbi in smp[f - 1]:
            c, mvv, k = o.la_block_multi(hv, w // 2, h // 2, f, nrefs, rng_la, (mbi % mb_w) * 8, (mbi // mb_w) * 8)
            ic = int(res_m["intra"][f, mbi])
            ok_m = ok_m and int(res_m["inter"][f, mbi]) == min(c, ic) and tuple(res_m["mv"][f, mbi]) == mvv and \
                int(res_m["ref"][f, mbi]) == k
    ok_m = ok_m and np.array_equal(res_m["prop"], o.propagate_multi_i32(res_m["intra"], res_m["inter"], res_m["mv"],
                                                                        res_m["ref"], mb_w, mb_h))
    learned = None
    if res_l is not None:
        ok_l = np.array_equal(res_l["intra"], res_m["intra"])
        for f in (1, 2, 3):
            for mbi in smp[f - 1]:
                c, mvv, k, msk = o.la_block_learned(hv, w // 2, h // 2, f, nrefs, rng_la, LA_LOCAL_RANGE, args.la_fb,
                                                    cen[f, :, mbi], (mbi % mb_w) * 8, (mbi // mb_w) * 8)
                ic = int(res_l["intra"][f, mbi])
                ok_l = ok_l and int(res_l["inter"][f, mbi]) == min(c, ic) and tuple(res_l["mv"][f, mbi]) == mvv and \
                    int(res_l["ref"][f, mbi]) == k and int(res_l["fb"][f, mbi]) == msk
        ok_l = ok_l and np.array_equal(res_l["prop"], o.propagate_multi_i32(res_l["intra"], res_l["inter"],
                                                                            res_l["mv"], res_l["ref"], mb_w, mb_h))
        inter_ex = res_m["inter"][1:].astype(np.int64)
        inter_l = res_l["inter"][1:].astype(np.int64)
        pairs = sum(min(f, nrefs) for f in range(1, n)) * nm
        fbits = int(np.bitwise_count(res_l["fb"][1:].astype(np.uint32)).sum())
        learned = {"mpx_s": round(w * h * n / (ms_l / 1e3) / 1e6, 2), "ms_per_step": round(ms_l, 4), "phase_ms": ph_l,
                   "regret": round(float(inter_l.sum() / inter_ex.sum() - 1.0), 6),
                   "blocks_worse": round(float((inter_l > inter_ex).mean()), 5),
                   "fallback_rate": round(fbits / pairs, 5), "local_range": LA_LOCAL_RANGE, "fallback_sad": args.la_fb,
                   "search": "network proposals (8 per block; for reference k the distance-1 pair's "
                             "proposals times k when la_scaled, else the pair (f, f-k)'s) searched over +-2 "
                             "half-res pels, exhaustive +-64 fall-back above the SAD threshold",
                   "la_scaled": la_scaled,
                   "parity_vs_oracle": "bit-exact" if ok_l and ok_c else "MISMATCH"}
    mpx = lambda ms: round(w * h * n / (ms / 1e3) / 1e6, 2)
    out = {"metric": "4K lookahead Mpx/s (40-frame window: half-res mb-tree costs, int32 propagation, "
                     "RC complexity)",
           "value": mpx(ms_m), "unit": "Mpx/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_m, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "3840x2160, 40 frames; half-res (1920x1080) 8x8 blocks, 3 references, "
                                  "+-64 full-pel range per reference (config.search: exhaustive, or learned-seeded "
                                  "with exhaustive fall-back), 8x8 Hadamard costs, cheapest reference; int32 "
                                  "propagation to the chosen references; full-res RC complexity (9 full-pel "
                                  "candidates per MB)",
                      "search": "exhaustive",
                      "device": dname},
           "phase_ms": ph_m,
           "paths": {"multi_ref": {"mpx_s": mpx(ms_m), "ms_per_step": round(ms_m, 4), "phase_ms": ph_m,
                                   "parity_vs_oracle": "bit-exact" if ok_m and ok_c else "MISMATCH"},
                     "reference_form": {"mpx_s": mpx(ms_r), "ms_per_step": round(ms_r, 4), "phase_ms": ph_r,
                                        "search": "seeded 4-point cascade diamond (steps 8/4/2/1) against the "
                                                  "previous frame, 4x4 Hadamard costs (mbtree.rs:154-263)",
                                        "parity_vs_oracle": "bit-exact" if ok_r and ok_c else "MISMATCH"}},
           "parity_vs_oracle": "bit-exact" if ok_m and ok_r and ok_c else "MISMATCH",
           "wall_ms_per_step": round(wall_m * 1e3 / args.steps, 4), "generate_s": round(t_gen, 1)}
    if learned is not None:
        out["paths"]["multi_ref_learned"] = learned
        if learned["regret"] <= MAX_REGRET and learned["mpx_s"] > out["value"]:
            # the headline is the faster multi-reference form that meets the regret bar
            # (the learned proposals seeding the search, as the configuration names)
            out["value"], out["ms_per_step"], out["phase_ms"] = learned["mpx_s"], learned["ms_per_step"], ph_l
            out["wall_ms_per_step"] = round(wall_l * 1e3 / args.steps, 4)
            out["config"]["search"] = "learned-seeded"
            out["parity_vs_oracle"] = "bit-exact" if learned["parity_vs_oracle"] == "bit-exact" and \
                out["parity_vs_oracle"] == "bit-exact" else "MISMATCH"
    if not args.no_cpu:
        cores = min(os.cpu_count() or 1, 16)
        # reference form: the whole window on the host (downsample, frame costs, int32
        # propagation, RC complexity), frames spread over a thread pool
        t0 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            hvs = list(ex.map(lambda f: o.downsample2x(frames[f], w, h), range(n)))
            cs = list(ex.map(lambda f: o.frame_costs_half(hvs[f], hvs[f - 1] if f else None, mb_w, mb_h), range(n)))
            list(ex.map(lambda f: o.rc_complexity(frames[f], frames[f - 1] if f else None, w, h), range(n)))
        o.propagate_i32(np.stack([c[0] for c in cs]), np.stack([c[1] for c in cs]), np.stack([c[2] for c in cs]),
                        mb_w, mb_h)
        dt_r = time.perf_counter() - t0
        # multi-reference form: its exhaustive search dominates; time 2 block rows of
        # frame 3 (3 references each) and scale to the window's 39 x 32,400 blocks
        blocks = [(x * 8, y * 8) for y in (40, 41) for x in range(mb_w)]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(cores) as ex:
            list(ex.map(lambda b: o.la_block_multi(hv, w // 2, h // 2, 3, nrefs, rng_la, b[0], b[1]), blocks))
        dt_s = time.perf_counter() - t0
        dt_m = dt_s * (n - 1) * nm / len(blocks)
        out["cpu_baseline"] = {"value": round(w * h * n / dt_m / 1e6, 3), "unit": "Mpx/s", "cores": cores,
                               "kind": "port",
                               "sample": f"multi-reference search of {len(blocks)} blocks (2 block rows of frame 3, "
                                         f"3 references, +-64) in {dt_s:.1f} s on {cores} threads, scaled to the "
                                         f"window's {(n - 1) * nm} blocks; the reference form's whole window took "
                                         f"{dt_r:.1f} s ({w * h * n / dt_r / 1e6:.0f} Mpx/s)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
