"""Trains the learned MV-proposal network on the GPU with the project's own
forward/backward/Adam kernels (csrc/rh264_mvnet.hip).

Data: frame pairs (frame t, frame t-1) of synthetic clips from synth.hd_clip,
generated by a CPU worker pool while the GPU trains (each epoch regenerates the
same seeded clips, so host memory stays at a few clips).  Labels come from the
plain exhaustive search on the GPU (integer stage, +-32 pels, lambda of QP 26):
the label of a macroblock is the 4-pel bin of its full-search vector
(mvnet.mv_bins).  Loss: mean softmax cross-entropy over macroblocks; Adam with a
cosine learning-rate decay.  Batches are drawn from a shuffle buffer of a few
clips so one step mixes pairs of different clips.

The trained parameters are written as an .npz holding the float32 array
`params` (plus the training log and arguments as JSON strings).

    python -m rusty_h264_amd.train_mvnet --pairs 20000 --out mvnet.npz
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

CLIP_FRAMES = 16  # 15 pairs per clip


def _clip(job):
    seed, w, h = job
    from rusty_h264_amd import synth
    return seed, synth.hd_clip(CLIP_FRAMES, seed=seed, w=w, h=h)


def _stream(pool, jobs, inflight):
    """Clips of `jobs` in completion order, at most `inflight` outstanding."""
    pending, it = [], iter(jobs)
    for job in it:
        pending.append(pool.apply_async(_clip, (job,)))
        if len(pending) >= inflight:
            break
    while pending:
        r = pending.pop(0)
        for job in it:
            pending.append(pool.apply_async(_clip, (job,)))
            break
        yield r.get()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=20000, help="training frame pairs per epoch")
    ap.add_argument("--w", type=int, default=1920)
    ap.add_argument("--h", type=int, default=1088)
    ap.add_argument("--batch", type=int, default=15)
    ap.add_argument("--shuffle-clips", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=1)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=500000)
    ap.add_argument("--val-clips", type=int, default=2)
    ap.add_argument("--workers", type=int, default=min(os.cpu_count() or 1, 15))
    ap.add_argument("--max-seconds", type=float, default=0.0, help="stop training after this long (0: no limit)")
    ap.add_argument("--log-every", type=int, default=50)
    ap.add_argument("--init", default=None, help="start from these parameters (.npz)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args()

    w, h, B = args.w, args.h, args.batch
    n_clips = (args.pairs + CLIP_FRAMES - 2) // (CLIP_FRAMES - 1)
    train_jobs = [(args.seed + i, w, h) for i in range(n_clips)]
    val_jobs = [(args.seed + 10_000_000 + i, w, h) for i in range(args.val_clips)]
    t0 = time.time()
    # the worker pool is forked before this process touches the GPU
    pool = mp.get_context("fork").Pool(args.workers)
    val = [c for _, c in pool.map(_clip, val_jobs)]
    print(f"pool of {args.workers}; {len(val)} validation clips in {time.time() - t0:.1f}s", flush=True)

    from rusty_h264_amd import accel, mvnet
    nblk = (w // 16) * (h // 16)
    pad = 32
    log = []
    with accel.Device(0) as dev:
        p = accel.search_params(w, h, pad, 16, 16, 32, 0, lam=accel.lambda_me(26))
        ms = accel.MotionSearch(dev, p)
        nf = CLIP_FRAMES - 1
        clip_buf = dev.alloc(CLIP_FRAMES * w * h)
        planes = dev.alloc(nf * accel.planes_bytes(w, h, pad))

        def labels_of(clip):
            """Full-search bins of frames 1..15 against frames 0..14."""
            dev.lib.rh264_h2d(dev.ctx, clip_buf.ptr, clip.ctypes.data, clip.nbytes)
            accel.build_planes(dev, clip_buf, nf, w, h, pad, out=planes)
            ms.run_full(clip_buf.ptr + w * h, planes, nf)  # frames 1.. of the clip
            return mvnet.mv_bins(ms.results(nf)["int_mv"].reshape(nf, nblk, 2))

        init = mvnet.load_weights(args.init) if args.init else mvnet.init_params(args.seed)
        net = mvnet.MVNet(dev, B, w, h, init)
        fb = dev.alloc(2 * B * w * h)
        lb = dev.alloc(4 * B * nblk)
        t8 = dev.alloc(4 * B * nblk * 8)
        val_lab = [labels_of(c) for c in val]

        def validate():
            hit1 = hit8 = cnt = 0
            for c, lab in zip(val, val_lab):
                for s in range(0, nf - B + 1, B):
                    batch = np.ascontiguousarray(np.concatenate([c[1 + s:1 + s + B], c[s:s + B]]))
                    dev.lib.rh264_h2d(dev.ctx, fb.ptr, batch.ctypes.data, batch.nbytes)
                    net.forward(int(fb), int(fb) + B * w * h, w * h, B, w, h, None, t8)
                    top = dev.download(t8, (B, nblk, 8), np.int32)
                    want = lab[s:s + B]
                    hit1 += int((top[..., 0] == want).sum())
                    hit8 += int((top == want[..., None]).any(-1).sum())
                    cnt += want.size
            return hit1 / max(cnt, 1), hit8 / max(cnt, 1)

        v1, v8 = validate()
        rec = {"step": 0, "val_top1": v1, "val_top8": v8, "seconds": round(time.time() - t0, 1)}
        log.append(rec)
        print(json.dumps(rec), flush=True)

        steps_per_epoch = (n_clips * nf) // B
        total = steps_per_epoch * args.epochs
        rng = np.random.default_rng(args.seed)
        step, stop, losses = 0, False, []
        t_train = time.time()
        for ep in range(args.epochs):
            order = list(train_jobs)
            rng.shuffle(order)
            buf_cur, buf_ref, buf_lab = [], [], []

            def drain():
                nonlocal step, stop
                if not buf_cur:
                    return
                cur = np.concatenate(buf_cur)
                ref = np.concatenate(buf_ref)
                lab = np.concatenate(buf_lab)
                perm = rng.permutation(cur.shape[0])
                nb = cur.shape[0] // B
                for j in range(nb):
                    idx = perm[j * B:(j + 1) * B]
                    batch = np.ascontiguousarray(np.concatenate([cur[idx], ref[idx]]))
                    dev.lib.rh264_h2d(dev.ctx, fb.ptr, batch.ctypes.data, batch.nbytes)
                    lab_b = np.ascontiguousarray(lab[idx])
                    dev.lib.rh264_h2d(dev.ctx, lb.ptr, lab_b.ctypes.data, lab_b.nbytes)
                    losses.append(net.backward(int(fb), int(fb) + B * w * h, w * h, B, w, h, lb))
                    frac = step / max(total, 1)
                    if args.max_seconds > 0:
                        frac = max(frac, (time.time() - t_train) / args.max_seconds)
                    net.adam(args.lr * 0.5 * (1.0 + np.cos(np.pi * min(frac, 1.0))) + 1e-6)
                    step += 1
                    if step % args.log_every == 0:
                        print(json.dumps({"step": step, "epoch": ep, "loss": float(np.mean(losses[-args.log_every:])),
                                          "seconds": round(time.time() - t0, 1)}), flush=True)
                    if args.max_seconds > 0 and time.time() - t_train > args.max_seconds:
                        stop = True
                        return
                buf_cur.clear()
                buf_ref.clear()
                buf_lab.clear()

            for _, clip in _stream(pool, order, 2 * args.workers):
                if stop:
                    break
                buf_lab.append(labels_of(clip))
                buf_cur.append(clip[1:])
                buf_ref.append(clip[:-1])
                if len(buf_cur) >= args.shuffle_clips:
                    drain()
            if not stop:
                drain()
            v1, v8 = validate()
            rec = {"step": step, "epoch": ep, "loss": float(np.mean(losses[-steps_per_epoch:])) if losses else None,
                   "val_top1": v1, "val_top8": v8, "seconds": round(time.time() - t0, 1)}
            log.append(rec)
            print(json.dumps(rec), flush=True)
            if stop:
                break
        params = net.params()
        net.close()
    pool.terminate()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, params=params, log=np.array(json.dumps(log)), args=np.array(json.dumps(vars(args))))
    print("wrote", args.out, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
