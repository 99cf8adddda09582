"""Trains the learned tile proposal on the device and evaluates it.

    python -m akaze_amd.train_proposal --width 1920 --height 1080 \
        --train 20000 --val 2000 --epochs 3 --out akaze_amd/proposal_1080p

Training images are device-rendered synthetic images (seeds 1..N), held-out
images use seeds from 10_000_000 on.  Labels are the tiles holding the plain
pipeline's kept keypoints.  The probability threshold is the largest one
whose keypoint recall on a calibration slice of the held-out set reaches
the target (default 0.98 + margin); recall and images/s are then measured on
the whole held-out set against the full scan.  Writes <out>.f32 (the 14513
parameters, little-endian float32) and <out>.json (threshold, recall, tile
fraction kept, throughput, loss curve).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import akaze_amd as A  # noqa: E402
from akaze_amd.proposal import (Dataset, ProposalNet, keypoint_recall, threshold_for_recall, tile_recall,  # noqa: E402
                                train)

VAL_SEED0 = 10_000_000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_images(bk, seed0, n, batch, net=None, thr=0.5, keep=False):
    """Runs n device-rendered images through the context; returns (images/s,
    list of keypoint arrays if keep)."""
    bk.set_proposal(net, thr)
    out = []
    t0 = time.perf_counter()
    for s in range(0, n, batch):
        k = min(batch, n - s)
        bk.synth(seed0 + s, k)
        bk.process_async(k)
        cnt, _ = bk.counts(k)
        if keep:
            for z in range(k):
                out.append(bk.read_results(z, int(cnt[z]))[0])
    return n / (time.perf_counter() - t0), out


def throughput(bk, batch, net, thr, reps=20):
    """Images/s of the batched pipeline on resident images (graph mode)."""
    bk.set_proposal(net, thr)
    bk.set_graph_mode(True)
    bk.synth(VAL_SEED0, batch)
    for _ in range(3):
        bk.process_async(batch)
    bk.wait()
    t0 = time.perf_counter()
    for _ in range(reps):
        bk.process_async(batch)
    bk.wait()
    dt = time.perf_counter() - t0
    bk.set_graph_mode(False)
    return reps * batch / dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--val", type=int, default=2000)
    ap.add_argument("--calib", type=int, default=256, help="held-out images used to pick the threshold")
    ap.add_argument("--epochs", type=float, default=3.0)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--pos-weight", type=float, default=4.0)
    ap.add_argument("--recall", type=float, default=0.98)
    ap.add_argument("--margin", type=float, default=0.005)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=os.path.join(HERE, "proposal_1080p"))
    args = ap.parse_args()
    W, H = args.width, args.height
    t_all = time.perf_counter()

    log("building training set: %d images %dx%d" % (args.train, W, H))
    t0 = time.perf_counter()
    tr = Dataset(W, H, args.train, seed0=1, batch=16, log=log)
    t_data = time.perf_counter() - t0
    log("training set ready in %.1f s, %d tiles per image, positive fraction %.4f" % (t_data, tr.P, tr.pos_frac))
    va = Dataset(W, H, args.val, seed0=VAL_SEED0, batch=16)
    log("held-out set ready, positive fraction %.4f" % va.pos_frac)

    net = ProposalNet(seed=args.seed)
    steps = int(args.epochs * args.train / args.batch)
    t0 = time.perf_counter()
    hist = train(net, tr, steps, batch=args.batch, lr=args.lr, pos_weight=args.pos_weight, seed=args.seed, log=log,
                 log_every=200)
    t_train = time.perf_counter() - t0
    log("trained %d steps in %.1f s (%.2f ms/step)" % (steps, t_train, 1e3 * t_train / steps))

    # threshold: start from high tile recall on the calibration slice, lower
    # until the keypoint recall of the pipeline reaches the target
    _, _, pcal = tile_recall(net, va, 0.5, nimg=args.calib)
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 16)
    _, full_cal = run_images(bk, VAL_SEED0, args.calib, 16, keep=True)
    chosen = None
    sweep = []
    for tile_target in (0.97, 0.98, 0.99, 0.995, 0.998, 0.999, 0.9995, 1.0):
        thr = threshold_for_recall(pcal, va.labels[:args.calib * va.P], tile_target)
        _, prop_cal = run_images(bk, VAL_SEED0, args.calib, 16, net=net, thr=thr, keep=True)
        rec = float(np.mean([keypoint_recall(f, p) for f, p in zip(full_cal, prop_cal)]))
        kept = float(np.mean(pcal >= thr))
        sweep.append({"tile_recall_target": tile_target, "threshold": thr, "kp_recall": rec, "tiles_kept": kept})
        log("threshold %.5f (tile recall %.4f): keypoint recall %.4f, tiles kept %.3f" % (thr, tile_target, rec, kept))
        if rec >= args.recall + args.margin:
            chosen = sweep[-1]
            break
    if chosen is None:
        chosen = {"threshold": 0.0, "kp_recall": 1.0, "tiles_kept": 1.0, "tile_recall_target": 1.0}
    thr = chosen["threshold"]

    # held-out evaluation on every image
    _, full = run_images(bk, VAL_SEED0, args.val, 16, keep=True)
    _, prop = run_images(bk, VAL_SEED0, args.val, 16, net=net, thr=thr, keep=True)
    recs = np.array([keypoint_recall(f, p) for f, p in zip(full, prop)])
    rec_all = float(sum(len(f) * r for f, r in zip(full, recs)) / max(sum(len(f) for f in full), 1))
    tile_rec, tiles_kept, _ = tile_recall(net, va, thr)
    ips_full = throughput(bk, 16, None, 0.5)
    ips_prop = throughput(bk, 16, net, thr)
    res = {
        "image": "%dx%d" % (W, H), "train_images": args.train, "heldout_images": args.val,
        "epochs": args.epochs, "batch": args.batch, "lr": args.lr, "pos_weight": args.pos_weight, "steps": steps,
        "threshold": thr, "heldout_kp_recall": rec_all, "heldout_kp_recall_min_image": float(recs.min()),
        "heldout_tile_recall": tile_rec, "tiles_kept": tiles_kept, "positive_tile_fraction": va.pos_frac,
        "images_per_s_batched_full_scan": ips_full, "images_per_s_batched_proposal": ips_prop,
        "train_seconds": t_train, "dataset_seconds": t_data, "ms_per_step": 1e3 * t_train / steps,
        "loss_first_100": float(np.mean(hist[:100])), "loss_last_100": float(np.mean(hist[-100:])),
        "loss_curve": [float(np.mean(hist[i:i + 100])) for i in range(0, len(hist), 100)],
        "threshold_sweep": sweep, "total_seconds": time.perf_counter() - t_all,
    }
    net.save(args.out + ".f32")
    with open(args.out + ".json", "w") as f:
        json.dump(res, f, indent=1)
    log(json.dumps({k: v for k, v in res.items() if k not in ("loss_curve", "threshold_sweep")}))


if __name__ == "__main__":
    main()
