"""Throughput benchmark: A-KAZE detect + MLDB on 1920x1080 images on one MI355X.

A step is one resident 1080p image (seeded synthetic generator) through the
whole pipeline: contrast factor, nonlinear scale space (166 FED steps),
Hessian response, extrema, cross-level suppression, orientation and 486-bit
MLDB descriptors, with the keypoints and descriptors copied to host memory.
Two contexts alternate so the host readback of one image overlaps the next
image's kernels.  The timed region starts with the images already in HBM.

Prints one JSON line (the last line on stdout).  Extra objects:
  roofline      FED kernel (the dominant kernel, up to 8 steps per launch):
                algorithmic bytes (12 B per pixel and step) over its duration,
                measured with HIP events around the FED launches on the
                context stream;
  cpu_baseline  the CPU oracle (the C restatement of the reference
                semantics, all host threads) on a bounded sample of the same
                workload;
  batched       16-image batches (one launch per stage, HIP graph) with the
                full scan and with the committed learned tile proposal, and
                the proposal's keypoint recall against the full scan.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
METRIC = "1080p images/s detect+MLDB: plain vs batched+learned proposal @>=98% kp recall"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(img, W, H, budget_s):
    """CPU oracle images/s on a bounded sample (at most `budget_s` of work)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    P = O.Pipeline(W, H, threads=threads)
    P.run(img)  # warm-up (page faults, thread pool)
    n, t0 = 0, time.perf_counter()
    while True:
        P.run(img)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return {"value": n / el, "unit": "images/s", "cores": threads, "kind": "port",
            "sample": "%d x %dx%d synthetic image, full detect+MLDB on the CPU oracle (%.1f s)" % (n, W, H, el)}


def batched_comparison(A, W, H, B=16, reps=20):
    """Images/s of 16-image batches (graph mode, resident device-rendered
    images) with the full scan and with the committed learned proposal, and
    the proposal's keypoint recall on those images."""
    from akaze_amd.proposal import ProposalNet, keypoint_recall
    here = os.path.join(REPO, "akaze_amd")
    rec = json.load(open(os.path.join(here, "proposal_1080p.json")))
    net = ProposalNet()
    net.load(os.path.join(here, "proposal_1080p.f32"))
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), B)
    bk.synth(20_000_000, B)
    out = {"images_per_call": B, "threshold": rec["threshold"]}
    kps = {}
    for name, nn in (("full_scan", None), ("learned_proposal", net)):
        bk.set_proposal(nn, rec["threshold"])
        bk.process_async(B)
        cnt, _ = bk.counts(B)
        kps[name] = [bk.read_results(z, int(cnt[z]))[0] for z in range(B)]
        bk.set_graph_mode(True)
        for _ in range(3):
            bk.process_async(B)
        bk.wait()
        t0 = time.perf_counter()
        for _ in range(reps):
            bk.process_async(B)
        bk.wait()
        out[name + "_images_per_s"] = round(reps * B / (time.perf_counter() - t0), 1)
        bk.set_graph_mode(False)
    full, prop = kps["full_scan"], kps["learned_proposal"]
    total = sum(len(f) for f in full)
    out["proposal_kp_recall"] = round(sum(len(f) * keypoint_recall(f, p) for f, p in zip(full, prop)) / total, 4)
    out["heldout_kp_recall_2000_images"] = round(rec["heldout_kp_recall"], 4)
    out["proposal_tiles_kept"] = round(rec["tiles_kept"], 4)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--graph", type=int, default=1, help="capture each context's pipeline in a HIP graph")
    ap.add_argument("--depth", type=int, default=2, help="contexts in flight")
    ap.add_argument("--batch", type=int, default=1, help="images per context call (one launch per stage)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched / learned-proposal comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import akaze_amd as A
    W, H = args.width, args.height
    if A.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    img = A.to_float(A.synth_image(1, W, H))

    B = max(1, args.batch)
    ctxs = []
    for _ in range(args.depth):
        bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), B)
        bk.upload([img] * B)
        bk.set_graph_mode(args.graph)
        ctxs.append(bk)

    pending = [False] * len(ctxs)
    counts = []

    def collect(bk):
        cnt, _ = bk.counts(B)
        for z in range(B):
            kps, desc, n = bk.read_results(z, int(cnt[z]))
            counts.append(n)

    def step(i):
        k = i % len(ctxs)
        bk = ctxs[k]
        if pending[k]:
            collect(bk)
        bk.process_async(B)
        pending[k] = True

    def drain():
        for k, bk in enumerate(ctxs):
            if pending[k]:
                collect(bk)
                pending[k] = False

    for i in range(args.warmup):
        step(i)
    drain()
    A.lib().akz_device_synchronize()
    counts.clear()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    drain()
    A.lib().akz_device_synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = args.steps * B / (t1 - t0)
    log("timed: %d calls of %d images in %.3f s -> %.1f images/s (%.3f ms/call), keypoints/image %s" %
        (args.steps, B, t1 - t0, value, ms, sorted(set(counts))))

    # roofline of the dominant kernel (FED step), live HIP events on the context stream
    bk = ctxs[0]
    fed_ms, fed_bytes, fed_launches = bk.profile_fed(B)
    fed_ms2, _, _ = bk.profile_fed(B)
    fed_ms = min(fed_ms, fed_ms2)
    avg_ms = fed_ms / fed_launches
    achieved = (fed_bytes / fed_launches) / (avg_ms * 1e-3) / 1e9
    log("FED: %d launches, %.3f ms per call of %d images, avg launch %.2f us, %.1f GB/s algorithmic" %
        (fed_launches, fed_ms, B, avg_ms * 1e3, achieved))
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_fed_multi (up to 8 FED steps per launch)", "avg_launch_us": round(avg_ms * 1e3, 3),
                "alg_bytes_per_launch": round(fed_bytes / fed_launches)}

    batched = None
    if not args.no_batched and B == 1:
        batched = batched_comparison(A, W, H)
        log("batched: %s" % batched)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(img, W, H, args.cpu_budget)
        log("cpu baseline: %s" % cpu)

    out = {"metric": METRIC, "value": round(value, 2), "unit": "images/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded generator: blobs, shapes, 1/f noise)",
           "config": {"workload": ("akaze_plain_detect_mldb_%dx%d" % (W, H)) if B == 1 else
                      ("akaze_batched_detect_mldb_%dx%d_x%d" % (W, H, B)), "image": "%dx%d" % (W, H),
                      "octaves": 4, "sublevels": 4, "diffusivity": "PM_G2", "descriptor": "MLDB-486",
                      "graph": bool(args.graph), "contexts_in_flight": args.depth, "images_per_call": B,
                      "keypoints_per_image": int(np.median(counts)) if counts else 0},
           "roofline": roofline, "cpu_baseline": cpu, "batched": batched}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
