"""Throughput benchmark: A-KAZE detect + MLDB on 1920x1080 images on one MI355X.

A step is one resident 1080p image (seeded synthetic generator) through the
whole pipeline: contrast factor, nonlinear scale space (166 FED steps),
Hessian response, extrema, cross-level suppression, orientation and 486-bit
MLDB descriptors, with the keypoints and descriptors copied to host memory.
Three contexts rotate so the host readback of one image overlaps the next
images' kernels.  The timed region starts with the images already in HBM.

Prints one JSON line (the last line on stdout).  Extra objects:
  roofline      FED kernel (the dominant kernel, up to 8 steps per launch):
                algorithmic bytes (12 B per pixel and step) over its duration,
                measured with HIP events around the FED launches on the
                context stream, beside the bytes the launches must move and
                the counter traffic (see fed_roofline);
  cpu_baseline  the CPU oracle (the C restatement of the reference
                semantics) on bounded samples of the same workload, with all
                host threads and with one thread;
  batched       16-image batches (one launch per stage, HIP graph) with the
                full scan and with the committed learned tile proposal, and
                the proposal's keypoint recall against the full scan.
"""
import argparse
import ctypes
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


def fed_roofline(avg_ms, alg_bytes, least_bytes, traffic, basis, traffic_profile="fed_plain_1080p_pmc.md"):
    """Roofline object of the FED launches at the measured average launch.

    Two byte counts per launch:
      alg_bytes    12 B per pixel and step (read L and g, write L, as if every
                   step were a pass of its own; the per-step unit of the
                   FED roofline), plus 12 B per pixel for a launch that also
                   builds its level's Lsmooth and conductance;
      least_bytes  what the launches must move however many steps each runs
                   (L and g read once, L written once per launch).
    basis "per_step": `achieved` counts alg_bytes (one image per call: the
    launches are short and latency-bound, far from either bound);
    basis "per_launch": `achieved` counts least_bytes, because temporal
    blocking (several steps per launch) lets a launch do more steps than its
    HBM traffic pays for, and per-step bytes would then exceed the HBM peak.
    `traffic` is the HBM bytes per launch of the committed counter runs."""
    alg_rate = alg_bytes / (avg_ms * 1e-3) / 1e9
    least_rate = least_bytes / (avg_ms * 1e-3) / 1e9
    achieved = alg_rate if basis == "per_step" else least_rate
    r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "basis": basis,
         "kernel": "FED launches: k_fed_reg / k_fed_multi (first launch of a level: k_fed_reg_sf / k_fed_multi_sf, "
                   "with its smooth + flow), up to 8 steps each, halo = steps",
         "avg_launch_us": round(avg_ms * 1e3, 3), "alg_bytes_per_launch": round(alg_bytes),
         "least_bytes_per_launch": round(least_bytes),
         "alg_GBs": round(alg_rate, 1), "least_GBs": round(least_rate, 1),
         "least_frac": round(least_rate / HBM_PEAK_GBS, 4)}
    if traffic:
        rate = traffic / (avg_ms * 1e-3) / 1e9
        r["traffic_GBs"] = round(rate, 1)
        r["traffic_frac"] = round(rate / HBM_PEAK_GBS, 4)
        # not measured by this run: the bytes come from the counter runs
        # committed beside the kernels, divided by this run's launch time
        r["traffic_source"] = "profiles/%s (rocprofv3 FETCH_SIZE and WRITE_SIZE runs)" % traffic_profile
    return r


def fed_traffic(profile="fed_plain_1080p_pmc.md", key="traffic_bytes_per_launch"):
    """HBM bytes per FED launch from the committed FETCH_SIZE / WRITE_SIZE
    counter runs (scaled by the copy-kernel calibration), or None."""
    path = os.path.join(REPO, "profiles", profile)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def host_threads():
    """Threads this process may run on (the box's CPU share), capped by
    OMP_NUM_THREADS when that is set."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, env) if env > 0 else n


def cpu_rate(O, img, W, H, threads, budget_s):
    P = O.Pipeline(W, H, threads=threads)
    P.run(img)  # warm-up (page faults, thread pool)
    n, t0 = 0, time.perf_counter()
    while True:
        P.run(img)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            return n, el


def cpu_baseline(img, W, H, budget_s):
    """CPU oracle images/s on bounded samples (about `budget_s` of work each)
    with every host thread available and with one thread."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O
    threads = host_threads()
    n, el = cpu_rate(O, img, W, H, threads, budget_s)
    n1, el1 = cpu_rate(O, img, W, H, 1, budget_s)
    O.lib().orc_set_threads(threads)
    sample = "%dx%d synthetic image, full detect+MLDB on the CPU oracle (-O3 -march=x86-64-v3, OpenMP)" % (W, H)
    return {"value": n / el, "unit": "images/s", "cores": threads, "kind": "port",
            "sample": "%d images in %.1f s on %d threads; %s" % (n, el, threads, sample),
            "one_thread": {"value": n1 / el1, "unit": "images/s", "cores": 1,
                           "sample": "%d images in %.1f s on 1 thread" % (n1, el1)}}


def batched_comparison(A, W, H, B=16, reps=20, autotune=False):
    """Images/s of 16-image batches (graph mode, resident device-rendered
    images) with the full scan and with the committed learned proposal, and
    the proposal's keypoint recall on those images.  The FED variants are the
    built-in batch table unless `autotune` (the autotuner times each level's
    FED alone; in the whole pipeline its choices measured 2% slower)."""
    from akaze_amd.proposal import ProposalNet, keypoint_recall
    here = os.path.join(REPO, "akaze_amd")
    rec = json.load(open(os.path.join(here, "proposal_1080p.json")))
    net = ProposalNet()
    net.load(os.path.join(here, "proposal_1080p.f32"))
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), B)
    variants = bk.autotune(B) if autotune else bk.fed_variants()
    bk.synth(20_000_000, B)
    out = {"images_per_call": B, "threshold": rec["threshold"], "autotuned": bool(autotune),
           "fed_variants": [int(v) for v in variants]}
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
    # FED roofline of the batched launches (live HIP events, as for the headline)
    bk.set_proposal(None)
    fms, fby, fl, fmin = min(bk.profile_fed(B), bk.profile_fed(B))
    out["fed_roofline"] = fed_roofline(fms / fl, fby / fl, fmin / fl,
                                       fed_traffic("akaze_batched16_1080p_kernel_stats.md",
                                                   "traffic_bytes_per_launch_batched16")
                                       if (W, H, B) == (1920, 1080, 16) else None, "per_launch",
                                       "akaze_batched16_1080p_kernel_stats.md")
    full, prop = kps["full_scan"], kps["learned_proposal"]
    total = sum(len(f) for f in full)
    out["proposal_kp_recall"] = round(sum(len(f) * keypoint_recall(f, p) for f, p in zip(full, prop)) / total, 4)
    out["heldout_kp_recall_2000_images"] = round(rec["heldout_kp_recall"], 4)
    out["proposal_tiles_kept"] = round(rec["tiles_kept"], 4)
    return out


RAGGED_SIZES = [(320, 240), (640, 480), (1280, 720), (1920, 1080), (2560, 1440)]


def device_image(A, seed, w, h):
    buf = A.DeviceBuffer(w * h * 4)
    A._check(A.lib().akz_synth_device_batch(seed, 1, w, h, buf.ptr, w * 4, 0), "akz_synth_device_batch")
    return buf


def ragged_workload(A, args):
    """64 images, sizes drawn uniformly from five classes, processed as one
    ragged call (per-class batches on concurrent streams, HIP graphs), then
    each image matched 2-NN to the next on the device."""
    rng = np.random.default_rng(2)
    picks = rng.integers(0, len(RAGGED_SIZES), 64)
    caps = [max(1, int(np.sum(picks == k))) for k in range(len(RAGGED_SIZES))]
    rg = A.Ragged(A.AKAZEOptions(), RAGGED_SIZES, caps)
    if args.autotune:
        rg.autotune()
    rg.set_work_lists(args.work_lists)
    rg.set_graph_mode(True)
    sizes = [RAGGED_SIZES[k] for k in picks]
    bufs = [device_image(A, 300 + i, w, h) for i, (w, h) in enumerate(sizes)]
    ptrs = [b.ptr for b in bufs]
    pitches = [w * 4 for (w, h) in sizes]
    m = A.Matcher()
    dout = A.DeviceBuffer(2 * 65536 * 16)
    npairs = [0]

    def step():
        counts = rg.process_device(ptrs, pitches, sizes)
        pairs = 0
        for i in range(63):
            _, _, dq = rg.results_device(i)
            _, _, dt = rg.results_device(i + 1)
            m.match2nn_device(dq, int(counts[i]), dt, int(counts[i + 1]), dout.ptr)
            pairs += int(counts[i]) * int(counts[i + 1])
        m.wait()
        npairs[0] = pairs
        return counts

    for _ in range(args.warmup):
        counts = step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dt = time.perf_counter() - t0
    mpix = sum(w * h for (w, h) in sizes) / 1e6
    value = 64 * args.steps / dt
    log("ragged: %d calls of 64 images in %.3f s -> %.1f images/s, %.1f Mpixel/s; class counts %s; keypoints %d..%d"
        % (args.steps, dt, value, mpix * args.steps / dt, caps, int(counts.min()), int(counts.max())))
    return {"value": round(value, 2), "unit": "images/s", "ms_per_step": round(1e3 * dt / args.steps, 4),
            "config": {"workload": "akaze_ragged64_5sizes_detect_mldb_match", "images_per_step": 64,
                       "sizes": ["%dx%d" % s for s in RAGGED_SIZES], "class_counts": caps,
                       "megapixels_per_step": round(mpix, 2), "match_pairs_per_step": npairs[0],
                       "autotuned": bool(args.autotune), "graph": True, "work_lists": bool(args.work_lists)}}


def drift_homographies(rng, w, h, n):
    """Frame-to-reference homographies drifting by <= 2 deg and <= 1% scale per frame."""
    Hs = [np.eye(3)]
    cx, cy = w / 2.0, h / 2.0
    T0 = np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, 1.0]])
    T1 = np.array([[1, 0, cx], [0, 1, cy], [0, 0, 1.0]])
    for _ in range(n - 1):
        th = np.deg2rad(rng.uniform(-2, 2))
        sc = 1.0 + rng.uniform(-0.01, 0.01)
        R = np.array([[sc * np.cos(th), -sc * np.sin(th), 0], [sc * np.sin(th), sc * np.cos(th), 0], [0, 0, 1.0]])
        Hs.append(T1 @ R @ T0 @ Hs[-1])
    return Hs


def streams_workload(A, args):
    """Four camera streams (two 1280x720, two 1920x1080) of 64 frames each:
    a seeded base image under drifting homographies plus noise of std 2/255.
    Ragged batches of 16 frames (4 per stream) run with the learned tile
    proposal; every 8th frame of a stream is audited by the full scan, whose
    keypoints measure the proposal's recall; every frame is matched 2-NN to its
    predecessor in the stream."""
    from akaze_amd.proposal import ProposalNet, keypoint_recall
    here = os.path.join(REPO, "akaze_amd")
    rec = json.load(open(os.path.join(here, "proposal_1080p.json")))
    net = ProposalNet()
    net.load(os.path.join(here, "proposal_1080p.f32"))
    ssz = [(1280, 720), (1280, 720), (1920, 1080), (1920, 1080)]
    classes = [(1280, 720), (1920, 1080)]
    F = args.frames
    rng = np.random.default_rng(7)
    frames = []
    for s, (w, h) in enumerate(ssz):
        base = device_image(A, 500 + s, w, h)
        fr = []
        for t, Hm in enumerate(drift_homographies(rng, w, h, F)):
            b = A.DeviceBuffer(w * h * 4)
            Hd = np.ascontiguousarray(Hm, np.float64)
            A._check(A.lib().akz_synth_device_warp(base.ptr, w, h, w * 4, Hd.ctypes.data, 2.0 / 255.0,
                                                   1000 * s + t, b.ptr, w * 4), "akz_synth_device_warp")
            fr.append(b)
        frames.append(fr)
    rg = A.Ragged(A.AKAZEOptions(), classes, [8, 8])
    audit = A.Ragged(A.AKAZEOptions(), classes, [2, 2])
    rg.set_work_lists(args.work_lists)
    audit.set_work_lists(args.work_lists)
    for k in range(2):
        c = ctypes.c_void_p()
        A._check(A.lib().akz_ragged_context(rg.r, k, ctypes.byref(c)), "akz_ragged_context")
        A._check(A.lib().akz_set_proposal(c, net.net, ctypes.c_float(rec["threshold"])), "akz_set_proposal")
    rg.set_graph_mode(True)
    audit.set_graph_mode(True)
    m = A.Matcher()
    dout = A.DeviceBuffer(2 * 65536 * 16)
    prev = [A.DeviceBuffer(65536 * 64) for _ in ssz]
    prev_n = [0] * len(ssz)
    recalls = []

    def run_all(measure_recall):
        nonlocal prev_n
        prev_n = [0] * len(ssz)
        for t0 in range(0, F, 4):
            ptrs, pit, sz, who = [], [], [], []
            for s, (w, h) in enumerate(ssz):
                for t in range(t0, t0 + 4):
                    ptrs.append(frames[s][t].ptr); pit.append(w * 4); sz.append((w, h)); who.append((s, t))
            counts = rg.process_device(ptrs, pit, sz)
            # 2-NN of every frame against its predecessor in the stream
            for i, (s, t) in enumerate(who):
                _, _, dq = rg.results_device(i)
                if t == t0:
                    if prev_n[s]:
                        m.match2nn_device(dq, int(counts[i]), prev[s].ptr, prev_n[s], dout.ptr)
                else:
                    _, _, dt = rg.results_device(i - 1)
                    m.match2nn_device(dq, int(counts[i]), dt, int(counts[i - 1]), dout.ptr)
            m.wait()
            for s in range(len(ssz)):
                i = 4 * s + 3
                _, _, d = rg.results_device(i)
                n = int(counts[i])
                A._check(A.lib().akz_memcpy_d2d(prev[s].ptr, d, n * 64), "akz_memcpy_d2d")
                prev_n[s] = n
            # audit: every 8th frame of every stream with the full scan
            if t0 % 8 == 0:
                aidx = [4 * s for s in range(len(ssz))]
                audit.process_device([ptrs[i] for i in aidx], [pit[i] for i in aidx], [sz[i] for i in aidx])
                if measure_recall:
                    for j, i in enumerate(aidx):
                        full = audit.read(j, 65536)[0]
                        prop = rg.read(i, 65536)[0]
                        recalls.append((len(full), keypoint_recall(full, prop)))

    run_all(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_all(False)
    dt = time.perf_counter() - t0
    nfr = len(ssz) * F * args.steps
    value = nfr / dt
    tot = sum(n for n, _ in recalls)
    recall = sum(n * r for n, r in recalls) / max(tot, 1)
    log("streams: %d frames in %.3f s -> %.1f frames/s; audited keypoint recall %.4f (min %.4f over %d frames)"
        % (nfr, dt, value, recall, min(r for _, r in recalls), len(recalls)))
    return {"value": round(value, 2), "unit": "frames/s", "ms_per_step": round(1e3 * dt / args.steps, 4),
            "config": {"workload": "akaze_streams4x%d_720p_1080p_proposal_audit_match" % F,
                       "frames_per_step": len(ssz) * F, "batch": 16, "audit_every": 8,
                       "proposal_threshold": rec["threshold"], "audited_kp_recall": round(recall, 4),
                       "audited_kp_recall_min": round(min(r for _, r in recalls), 4), "graph": True}}


def match_workload(A, args):
    """2-NN Hamming matching of MLDB descriptors: 3840x2160 image against
    its homography warp (about 8k x 8k), and 50k x 50k random descriptors."""
    W, H = 3840, 2160
    base = device_image(A, 77, W, H)
    warp = A.DeviceBuffer(W * H * 4)
    Hm = np.ascontiguousarray(drift_homographies(np.random.default_rng(3), W, H, 6)[-1], np.float64)
    A._check(A.lib().akz_synth_device_warp(base.ptr, W, H, W * 4, Hm.ctypes.data, 2.0 / 255.0, 9, warp.ptr, W * 4),
             "akz_synth_device_warp")
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 2)
    ptr, pitch, zs = bk.input_buffer()
    for z, b in enumerate((base, warp)):
        A._check(A.lib().akz_memcpy_2d(ptr + z * zs, pitch, b.ptr, W * 4, W * 4, H), "akz_memcpy_2d")
    bk.process_async(2)
    counts, _ = bk.counts(2)
    res = []
    m = A.Matcher()
    rng = np.random.default_rng(5)
    R = rng.integers(0, 256, (50000, 64), dtype=np.uint8)
    R[:, 61:] = 0
    R[:, 60] &= 0x3F
    dR = A.DeviceBuffer(R.nbytes)
    dR.upload(R)
    out = A.DeviceBuffer(2 * 50000 * 16)
    for name in ("image_vs_warp_4k", "random_50k"):
        if name == "random_50k":
            dq, nq, dt, nt = dR.ptr, 50000, dR.ptr, 50000
        else:
            _, _, dq = bk.results_device(0)
            _, _, dt = bk.results_device(1)
            nq, nt = int(counts[0]), int(counts[1])
        m.match2nn_device(dq, nq, dt, nt, out.ptr)
        m.wait()
        reps = max(1, int(2e11 // max(nq * nt, 1)))
        reps = min(reps, 200)
        t0 = time.perf_counter()
        ms = 0.0
        for _ in range(reps):
            m.match2nn_device(dq, nq, dt, nt, out.ptr)
            m.wait()
            ms += m.last_ms()
        wall = time.perf_counter() - t0
        res.append({"case": name, "nq": nq, "nt": nt, "reps": reps, "kernel_ms": round(ms / reps, 4),
                    "pairs_per_s": round(nq * nt / (ms / reps * 1e-3), 1),
                    "wall_pairs_per_s": round(nq * nt * reps / wall, 1)})
        log("match %s: %d x %d in %.3f ms -> %.3e pairs/s" % (name, nq, nt, ms / reps, nq * nt / (ms / reps * 1e-3)))
    big = res[-1]
    # integer VALU peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz; ~37 ops per
    # pair (16 xor, 16 popcount-accumulate, key build and three min/max)
    peak_ops = 256 * 4 * 16 * 2.4e9
    ach = big["pairs_per_s"] * 37
    return {"value": big["pairs_per_s"], "unit": "pairs/s", "ms_per_step": big["kernel_ms"],
            "config": {"workload": "akaze_match2nn_mldb486", "cases": res},
            "roofline": {"bound": "valu", "achieved": round(ach / 1e12, 2), "peak": round(peak_ops / 1e12, 2),
                         "unit": "Tops/s", "frac": round(ach / peak_ops, 4), "traffic": None,
                         "kernel": "k_match", "ops_per_pair": 37}}


def hbm_probe(nbytes=1 << 30, reps=5):
    """Practical HBM rates of the box: best of `reps` device-to-device copies,
    float sums (read only) and fills (write only) of a 1 GiB buffer, timed
    with events on torch's stream.  Context for the roofline fractions, which
    are taken against the 8 TB/s published peak."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    out = {"probe": "hbm", "bytes": nbytes}
    for name, fn, moved in (("copy", lambda: b.copy_(a), 2 * nbytes), ("read", lambda: a.sum(), nbytes),
                            ("write", lambda: b.fill_(2.0), nbytes)):
        fn()
        best = 1e30
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[name + "_GBs"] = round(moved / (best * 1e-3) / 1e9, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--graph", type=int, default=1, help="capture each context's pipeline in a HIP graph")
    ap.add_argument("--depth", type=int, default=3, help="contexts in flight (3 measured best on MI355X)")
    ap.add_argument("--batch", type=int, default=1, help="images per context call (one launch per stage)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched / learned-proposal comparison")
    ap.add_argument("--workload", default="plain", choices=["plain", "ragged", "streams", "match"],
                    help="plain: the headline 1080p single-image line; the others measure the ragged-batch, "
                         "camera-stream and matcher configurations")
    ap.add_argument("--frames", type=int, default=64, help="frames per stream (streams workload)")
    ap.add_argument("--work-lists", type=int, default=0,
                    help="ragged / streams workloads: persistent work-list kernels over the whole call (1) or one "
                         "batch call per size class on concurrent streams (0, faster on MI355X)")
    ap.add_argument("--autotune", action="store_true", help="autotune FED kernels first")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 256 MiB device copy kernel first (byte-counter calibration for profiles)")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only time 1 GiB device copy / read / fill passes (the practical HBM rates) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.hbm_probe:
        print(json.dumps(hbm_probe()), flush=True)
        return

    import akaze_amd as A
    if args.workload != "plain":
        fn = {"ragged": ragged_workload, "streams": streams_workload, "match": match_workload}[args.workload]
        r = fn(A, args)
        out = {"metric": METRIC, "value": r["value"], "unit": r["unit"], "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32" if args.workload != "match" else "u32",
               "data": "synthetic (device-rendered seeded images)", "config": r["config"],
               "roofline": r.get("roofline"), "cpu_baseline": None}
        print(json.dumps(out), flush=True)
        return
    W, H = args.width, args.height
    if A.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    img = A.to_float(A.synth_image(1, W, H))
    if args.copy_calibration:
        nf = 64 << 20
        src, dst = A.DeviceBuffer(nf * 4), A.DeviceBuffer(nf * 4)
        ms = ctypes.c_float()
        A._check(A.lib().akz_copy_kernel_run(src.ptr, dst.ptr, nf, 3, ctypes.byref(ms)), "akz_copy_kernel_run")
        log("copy calibration: 3 x 256 MiB copies, %.3f ms each, %.1f GB/s" % (ms.value, 2 * nf * 4 / ms.value / 1e6))
        del src, dst

    B = max(1, args.batch)
    ctxs = []
    for _ in range(args.depth):
        bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), B)
        if args.autotune:
            if ctxs:
                bk.set_fed_variants(ctxs[0].fed_variants())
            else:
                log("autotuned FED variants per level: %s" % list(bk.autotune(B)))
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
    fed_ms, fed_bytes, fed_launches, fed_min = min(bk.profile_fed(B), bk.profile_fed(B))
    avg_ms = fed_ms / fed_launches
    # the counter runs were made on one 1920x1080 image per call
    roofline = fed_roofline(avg_ms, fed_bytes / fed_launches, fed_min / fed_launches,
                            fed_traffic() if (W, H, B) == (1920, 1080, 1) else None,
                            "per_step" if B == 1 else "per_launch")
    log("FED: %d launches, %.3f ms per call of %d images, avg launch %.2f us, %.1f GB/s per-step algorithmic, "
        "%.1f GB/s must-move" % (fed_launches, fed_ms, B, avg_ms * 1e3, roofline["alg_GBs"], roofline["least_GBs"]))

    batched = None
    if not args.no_batched and B == 1:
        batched = batched_comparison(A, W, H, autotune=args.autotune)
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
                      "autotuned": bool(args.autotune),
                      "keypoints_per_image": int(np.median(counts)) if counts else 0},
           "roofline": roofline, "cpu_baseline": cpu, "batched": batched}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
