"""The camera-stream workload: four streams (two 1280x720, two 1920x1080),
each a seeded base image under homographies drifting by at most 2 degrees
and 1% scale per frame plus noise of std 2/255, rendered and warped on the
device (akz_synth_device_warp).  Frames go through ragged batches of 16
(four per stream) with the learned tile proposal; every 8th frame of a
stream is audited by the full scan; every frame is matched 2-NN to its
predecessor in the stream on the device.  Checked here:

  * each audited full-scan frame equals the oracle on the same frame;
  * the proposal reproduces at least 98% of the audited keypoints;
  * every predecessor match equals the oracle's 2-NN on the same descriptors.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STREAMS = [(1280, 720), (1280, 720), (1920, 1080), (1920, 1080)]
CLASSES = [(1280, 720), (1920, 1080)]
FRAMES = 16


def drift(rng, w, h, n):
    Hs = [np.eye(3)]
    T0 = np.array([[1, 0, -w / 2.0], [0, 1, -h / 2.0], [0, 0, 1.0]])
    T1 = np.array([[1, 0, w / 2.0], [0, 1, h / 2.0], [0, 0, 1.0]])
    for _ in range(n - 1):
        th = np.deg2rad(rng.uniform(-2, 2))
        sc = 1.0 + rng.uniform(-0.01, 0.01)
        R = np.array([[sc * np.cos(th), -sc * np.sin(th), 0], [sc * np.sin(th), sc * np.cos(th), 0], [0, 0, 1.0]])
        Hs.append(T1 @ R @ T0 @ Hs[-1])
    return Hs


@pytest.mark.parametrize("work_lists", [False, True])
def test_camera_streams_with_proposal_audit_and_matching(gpu_lib, work_lists):
    A = gpu_lib
    from akaze_amd.proposal import ProposalNet, keypoint_recall
    rec = json.load(open(os.path.join(REPO, "akaze_amd", "proposal_1080p.json")))
    net = ProposalNet()
    net.load(os.path.join(REPO, "akaze_amd", "proposal_1080p.f32"))
    rng = np.random.default_rng(7)
    frames = []
    for s, (w, h) in enumerate(STREAMS):
        base = A.DeviceBuffer(w * h * 4)
        A._check(A.lib().akz_synth_device_batch(500 + s, 1, w, h, base.ptr, w * 4, 0), "akz_synth_device_batch")
        fr = []
        for t, Hm in enumerate(drift(rng, w, h, FRAMES)):
            b = A.DeviceBuffer(w * h * 4)
            Hd = np.ascontiguousarray(Hm, np.float64)
            A._check(A.lib().akz_synth_device_warp(base.ptr, w, h, w * 4, Hd.ctypes.data, 2.0 / 255.0, 1000 * s + t,
                                                   b.ptr, w * 4), "akz_synth_device_warp")
            fr.append(b)
        frames.append(fr)
    rg = A.Ragged(A.AKAZEOptions(), CLASSES, [8, 8])
    audit = A.Ragged(A.AKAZEOptions(), CLASSES, [2, 2])
    import ctypes
    for k in range(len(CLASSES)):
        c = ctypes.c_void_p()
        A._check(A.lib().akz_ragged_context(rg.r, k, ctypes.byref(c)), "akz_ragged_context")
        A._check(A.lib().akz_set_proposal(c, net.net, ctypes.c_float(rec["threshold"])), "akz_set_proposal")
    rg.set_work_lists(work_lists)
    audit.set_work_lists(work_lists)
    rg.set_graph_mode(True)
    audit.set_graph_mode(True)
    m = A.Matcher()
    dout = A.DeviceBuffer(2 * 65536 * 16)
    prev = [None] * len(STREAMS)        # host descriptors of each stream's last frame
    prev_dev = [A.DeviceBuffer(65536 * 64) for _ in STREAMS]
    recalls, n_matched, n_audited = [], 0, 0
    for t0 in range(0, FRAMES, 4):
        ptrs, pit, sz, who = [], [], [], []
        for s, (w, h) in enumerate(STREAMS):
            for t in range(t0, t0 + 4):
                ptrs.append(frames[s][t].ptr); pit.append(w * 4); sz.append((w, h)); who.append((s, t))
        print("stream batch t0=%d: 16 frames" % t0, flush=True)
        counts = rg.process_device(ptrs, pit, sz)
        res = [rg.read(i, int(counts[i]))[:2] for i in range(len(ptrs))]
        # predecessor matching on the device, checked against the oracle
        for i, (s, t) in enumerate(who):
            _, _, dq = rg.results_device(i)
            if t == t0:
                if prev[s] is None:
                    continue
                dt_ptr, tdesc = prev_dev[s].ptr, prev[s]
            else:
                _, _, dt_ptr = rg.results_device(i - 1)
                tdesc = res[i - 1][1]
            qdesc = res[i][1]
            m.match2nn_device(dq, len(qdesc), dt_ptr, len(tdesc), dout.ptr)
            m.wait()
            got = dout.download(np.zeros(2 * max(len(qdesc), 1), A.DM_DTYPE))[:2 * len(qdesc)].reshape(-1, 2)
            want = O.match2nn(qdesc, tdesc)
            assert np.array_equal(got["trainIdx"], want["trainIdx"]), (s, t)
            assert np.array_equal(got["distance"], want["distance"]), (s, t)
            assert np.array_equal(got["queryIdx"][:, 0], np.arange(len(qdesc))), (s, t)
            n_matched += 1
        for s in range(len(STREAMS)):
            i = 4 * s + 3
            _, _, d = rg.results_device(i)
            A._check(A.lib().akz_memcpy_d2d(prev_dev[s].ptr, d, len(res[i][1]) * 64), "akz_memcpy_d2d")
            prev[s] = res[i][1]
        # audit every 8th frame with the full scan; the audited frames are
        # checked against the oracle, and the proposal's recall against them
        if t0 % 8 == 0:
            aidx = [4 * s for s in range(len(STREAMS))]
            acounts = audit.process_device([ptrs[i] for i in aidx], [pit[i] for i in aidx], [sz[i] for i in aidx])
            for j, i in enumerate(aidx):
                w, h = sz[i]
                full_kps, full_desc, _ = audit.read(j, int(acounts[j]))
                host = frames[who[i][0]][who[i][1]].download(np.zeros((h, w), np.float32))
                okp, odesc = O.Pipeline(w, h).run(host)
                assert full_kps.tobytes() == okp.tobytes(), who[i]
                assert np.array_equal(full_desc, odesc), who[i]
                r = keypoint_recall(full_kps, res[i][0])
                recalls.append((len(full_kps), r))
                n_audited += 1
                print("audit stream %d frame %d (%dx%d): %d keypoints, proposal recall %.4f"
                      % (who[i][0], who[i][1], w, h, len(full_kps), r), flush=True)
    tot = sum(n for n, _ in recalls)
    recall = sum(n * r for n, r in recalls) / tot
    print("streams: %d audited frames, keypoint recall %.4f (min %.4f), %d predecessor matches checked"
          % (n_audited, recall, min(r for _, r in recalls), n_matched), flush=True)
    assert n_audited == len(STREAMS) * FRAMES // 8
    assert n_matched == len(STREAMS) * (FRAMES - 1)
    assert recall >= 0.98
