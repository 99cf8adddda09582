"""GPU checks of the learned tile proposal against the CPU oracle.

Features and labels of the device pipeline equal the oracle's; the bf16
MFMA forward pass and gradients agree with the oracle's bf16 emulation
within bf16 rounding noise (and with its exact double-precision mode in
direction); the Adam step is the oracle's update of the device gradient;
detection with a tile mask equals the oracle's suppression, orientation and
MLDB run on the candidates the mask keeps; a threshold of 0 reproduces the
full scan bit for bit; and a short training run learns the labels.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def prop_geom_of(ctx):
    levels = []
    for l in range(ctx.nlev):
        w, h, o = ctx.level_size(l)[:3]
        levels.append(((w + 15) // 16, (h + 15) // 16, o))
    return O.make_prop_geom(levels)


def test_features_and_labels_match_oracle(gpu_lib):
    A = gpu_lib
    W, H = 640, 480
    imgs = [A.to_float(A.synth_image(21 + i, W, H)) for i in range(2)]
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 2)
    bk.set_keep_determinant(True)
    bk.upload(imgs)
    bk.process_async(2)
    P = bk.proposal_positions()
    feat = A.DeviceBuffer(2 * P * 8)
    lab = A.DeviceBuffer(2 * P)
    bk.extract(2, feat.ptr, lab.ptr)
    f = np.zeros(2 * P * 4, np.uint16)
    l = np.zeros(2 * P, np.uint8)
    feat.download(f)
    lab.download(l)
    for z in range(2):
        orc = O.Pipeline(W, H)
        orc.scale_space(imgs[z])
        orc.detect()
        assert orc.prop_geom().P == P
        assert np.array_equal(l[z * P:(z + 1) * P], orc.prop_labels()), z
        of = O.bf16_to_f32(orc.prop_features())
        gf = O.bf16_to_f32(f[z * P * 4:(z + 1) * P * 4])
        diff = np.abs(gf - of)
        # det channels: same maxima, log2 may differ in the last float bit;
        # the Lt mean sums in another order: at most one bf16 step apart
        assert np.all(diff <= np.abs(of) * 2 ** -7), z
        assert np.mean(diff > 0) < 0.01, (z, np.mean(diff > 0))


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(O.PROP_NPARAM, np.float32)
    OFF = O.PROP_OFF
    for name, cin, cout in (("W1", 4, 16), ("W2", 16, 32), ("W3", 32, 32)):
        n = 9 * cin * cout
        p[OFF[name]:OFF[name] + n] = rng.normal(0, np.sqrt(2.0 / (9 * cin)), n)
    for name, n in (("b1", 16), ("b2", 32), ("b3", 32)):
        p[OFF[name]:OFF[name] + n] = rng.normal(0, 0.1, n)
    p[OFF["w4"]:OFF["w4"] + 32] = rng.normal(0, 0.3, 32)
    p[OFF["b4"]] = -0.3
    return p


def test_forward_gradient_and_adam_match_oracle(gpu_lib):
    A = gpu_lib
    from akaze_amd.proposal import ProposalNet
    W, H = 320, 240
    nimg = 3
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 1)
    P = bk.proposal_positions()
    g = prop_geom_of(bk)
    assert g.P == P
    rng = np.random.default_rng(4)
    x0 = O.f32_to_bf16(rng.random(nimg * P * 4).astype(np.float32))
    lab = (rng.random(nimg * P) < 0.25).astype(np.uint8)
    feat = A.DeviceBuffer(x0.nbytes)
    feat.upload(x0)
    dlab = A.DeviceBuffer(lab.nbytes)
    dlab.upload(lab)
    net = ProposalNet(seed=3)
    params = random_params(8)
    net.set_params(params)
    prob = net.forward(bk, feat.ptr, nimg)
    _, _, oprob = O.prop_loss_grad(g, nimg, params, x0, None, emulate=True, want_grad=False)
    d = np.abs(prob - oprob)
    assert d.max() < 2e-2 and d.mean() < 2e-3, (d.max(), d.mean())
    loss, grad = net.train_step(bk, feat.ptr, dlab.ptr, nimg, pos_weight=2.0, lr=0.0, want_grad=True)
    oloss, ograd, _ = O.prop_loss_grad(g, nimg, params, x0, lab, wpos=2.0, emulate=True)
    assert abs(loss - oloss) <= 1e-3 * abs(oloss), (loss, oloss)
    OFF = O.PROP_OFF
    bounds = [OFF["W1"], OFF["b1"], OFF["W2"], OFF["b2"], OFF["W3"], OFF["b3"], OFF["w4"], OFF["b4"],
              O.PROP_NPARAM]
    for a, b in zip(bounds[:-1], bounds[1:]):
        gg, oo = grad[a:b].astype(np.float64), ograd[a:b]
        rel = np.linalg.norm(gg - oo) / max(np.linalg.norm(oo), 1e-30)
        assert rel < 3e-2, (a, rel)
    _, exact, _ = O.prop_loss_grad(g, nimg, params, x0, lab, wpos=2.0, emulate=False)
    cos = float(grad.astype(np.float64) @ exact / np.linalg.norm(grad) / np.linalg.norm(exact))
    assert cos > 0.995, cos
    # one Adam step: the oracle's update of the device gradient
    lr = 1e-3
    loss1, g1 = net.train_step(bk, feat.ptr, dlab.ptr, nimg, pos_weight=2.0, lr=lr, want_grad=True)
    after = net.get_params()
    expect, _, _ = O.prop_adam(params, np.zeros_like(params), np.zeros_like(params), g1.astype(np.float64), lr, 1)
    assert np.allclose(after, expect, rtol=1e-6, atol=1e-9)


def masked_oracle(A, orc, probs, thr, g):
    """Oracle detection restricted to the tiles with prob >= thr."""
    cand = orc.candidates()
    ls = orc.level_start()
    keep = np.zeros(len(cand), bool)
    for i, c in enumerate(cand):
        l = int(c["class_id"])
        r = 1 << g.oct[l]
        x, y = int(c["x"]) // r, int(c["y"]) // r
        keep[i] = probs[g.base[l] + (y // 16) * g.tw[l] + x // 16] >= thr
    sub = cand[keep]
    nls = np.array([keep[:ls[l]].sum() for l in range(len(ls))], np.int32)
    kps, _ = O.nms(sub, nls)
    return O.orient_describe(orc, kps)


def test_masked_detection_equals_oracle(gpu_lib):
    A = gpu_lib
    from akaze_amd.proposal import ProposalNet
    W, H = 640, 480
    img = A.to_float(A.synth_image(33, W, H))
    orc = O.Pipeline(W, H)
    okp, odesc = orc.run(img)
    net = ProposalNet(seed=5)
    net.set_params(random_params(12))
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 1)
    bk.upload([img])
    # threshold 0 keeps every tile: the full scan, bit for bit
    bk.set_proposal(net, 0.0)
    bk.process_async(1)
    kps, desc, n = bk.read_results(0, len(okp) + 16)
    assert kps.tobytes() == okp.tobytes() and np.array_equal(desc, odesc)
    probs = bk.read_probs(0)
    g = prop_geom_of(bk)
    for q in (0.5, 0.8):
        thr = float(np.quantile(probs, q))
        bk.set_proposal(net, thr)
        bk.process_async(1)
        kps, desc, n = bk.read_results(0, len(okp) + 16)
        assert np.array_equal(bk.read_probs(0), probs)
        mk, md = masked_oracle(A, orc, probs, thr, g)
        assert 0 < n < len(okp), (q, n, len(okp))
        assert kps.tobytes() == mk.tobytes(), q
        assert np.array_equal(desc, md), q
    bk.set_proposal(None)
    bk.process_async(1)
    kps, desc, n = bk.read_results(0, len(okp) + 16)
    assert kps.tobytes() == okp.tobytes()


def test_detector_tile_maxima_match_full_determinant_features(gpu_lib):
    """With a proposal the fused detector keeps the determinant maxima of the
    16x16 tiles itself (order-preserving keys, atomicMax across its 30-row
    tiles) and no whole determinant image is written; the probabilities are
    those of a forward pass over features taken from whole determinant
    images.  Two image sets in turn: keys left from the first set would raise
    the second set's maxima."""
    A = gpu_lib
    from akaze_amd.proposal import ProposalNet
    W, H, B = 640, 480, 2
    sets = [[A.to_float(A.synth_image(41 + 2 * k + i, W, H)) for i in range(B)] for k in range(2)]
    net = ProposalNet(seed=7)
    net.set_params(random_params(13))
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), B)
    P = bk.proposal_positions()
    feat = A.DeviceBuffer(B * P * 8)
    refs = []
    bk.set_keep_determinant(True)
    for imgs in sets:
        bk.upload(imgs)
        bk.process_async(B)
        bk.extract(B, feat.ptr)
        refs.append(net.forward(bk, feat.ptr, B))
    bk.set_keep_determinant(False)
    bk.set_proposal(net, 0.5)
    for k, imgs in enumerate(sets):
        bk.upload(imgs)
        bk.process_async(B)
        for z in range(B):
            assert np.array_equal(bk.read_probs(z), refs[k][z * P:(z + 1) * P]), (k, z)
    bk.set_proposal(None)


def test_short_training_learns_the_labels(gpu_lib):
    """64 training and 32 held-out 640x480 device-rendered images: the loss
    falls and the held-out tile recall at the threshold that keeps 99% of the
    training positives stays high while tiles are skipped."""
    from akaze_amd.proposal import ProposalNet, Dataset, train, tile_recall, threshold_for_recall
    W, H = 640, 480
    tr = Dataset(W, H, 64, seed0=1000, batch=16)
    te = Dataset(W, H, 32, seed0=900000, batch=16)
    assert 0.01 < tr.pos_frac < 0.6, tr.pos_frac
    net = ProposalNet(seed=1)
    hist = train(net, tr, steps=300, batch=16, lr=3e-3, pos_weight=3.0)
    assert np.mean(hist[-30:]) < 0.8 * np.mean(hist[:10]), (hist[:10], hist[-30:])
    _, _, ptr = tile_recall(net, tr, 0.5)
    thr = threshold_for_recall(ptr, tr.labels, 0.99)
    rec, kept, _ = tile_recall(net, te, thr)
    assert rec > 0.95 and kept < 0.9, (rec, kept, thr)


def test_trained_network_recall_and_gain(gpu_lib):
    """The committed 1080p network on 32 held-out device-rendered images:
    keypoint recall against the full scan at the committed threshold, tiles
    kept against the candidate-tile rule (a tile is kept when its maximum
    response passes the detector threshold, which keeps every keypoint), and
    the images/s of both batched pipelines (printed)."""
    import json
    import os
    import time
    A = gpu_lib
    from akaze_amd.proposal import ProposalNet, keypoint_recall
    here = os.path.dirname(A.__file__)
    rec = json.load(open(os.path.join(here, "proposal_1080p.json")))
    net = ProposalNet()
    net.load(os.path.join(here, "proposal_1080p.f32"))
    W, H, B, n = 1920, 1080, 16, 32
    seed0 = 10_000_000 + 50_000
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), B)
    bk.set_keep_determinant(True)
    P = bk.proposal_positions()
    feat = A.DeviceBuffer(B * P * 8)
    full, prop, rule_kept, net_kept = [], [], [], []
    for s in range(0, n, B):
        bk.synth(seed0 + s, B)
        bk.set_proposal(None)
        bk.process_async(B)
        cnt, _ = bk.counts(B)
        full += [bk.read_results(z, int(cnt[z]))[0] for z in range(B)]
        bk.extract(B, feat.ptr)
        x = np.zeros(B * P * 4, np.uint16)
        feat.download(x)
        rule_kept.append(np.mean(O.bf16_to_f32(x).reshape(-1, 4)[:, 1] > 0.125))
        bk.set_proposal(net, rec["threshold"])
        bk.process_async(B)
        cnt, _ = bk.counts(B)
        prop += [bk.read_results(z, int(cnt[z]))[0] for z in range(B)]
        net_kept += [np.mean(bk.read_probs(z) >= rec["threshold"]) for z in range(B)]
    total = sum(len(f) for f in full)
    recall = sum(len(f) * keypoint_recall(f, p) for f, p in zip(full, prop)) / total
    speeds = {}
    for name, nn in (("full_scan", None), ("proposal", net)):
        bk.set_proposal(nn, rec["threshold"])
        bk.set_graph_mode(True)
        for _ in range(3):
            bk.process_async(B)
        bk.wait()
        t0 = time.perf_counter()
        for _ in range(10):
            bk.process_async(B)
        bk.wait()
        speeds[name] = 10 * B / (time.perf_counter() - t0)
        bk.set_graph_mode(False)
    print("held-out keypoint recall %.4f, tiles kept: network %.3f vs candidate-tile rule %.3f, images/s %s"
          % (recall, np.mean(net_kept), np.mean(rule_kept), {k: round(v, 1) for k, v in speeds.items()}))
    assert recall >= 0.98, recall
    assert np.mean(net_kept) < np.mean(rule_kept), (np.mean(net_kept), np.mean(rule_kept))
