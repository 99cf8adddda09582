"""CPU checks of the learned-proposal oracle (oracle/proposal_oracle.c).

The proposal has no reference counterpart, so its oracle is pinned by
mathematics: central finite differences of the double-precision loss, closed
forms of hand-built networks (a bias-only head; single-tap identity paths
that expose the neighbour indexing and zero padding at map and level edges),
the first Adam step, and tile statistics recomputed with numpy.
"""
import numpy as np
import pytest

import oracle_lib as O

OFF = O.PROP_OFF
GEOM_LEVELS = [(5, 4, 0), (5, 4, 0), (3, 2, 1)]


def random_problem(seed, nimg=2):
    rng = np.random.default_rng(seed)
    g = O.make_prop_geom(GEOM_LEVELS)
    p = np.zeros(O.PROP_NPARAM, np.float32)
    for name, cin, cout in (("W1", 4, 16), ("W2", 16, 32), ("W3", 32, 32)):
        n = 9 * cin * cout
        p[OFF[name]:OFF[name] + n] = rng.normal(0, np.sqrt(2.0 / (9 * cin)), n)
    p[OFF["b1"]:OFF["b1"] + 16] = rng.normal(0, 0.1, 16)
    p[OFF["b2"]:OFF["b2"] + 32] = rng.normal(0, 0.1, 32)
    p[OFF["b3"]:OFF["b3"] + 32] = rng.normal(0, 0.1, 32)
    p[OFF["w4"]:OFF["w4"] + 32] = rng.normal(0, 0.3, 32)
    p[OFF["b4"]] = -0.2
    x0 = O.f32_to_bf16(rng.random(g.P * nimg * 4).astype(np.float32))
    lab = (rng.random(g.P * nimg) < 0.3).astype(np.uint8)
    return g, p, x0, lab


def test_gradient_matches_finite_differences():
    g, p, x0, lab = random_problem(1)
    loss, grad, _ = O.prop_loss_grad(g, 2, p, x0, lab, wpos=2.5, emulate=False)
    assert np.isfinite(loss) and loss > 0
    rng = np.random.default_rng(2)
    idx = []
    for name, n in (("W1", 576), ("b1", 16), ("W2", 4608), ("b2", 32), ("W3", 9216), ("b3", 32), ("w4", 32)):
        idx += list(OFF[name] + rng.choice(n, size=min(n, 12), replace=False))
    idx.append(OFF["b4"])
    errs = []
    for i in idx:
        h = np.float32(1e-5) * max(np.float32(1.0), abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] = p[i] + h
        pm[i] = p[i] - h
        lp = O.prop_loss_grad(g, 2, pp, x0, lab, wpos=2.5, emulate=False, want_grad=False)[0]
        lm = O.prop_loss_grad(g, 2, pm, x0, lab, wpos=2.5, emulate=False, want_grad=False)[0]
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        errs.append(abs(fd - grad[i]) / max(1e-4, abs(grad[i]), abs(fd)))
    errs = np.array(errs)
    # a relu kink inside the +-h interval could spoil a central difference:
    # allow one of the ~90 sampled parameters to miss
    assert np.sum(errs > 1e-5) <= 1, np.sort(errs)[-8:]
    assert np.median(errs) < 1e-7


def test_bias_only_head_closed_form():
    g, _, x0, lab = random_problem(3)
    p = np.zeros(O.PROP_NPARAM, np.float32)
    p[OFF["b4"]] = 0.7
    wpos = 3.0
    loss, grad, prob = O.prop_loss_grad(g, 2, p, x0, lab, wpos=wpos, emulate=False)
    b = float(np.float32(0.7))
    s = 1.0 / (1.0 + np.exp(-b))
    sp = lambda z: np.log1p(np.exp(z))
    y = lab.astype(np.float64)
    assert np.allclose(prob, s, rtol=0, atol=1e-7)
    assert abs(loss - np.mean(wpos * y * sp(-b) + (1 - y) * sp(b))) < 1e-12
    assert abs(grad[OFF["b4"]] - np.mean(wpos * y * (s - 1) + (1 - y) * s)) < 1e-12
    assert np.all(grad[:OFF["b4"]] == 0.0)  # relu(0) blocks every other path


def tap_path(tap):
    """Channel 0 of x0 carried by tap `tap` of layer 1 and the centre taps of
    layers 2 and 3 to the head: z[p] = relu(x0[nbr(p, tap)][0])."""
    p = np.zeros(O.PROP_NPARAM, np.float32)
    p[OFF["W1"] + (tap * 4 + 0) * 16 + 0] = 1.0
    p[OFF["W2"] + (4 * 16 + 0) * 32 + 0] = 1.0
    p[OFF["W3"] + (4 * 32 + 0) * 32 + 0] = 1.0
    p[OFF["w4"] + 0] = 1.0
    return p


@pytest.mark.parametrize("tap", [4, 5, 1, 6])
def test_single_tap_paths(tap):
    g = O.make_prop_geom(GEOM_LEVELS)
    nimg = 2
    rng = np.random.default_rng(tap)
    vals = rng.integers(1, 200, g.P * nimg).astype(np.float32) / 256.0  # exact in bf16
    x0 = np.zeros((g.P * nimg, 4), np.float32)
    x0[:, 0] = vals
    x0b = O.f32_to_bf16(x0.ravel())
    _, _, prob = O.prop_loss_grad(g, nimg, tap_path(tap), x0b, None, emulate=False, want_grad=False)
    dy, dx = tap // 3 - 1, tap % 3 - 1
    expect = np.zeros(g.P * nimg)
    for z in range(nimg):
        for lvl, (tw, th, _) in enumerate(GEOM_LEVELS):
            for ty in range(th):
                for tx in range(tw):
                    y, x = ty + dy, tx + dx
                    v = 0.0
                    if 0 <= y < th and 0 <= x < tw:
                        v = vals[z * g.P + g.base[lvl] + y * tw + x]
                    expect[z * g.P + g.base[lvl] + ty * tw + tx] = v
    assert np.allclose(prob, 1.0 / (1.0 + np.exp(-expect)), rtol=0, atol=1e-7)


def test_bf16_emulation_is_close_but_not_identical():
    g, p, x0, lab = random_problem(5)
    l0, g0, p0 = O.prop_loss_grad(g, 2, p, x0, lab, emulate=False)
    l1, g1, p1 = O.prop_loss_grad(g, 2, p, x0, lab, emulate=True)
    assert np.max(np.abs(p0 - p1)) < 2e-2
    assert not np.array_equal(p0, p1)
    cos = float(g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1))
    assert cos > 0.999, cos


def test_adam_first_step_moves_by_lr_sign():
    rng = np.random.default_rng(9)
    n = 1000
    p = rng.normal(0, 1, n).astype(np.float32)
    g = rng.normal(0, 1, n)
    g[::7] = 0.0
    lr = 1e-3
    p1, m1, v1 = O.prop_adam(p, np.zeros(n, np.float32), np.zeros(n, np.float32), g, lr, 1)
    step = p - p1
    nz = g != 0
    assert np.allclose(step[nz], lr * np.sign(g[nz]), rtol=1e-4, atol=0)
    assert np.all(step[~nz] == 0)
    assert np.allclose(m1, 0.1 * g.astype(np.float32), rtol=1e-6)


def test_features_and_labels_on_a_pipeline():
    W, H = 320, 240
    import akaze_amd as A
    img = A.to_float(A.synth_image(11, W, H))
    orc = O.Pipeline(W, H)
    orc.scale_space(img)
    orc.detect()
    g = orc.prop_geom()
    x0 = O.bf16_to_f32(orc.prop_features()).reshape(g.P, 4)
    lab = orc.prop_labels()
    thr = 0.001
    for lvl in range(g.nlev):
        tw, th = g.tw[lvl], g.th[lvl]
        det = orc.level(lvl, 4)
        lt = orc.level(lvl, 0)
        for ty in (0, th - 1):
            for tx in (0, tw - 1):
                p = g.base[lvl] + ty * tw + tx
                m = det[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16].max()
                f = np.log2(1.0 + max(m, 0.0) / thr) * 0.125
                assert abs(x0[p, 1] - f) <= abs(f) * 2 ** -8 + 1e-30, (lvl, ty, tx)
                mean = lt[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16].astype(np.float64).mean()
                assert abs(x0[p, 3] - mean) <= mean * 2 ** -8, (lvl, ty, tx)
    # level 0 has no finer neighbour: its first channel repeats its own maximum
    assert np.array_equal(x0[:g.base[1], 0], x0[:g.base[1], 1])
    last = g.nlev - 1
    assert np.array_equal(x0[g.base[last]:, 2], x0[g.base[last]:, 1])
    # every kept keypoint's tile is labelled, nothing else
    cand = orc.candidates()
    kept = orc.keypoints()
    assert 0 < lab.sum() <= len(kept)
    mask = np.zeros_like(lab)
    for k in kept:
        l = int(k["class_id"])
        ratio = 1 << g.oct[l]
        # the kept keypoint lies within one pixel of its candidate pixel
        ok = cand[(cand["class_id"] == l) & (np.abs(cand["x"] / ratio - k["x"] / ratio) <= 1.0)
                  & (np.abs(cand["y"] / ratio - k["y"] / ratio) <= 1.0)]
        assert len(ok) >= 1
    kidx = np.ctypeslib.as_array(O.lib().orc_kept_candidates(orc.ctx), shape=(len(kept),))
    for v in kidx:
        c = cand[v]
        l = int(c["class_id"])
        ratio = 1 << g.oct[l]
        mask[g.base[l] + (int(c["y"]) // ratio // 16) * g.tw[l] + int(c["x"]) // ratio // 16] = 1
    assert np.array_equal(mask, lab)
