"""Pins the CPU oracle stage by stage (no GPU).

Each stage is checked against something other than its own formula: the
reference's FED table and compiled fed.cpp, closed forms (ramps, quadratics,
constant images), independent numpy/float64 implementations, and hand-worked
suppression cases.
"""
import math
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- FED schedule
def load_golden():
    rows = {}
    with open(os.path.join(HERE, "golden", "fed_tau_4x4.txt")) as fh:
        for line in fh:
            head, taus = line.split(":")
            lvl = int(head.split()[1])
            n = int(head.split("n=")[1])
            vals = np.array([np.float32(float(v)) for v in taus.split()], np.float32)
            assert len(vals) == n
            rows[lvl] = vals
    return rows


def test_fed_schedule_matches_reference_table():
    """Appendix-A golden table, produced by the reference's own fed.cpp."""
    g = O.geometry(640, 480)
    gold = load_golden()
    assert sorted(gold) == list(range(1, 16))
    total = 0
    for lvl, vals in gold.items():
        mine = np.array(g.tau[lvl][:g.nsteps[lvl]], np.float32)
        assert np.array_equal(mine, vals), lvl
        total += len(vals)
    assert total == 166


@pytest.mark.skipif(not os.path.exists(O.FEDREF_SO), reason="reference fed.cpp not built (no reference sources)")
def test_fed_against_compiled_reference():
    import ctypes
    L = ctypes.CDLL(O.FEDREF_SO)
    f = L.fedref_tau_by_process_time
    f.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                  ctypes.c_int]
    rng = np.random.default_rng(0)
    buf = (ctypes.c_float * 256)()
    for T in np.concatenate([rng.uniform(0.01, 200.0, 300), [0.53, 2.0, 67.86]]).astype(np.float32):
        for tmax in (0.25, 0.5):
            n = f(float(T), 1, tmax, 1, buf, 256)
            mine = O.fed_tau(float(T), 1, tmax, True)
            assert n == len(mine), T
            if n == 1:
                continue  # fed.cpp:108-114 reads tauh[-1] for n == 1; the oracle uses the one step
            assert np.array_equal(np.array(buf[:n], np.float32), mine), T


def test_fed_properties():
    """Step sizes sum to the stopping time and the cycle is a permutation."""
    for T in (0.5, 3.0, 47.99, 67.86):
        tau = O.fed_tau(T)
        unr = O.fed_tau(T, reordering=False)
        assert abs(float(np.sum(tau.astype(np.float64))) - T) < 1e-4 * T
        assert sorted(tau.tolist()) == sorted(unr.tolist())
        n = len(tau)
        assert n == math.ceil(math.sqrt(3 * T / 0.25 + 0.25) - 0.5 - 1e-8)


# ---------------------------------------------------------------- geometry
def test_level_geometry():
    g = O.geometry(1920, 1080)
    assert g.nlev == 16
    assert [(g.w[i], g.h[i]) for i in (0, 4, 8, 12)] == [(1920, 1080), (960, 540), (480, 270), (240, 135)]
    assert [g.deriv_step[i] for i in range(16)] == [2, 3, 3, 4] * 4
    assert [round(g.border[i], 3) for i in range(4)] == [28.284, 42.426, 42.426, 56.569]
    g = O.geometry(320, 240)
    assert g.nlev == 12  # 40x30 < 80x40 ends the pyramid (AKAZE.cpp:157-160)
    g = O.geometry(900, 675)
    assert [(g.w[i], g.h[i]) for i in (0, 4, 8, 12)] == [(900, 675), (450, 337), (225, 168), (112, 84)]


# ---------------------------------------------------------------- Gaussian
def test_gauss_taps_closed_form():
    for var, R in ((1.0, 2), (2.56, 4)):
        k = O.gauss_taps(var, R).astype(np.float64)
        ref = np.exp(-np.arange(-R, R + 1) ** 2 / (2 * var))
        ref /= ref.sum()
        assert np.allclose(k, ref, rtol=1e-6, atol=0)
        assert np.array_equal(k, k[::-1])


def test_lowpass_matches_2d_convolution():
    rng = np.random.default_rng(1)
    img = rng.random((37, 53)).astype(np.float32)
    for var, ksize in ((1.0, 5), (2.56, 9)):
        R = {5: 2, 9: 4}[ksize]
        k = O.gauss_taps(var, R).astype(np.float64)
        pad = np.pad(img.astype(np.float64), R, mode="edge")
        K = np.outer(k, k)
        ref = np.zeros(img.shape)
        for dy in range(2 * R + 1):
            for dx in range(2 * R + 1):
                ref += K[dy, dx] * pad[dy:dy + img.shape[0], dx:dx + img.shape[1]]
        out = O.lowpass(img, var, ksize)
        assert np.max(np.abs(out - ref)) < 2e-6
    c = np.full((20, 30), 0.37, np.float32)
    assert np.max(np.abs(O.lowpass(c, 2.56, 9) - 0.37)) < 1e-6


# ---------------------------------------------------------------- contrast
def numpy_contrast(img, perc=0.7, nbins=300):
    k = O.gauss_taps(1.0, 2).astype(np.float64)
    pad = np.pad(img.astype(np.float64), 2, mode="edge")
    tmp = sum(k[i] * pad[:, i:i + img.shape[1]] for i in range(5))
    blur = sum(k[i] * tmp[i:i + img.shape[0], :] for i in range(5))
    b = blur
    dx = 3 * (b[:-2, :-2] - b[:-2, 2:] + b[2:, :-2] - b[2:, 2:]) + 10 * (b[1:-1, :-2] - b[1:-1, 2:])
    dy = 3 * (b[:-2, :-2] + b[:-2, 2:] - b[2:, :-2] - b[2:, 2:]) + 10 * (b[:-2, 1:-1] - b[2:, 1:-1])
    g = np.sqrt(dx * dx + dy * dy)
    mx = g.max()
    idx = np.minimum((nbins * g / mx).astype(np.int64), nbins - 1)
    hist = np.bincount(idx.ravel(), minlength=nbins)
    thr = int(g.size * perc)
    cs = np.cumsum(hist)
    kbin = int(np.searchsorted(cs, thr) + 1)
    return mx * kbin / nbins, hist, mx


def test_contrast_against_numpy():
    rng = np.random.default_rng(2)
    for shape in ((24, 31), (60, 80)):
        img = rng.random(shape).astype(np.float32)
        k, hist, mx = O.contrast(img)
        kr, hr, mr = numpy_contrast(img)
        assert abs(float(mx) - mr) < 1e-5 * mr
        assert np.sum(np.abs(hist - hr)) <= 2  # float64 vs float32 may move a bin edge
        assert abs(float(k) - kr) < 1e-2 * kr
        assert hist.sum() == (shape[0] - 2) * (shape[1] - 2)  # zeros included (HistContrast)


def test_contrast_ramp_closed_form():
    """Interior unnormalised Scharr magnitude of a ramp a*x is 32a."""
    a = 0.01
    img = (a * np.arange(64, dtype=np.float32))[None, :].repeat(48, 0).astype(np.float32)
    k, hist, mx = O.contrast(img)
    assert abs(float(mx) - 32 * a) < 1e-5
    # more than 70% of the interior sits at the maximum: the percentile is the top bin
    assert abs(float(k) - 32 * a) < 1e-5


# ---------------------------------------------------------------- conductance
FLOW_FORMULAS = {
    # Flow, cuda_akaze.cu:261-268: PM_G1, PM_G2, WEICKERT, CHARBONNIER of
    # d = |grad|^2 / k^2
    0: lambda d: math.exp(-d),
    1: lambda d: 1.0 / (1.0 + d),
    2: lambda d: 1.0 - math.exp(-3.315 / d ** 4),
    3: lambda d: 1.0 / math.sqrt(1.0 + d),
}


@pytest.mark.parametrize("kc", [0.05, 0.1])
@pytest.mark.parametrize("dtype", [0, 1, 2, 3])
def test_flow_on_ramp(dtype, kc):
    """On L = a x + b y the unnormalised Scharr gives (32 a, 32 b) inside and
    (0, 32 b) on the reflect-101 edge columns, so g is the closed form of the
    diffusivity at d = |grad|^2 / k^2."""
    a, b = 0.003, -0.002
    y, x = np.mgrid[0:30, 0:40].astype(np.float32)
    L = (a * x + b * y).astype(np.float32)
    g = O.flow(L, kc, dtype)
    f = FLOW_FORMULAS[dtype]
    d_in = ((32 * a) ** 2 + (32 * b) ** 2) / kc ** 2
    d_edge = (32 * b) ** 2 / kc ** 2
    assert np.allclose(g[1:-1, 1:-1], f(d_in), rtol=2e-5, atol=1e-7)
    assert np.allclose(g[1:-1, 0], f(d_edge), rtol=2e-5, atol=1e-7)
    # the four formulas differ from each other far beyond the tolerance here
    others = [FLOW_FORMULAS[t](d_in) for t in FLOW_FORMULAS if t != dtype]
    assert all(abs(v - f(d_in)) > 1e-3 * abs(f(d_in)) for v in others)


# ---------------------------------------------------------------- FED step
def numpy_fed_step(L, g, tau):
    Lp = np.pad(L.astype(np.float64), 1, mode="edge")
    gp = np.pad(g.astype(np.float64), 1, mode="edge")
    c = Lp[1:-1, 1:-1]
    gc = gp[1:-1, 1:-1]
    flux = ((gc + gp[1:-1, 2:]) * (Lp[1:-1, 2:] - c) - (gc + gp[1:-1, :-2]) * (c - Lp[1:-1, :-2]) +
            (gc + gp[2:, 1:-1]) * (Lp[2:, 1:-1] - c) - (gc + gp[:-2, 1:-1]) * (c - Lp[:-2, 1:-1]))
    return c + 0.5 * tau * flux


def test_fed_step_invariants():
    rng = np.random.default_rng(3)
    L = rng.random((33, 47)).astype(np.float32)
    g = rng.random((33, 47)).astype(np.float32)
    out = O.fed_step(L, g, 0.2)
    assert np.max(np.abs(out - numpy_fed_step(L, g, 0.2))) < 1e-6
    # zero flux across the border: the sum is conserved
    assert abs(out.astype(np.float64).sum() - L.astype(np.float64).sum()) < 1e-4
    # a constant image is a fixed point, exactly
    c = np.full((10, 12), 0.25, np.float32)
    assert np.array_equal(O.fed_step(c, g[:10, :12].copy(), 5.0), c)
    # transposition symmetry
    outT = O.fed_step(np.ascontiguousarray(L.T), np.ascontiguousarray(g.T), 0.2)
    assert np.max(np.abs(outT - out.T)) < 1e-6


# ---------------------------------------------------------------- downsampling
def area_resize(img, ow, oh):
    ih, iw = img.shape

    def weights(n_in, n_out):
        W = np.zeros((n_out, n_in))
        for o in range(n_out):
            lo, hi = o * n_in / n_out, (o + 1) * n_in / n_out
            for i in range(int(math.floor(lo)), min(int(math.ceil(hi)), n_in)):
                W[o, i] = max(0.0, min(hi, i + 1) - max(lo, i))
        return W / W.sum(1, keepdims=True)
    return weights(ih, oh) @ img.astype(np.float64) @ weights(iw, ow).T


def test_halfsample_exact_and_area():
    y, x = np.mgrid[0:20, 0:30].astype(np.float32)
    ramp = (0.5 * x + 0.25 * y).astype(np.float32)
    hs = O.halfsample(ramp, 15, 10)
    yy, xx = np.mgrid[0:10, 0:15]
    assert np.array_equal(hs, (0.5 * (2 * xx + 0.5) + 0.25 * (2 * yy + 0.5)).astype(np.float32))
    rng = np.random.default_rng(4)
    for iw, ih in ((31, 20), (30, 21), (45, 37)):
        img = rng.random((ih, iw)).astype(np.float32)
        ow, oh = iw // 2, ih // 2
        out = O.halfsample(img, ow, oh)
        ref = area_resize(img, ow, oh)
        rows = [r for r in range(oh) if not (iw % 2 and ih % 2 and r % 16 == 15)]
        assert np.max(np.abs(out[rows] - ref[rows])) < 1e-5, (iw, ih)
    # odd width and height: rows y % 16 == 15 read column 2x+1 twice (cuda_akaze.cu:413)
    img = rng.random((37, 45)).astype(np.float32)
    out = O.halfsample(img, 22, 18)
    r = 15
    ow, oh, iw, ih = 22, 18, 45, 37
    xs = np.arange(ow)
    f0, f2 = (ow - xs).astype(np.float64), (1 + xs).astype(np.float64)

    def hrow(row, quirk=False):
        p = img[row].astype(np.float64)
        return f0 * p[2 * xs] + ow * p[2 * xs + 1] + f2 * p[2 * xs + (1 if quirk else 2)]
    expect = ((oh - r) * hrow(2 * r) + oh * hrow(2 * r + 1) + (1 + r) * hrow(2 * r + 2, True)) / (iw * ih)
    assert np.max(np.abs(out[r] - expect)) < 1e-5


# ---------------------------------------------------------------- Hessian
@pytest.mark.parametrize("s", [2, 3, 4])
def test_hessian_of_quadratics(s):
    y, x = np.mgrid[0:40, 0:50].astype(np.float64)
    inner = (slice(2 * s, -2 * s), slice(2 * s, -2 * s))
    f = ((x - 25) ** 2 + (y - 20) ** 2) * 1e-3
    lx, ly, det = O.hessian(f.astype(np.float32), s)
    assert np.allclose(lx[inner], (2 * s * (x - 25) * 1e-3)[inner], atol=1e-5)
    assert np.allclose(det[inner], 4 * s ** 4 * 1e-6, rtol=1e-3)
    f = (x - 25) * (y - 20) * 1e-3
    _, _, det = O.hessian(f.astype(np.float32), s)
    assert np.allclose(det[inner], -(s ** 4) * 1e-6, rtol=1e-3)


# ---------------------------------------------------------------- extrema
def test_extrema_planted_peak_and_subpixel():
    y, x = np.mgrid[0:100, 0:120].astype(np.float64)
    x0, y0, ox, oy = 60, 45, 0.3, -0.2
    det = (0.02 - 1e-3 * ((x - x0 - ox) ** 2 + (y - y0 - oy) ** 2)).astype(np.float32)
    c = O.find_extrema(det, border=28.28, thr=0.001, level=2, octave=0, size=3.0)
    assert len(c) == 1
    assert (c[0]["x"], c[0]["y"]) == (x0, y0)
    octf = np.frombuffer(np.int32(c[0]["octave"]).tobytes(), np.float32)[0]
    assert abs(abs(octf) - ox) < 1e-4 and np.sign(octf) == 1
    assert abs(c[0]["angle"] - oy) < 1e-4
    assert c[0]["size"] == 6.0 and c[0]["class_id"] == 2
    # a peak inside the border is rejected; below threshold is rejected
    assert len(O.find_extrema(det, border=60.0, thr=0.001, level=0, octave=0, size=3.0)) == 0
    assert len(O.find_extrema(det, border=10.0, thr=0.05, level=0, octave=0, size=3.0)) == 0


def test_extrema_weak_fit_gets_negative_size():
    det = np.zeros((80, 80), np.float32)
    # a strict maximum whose quadratic fit (strong cross term) lands 2.5 px away
    det[39:42, 39:42] = np.array([[0.98, 0.90, 0.70], [0.90, 1.00, 0.95], [0.70, 0.95, 0.98]], np.float32)
    c = O.find_extrema(det, border=20.0, thr=0.001, level=0, octave=1, size=4.0)
    assert len(c) == 1 and c[0]["size"] < 0 and c[0]["x"] == 80


# ---------------------------------------------------------------- suppression
def cand(x, y, resp, level, size, octave=0):
    c = np.zeros(1, O.KP_DTYPE)
    c["x"], c["y"], c["response"], c["class_id"], c["size"] = x, y, resp, level, size
    c["octave"] = np.frombuffer(np.float32(octave).tobytes(), np.int32)[0]
    return c


def test_nms_hand_cases():
    # (a) weaker later neighbour is dropped
    cs = np.concatenate([cand(50, 50, 1.0, 0, 10), cand(52, 51, 0.5, 0, 10)])
    out, waves = O.nms(cs, [0, 2])
    assert len(out) == 1 and (out[0]["x"], out[0]["y"]) == (50, 50)
    # (b) stronger later neighbour wins, and the emitted point is the one it
    # replaced (sortFiltered_kernel emits kpts[memberarray[i]])
    cs = np.concatenate([cand(50, 50, 0.5, 0, 10), cand(52, 51, 1.0, 0, 10)])
    out, waves = O.nms(cs, [0, 2])
    assert len(out) == 1 and (out[0]["x"], out[0]["y"], out[0]["response"]) == (50, 50, 0.5)
    assert waves == 2
    # (c) two contenders for the same kept point: the lower index wins, the
    # other retries and, seeing no kept neighbour, is added
    cs = np.concatenate([cand(50, 50, 0.2, 0, 10), cand(46, 52, 1.0, 0, 10), cand(54, 52, 0.9, 0, 10)])
    out, waves = O.nms(cs, [0, 3])
    assert [(o["x"], o["y"]) for o in out] == [(50, 50), (54, 52)]
    assert waves == 3
    # (d) a weak candidate has no neighbours and is never emitted
    cs = np.concatenate([cand(50, 50, 1.0, 0, -10), cand(70, 50, 1.0, 0, 10)])
    out, _ = O.nms(cs, [0, 2])
    assert [(o["x"], o["y"]) for o in out] == [(70, 50)]
    # (e) previous-level neighbour across levels
    cs = np.concatenate([cand(100, 100, 0.3, 0, 8), cand(102, 100, 0.6, 1, 12)])
    out, _ = O.nms(cs, [0, 1, 2])
    assert len(out) == 1 and out[0]["response"] == np.float32(0.3)


def test_nms_properties_random():
    rng = np.random.default_rng(5)
    for _ in range(20):
        n0, n1 = rng.integers(1, 60, 2)
        c0 = np.concatenate([cand(int(x), int(y), float(r), 0, 6.0) for x, y, r in
                             zip(rng.integers(0, 40, n0), rng.integers(0, 40, n0), rng.random(n0))])
        c1 = np.concatenate([cand(int(x), int(y), float(r), 1, 9.0) for x, y, r in
                             zip(rng.integers(0, 40, n1), rng.integers(0, 40, n1), rng.random(n1))])
        order0 = np.lexsort((c0["x"], c0["y"]))
        order1 = np.lexsort((c1["x"], c1["y"]))
        cs = np.concatenate([c0[order0], c1[order1]])
        out, waves = O.nms(cs, [0, n0, n0 + n1])
        pos = {(float(c["x"]), float(c["y"]), int(c["class_id"])) for c in cs}
        assert all((float(o["x"]), float(o["y"]), int(o["class_id"])) in pos for o in out)
        assert len(out) <= len(cs) and waves <= len(cs) + 1


# ---------------------------------------------------------------- orientation / MLDB
@pytest.mark.parametrize("theta", [0.1, 1.0, 2.5, 3.5, 5.9])
def test_orientation_of_uniform_gradient(theta):
    h, w = 80, 80
    Lx = np.full((h, w), math.cos(theta), np.float32)
    Ly = np.full((h, w), math.sin(theta), np.float32)
    k = np.zeros(1, O.KP_DTYPE)
    k["x"], k["y"], k["size"], k["octave"], k["class_id"] = 40, 40, 6.0, 0, 0
    O.lib().orc_orientation(k.ctypes.data, O.fptr(Lx), O.fptr(Ly), w, h)
    assert abs(float(k[0]["angle"]) - theta) < 1e-5


def test_mldb_bits_of_ramps():
    h, w = 100, 100
    zeros = np.zeros((h, w), np.float32)
    k = np.zeros(1, O.KP_DTYPE)
    k["x"], k["y"], k["size"], k["angle"], k["octave"], k["class_id"] = 50, 50, 4.8, 0.0, 0, 0
    desc = np.zeros(61, np.uint8)
    vals = np.zeros(87, np.float32)
    flat = np.full((h, w), 7.0, np.float32)
    O.lib().orc_mldb(k.ctypes.data, O.fptr(flat), O.fptr(zeros), O.fptr(zeros), w, h, 10,
                     desc.ctypes.data_as(O.c_u8_p), O.fptr(vals))
    assert not desc.any()
    # decreasing integer ramp along image x: grid row index k moves image x at
    # angle 0, so cells further down the grid are darker; sums stay exact
    ramp = (1000.0 - np.arange(w, dtype=np.float32))[None, :].repeat(h, 0).astype(np.float32)
    O.lib().orc_mldb(k.ctypes.data, O.fptr(ramp), O.fptr(zeros), O.fptr(zeros), w, h, 10,
                     desc.ctypes.data_as(O.c_u8_p), O.fptr(vals))
    bits = np.unpackbits(desc, bitorder="little")[:486]
    i1, i2 = O.compare_tables()
    rows = {}
    for c in range(4):
        rows[c] = c // 2
    for c in range(9):
        rows[4 + c] = c // 3
    for c in range(16):
        rows[13 + c] = c // 4
    expect = np.array([1 if (i1[b] % 3 == 0 and rows[i1[b] // 3] < rows[i2[b] // 3]) else 0 for b in range(486)])
    assert np.array_equal(bits, expect)
    assert not np.unpackbits(desc, bitorder="little")[486:].any()


def test_compare_tables():
    i1, i2 = O.compare_tables()
    pairs = list(zip(i1[:486].tolist(), i2[:486].tolist()))
    assert len(set(pairs)) == 486
    assert all(a < b and a % 3 == b % 3 for a, b in pairs)
    assert sum(1 for a, b in pairs if b < 12) == 18 and sum(1 for a, b in pairs if 12 <= a < 39 and b < 39) == 108


# ---------------------------------------------------------------- matcher
def test_matcher_brute_force():
    rng = np.random.default_rng(6)
    q = rng.integers(0, 256, (40, 61), dtype=np.uint8)
    t = rng.integers(0, 256, (90, 61), dtype=np.uint8)
    t[10] = t[20]
    q[3] = t[10]
    out = O.match2nn(q, t)
    D = np.unpackbits(q[:, None, :] ^ t[None, :, :], axis=2).sum(2)
    for i in range(len(q)):
        order = sorted(range(len(t)), key=lambda j: (D[i, j], j))
        assert (out[i, 0]["trainIdx"], out[i, 0]["distance"]) == (order[0], D[i, order[0]])
        assert (out[i, 1]["trainIdx"], out[i, 1]["distance"]) == (order[1], D[i, order[1]])
    assert (out[3, 0]["trainIdx"], out[3, 1]["trainIdx"], out[3, 1]["distance"]) == (10, 20, 0)
    one = O.match2nn(q[:2], t[:1])  # fewer than two train rows: (0, 512) sentinel
    assert one[0, 1]["trainIdx"] == 0 and one[0, 1]["distance"] == 512


# ---------------------------------------------------------------- math
def test_atan2_and_sincos_accuracy():
    rng = np.random.default_rng(8)
    for y, x in zip(rng.standard_normal(2000), rng.standard_normal(2000)):
        assert abs(float(O.atan2f(float(y), float(x))) - math.atan2(np.float32(y), np.float32(x))) < 5e-7
    for a in np.linspace(0, 2 * math.pi, 2000, dtype=np.float32):
        s, c = O.sincosf(float(a))
        assert abs(float(s) - math.sin(a)) < 2e-7 and abs(float(c) - math.cos(a)) < 2e-7
    assert O.atan2f(0.0, -1.0) == np.float32(math.pi) and O.atan2f(-0.0, -1.0) == -np.float32(math.pi)
    assert O.atan2f(0.0, 0.0) == 0.0


def test_pipeline_smoke_small():
    import akaze_amd as A
    img = A.to_float(A.synth_image(0, 320, 240))
    P = O.Pipeline(320, 240)
    kps, desc = P.run(img)
    assert P.nlev == 12 and len(kps) > 10 and desc.shape == (len(kps), 61)
    assert np.all((kps["angle"] >= 0) & (kps["angle"] < 2 * math.pi + 1e-6))
    assert not (desc[:, 60] & 0xC0).any()
