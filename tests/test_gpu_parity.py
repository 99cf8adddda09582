"""GPU parity: every stage of the HIP pipeline against the CPU oracle.

The library and the oracle evaluate the same float expressions in the same
order without FMA contraction, so levels, contrast factors, candidates,
keypoints and descriptors are compared for exact equality.  The tolerances
the specification allows (1e-4 relative on levels, 99% of keypoints within
0.5 px, at most 1% MLDB bits) are asserted as well, and a failure reports
the gap against them.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

SIZES = [(640, 480, 0), (1920, 1080, 1), (900, 675, 2), (320, 240, 3), (3840, 2160, 4)]


def run_both(A, W, H, seed, fed_steps=0):
    img = A.to_float(A.synth_image(seed, W, H))
    opts = A.AKAZEOptions(img_width=W, img_height=H, fed_steps_per_launch=fed_steps)
    ak = A.AKAZE(opts)
    assert ak.Create_Nonlinear_Scale_Space(img) == 0
    orc = O.Pipeline(W, H)
    orc.scale_space(img)
    return ak, orc, img


def rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("fed_steps", [0, 1])
@pytest.mark.parametrize("W,H,seed", SIZES)
def test_scale_space_levels(gpu_lib, W, H, seed, fed_steps):
    """fed_steps 0: blocked multi-step FED kernel; 1: one step per launch."""
    A = gpu_lib
    ak, orc, _ = run_both(A, W, H, seed, fed_steps)
    assert ak.nlev == orc.nlev
    for lvl in range(ak.nlev):
        assert ak.read_kcontrast(lvl) == orc.kcontrast(lvl), lvl
        for which in (0, 1):
            g = ak.read_level(lvl, which)
            c = orc.level(lvl, which)
            assert g.shape == c.shape
            err = rel_err(g, c)
            assert err <= 1e-4, (lvl, which, err)
            assert np.array_equal(g, c), (lvl, which, err, int(np.sum(g != c)))


@pytest.mark.parametrize("W,H,seed", SIZES)
def test_detection_and_description(gpu_lib, W, H, seed):
    A = gpu_lib
    ak, orc, _ = run_both(A, W, H, seed)
    n = ak.detect()
    orc.detect()
    for lvl in range(ak.nlev):
        for which in (2, 3, 4):
            g = ak.read_level(lvl, which)
            c = orc.level(lvl, which)
            assert np.array_equal(g, c), (lvl, which, rel_err(g, c))
    cand, ls, waves = ak.read_candidates()
    ocand = orc.candidates()
    assert np.array_equal(ls, orc.level_start())
    assert len(cand) == len(ocand)
    assert cand.tobytes() == ocand.tobytes()
    assert waves == orc.waves()
    kps, desc = ak.describe()
    okp = orc.keypoints()
    odesc = orc.describe()
    okp = orc.keypoints()
    assert n == len(okp) and len(kps) == len(okp)
    for f in ("x", "y", "size", "response", "octave", "class_id", "angle"):
        assert np.array_equal(kps[f], okp[f]), f
    assert np.array_equal(desc, odesc)


def test_matcher_exact(gpu_lib):
    A = gpu_lib
    rng = np.random.default_rng(7)
    for nq, nt in [(1, 0), (1, 1), (5, 3), (300, 1000), (2048, 2048), (1000, 70000)]:
        q = rng.integers(0, 256, (nq, 61), dtype=np.uint8)
        t = rng.integers(0, 256, (nt, 61), dtype=np.uint8)
        q[:, 60] &= 0x3F
        t[:, 60] &= 0x3F
        if nt > 3:
            t[1] = t[2]          # duplicate train rows: ties broken by lower index
            q[0] = t[3]          # an exact match
        m = A.Matcher().match2nn(q, t)
        o = O.match2nn(q, t)
        assert m.tobytes() == o.tobytes(), (nq, nt)


def test_graph_mode_replays(gpu_lib):
    """The captured HIP graph gives the oracle's result on every replay,
    with eager work and host synchronisation between replays."""
    A = gpu_lib
    W, H = 640, 480
    img = A.to_float(A.synth_image(5, W, H))
    orc = O.Pipeline(W, H)
    okp, odesc = orc.run(img)
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H))
    ptr, pitch = ak.input_buffer()
    padded = np.zeros((H, pitch // 4), np.float32)
    padded[:, :W] = img
    A.lib().akz_memcpy_h2d(ptr, padded.ctypes.data, padded.nbytes)
    ak.set_graph_mode(True)
    other = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H))
    for rep in range(4):
        ak.process_device_async(ptr, pitch)
        ak.wait()
        kps, desc, n = ak.read_results(len(okp) + 16)
        assert n == len(okp), rep
        assert kps.tobytes() == okp.tobytes(), rep
        assert np.array_equal(desc, odesc), rep
        other.run(A.to_float(A.synth_image(100 + rep, W, H)))  # eager work between replays


@pytest.mark.parametrize("W,H,thr", [(1280, 720, 1e-4), (1920, 1080, 1e-6)])
def test_many_candidates(gpu_lib, W, H, thr):
    """Dense candidates: the suppression's global-memory path (> 8192
    candidates) and its sort fallback (> 65536) against the oracle."""
    A = gpu_lib
    rng = np.random.default_rng(W)
    img = rng.random((H, W)).astype(np.float32)
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H, dthreshold=thr, min_dthreshold=thr * 0.1))
    ak.Create_Nonlinear_Scale_Space(img)
    n = ak.detect()
    orc = O.Pipeline(W, H, dthreshold=thr, min_dthreshold=thr * 0.1)
    orc.scale_space(img)
    on = orc.detect()
    cand, ls, waves = ak.read_candidates()
    assert len(cand) == orc.candidates().shape[0] and len(cand) > 8192
    assert cand.tobytes() == orc.candidates().tobytes()
    assert waves == orc.waves()
    kps, desc = ak.describe()
    odesc = orc.describe()
    assert n == on and kps.tobytes() == orc.keypoints().tobytes()
    assert np.array_equal(desc, odesc)


@pytest.mark.parametrize("W,H,seed,pool", [(640, 480, 0, 0), (1920, 1080, 1, 0), (1920, 1080, 1, 8)])
def test_suppression_pool_spill(gpu_lib, W, H, seed, pool):
    """Neighbour lists of three or more that do not fit the suppression's LDS
    pool send the image down the global-memory path: same keypoints."""
    A = gpu_lib
    ak, orc, _ = run_both(A, W, H, seed)
    ak.set_suppression_pool(pool)
    n = ak.detect()
    orc.detect()
    cand, ls, waves = ak.read_candidates()
    assert waves == orc.waves()
    kps, desc = ak.describe()
    odesc = orc.describe()
    assert n == len(orc.keypoints()) and kps.tobytes() == orc.keypoints().tobytes()
    assert np.array_equal(desc, odesc)


@pytest.mark.parametrize("max_waves", [0, 1, 2, 5, 13])
def test_suppression_wave_limit(gpu_lib, max_waves):
    """A wave limit stops the suppression where the oracle's stops: wave 0
    alone, a few block waves, and a stop inside the single-wavefront tail."""
    A = gpu_lib
    W, H, seed = 1920, 1080, 1
    img = A.to_float(A.synth_image(seed, W, H))
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H, nms_max_waves=max_waves))
    ak.Create_Nonlinear_Scale_Space(img)
    n = ak.detect()
    orc = O.Pipeline(W, H, nms_max_waves=max_waves)
    orc.scale_space(img)
    on = orc.detect()
    _, _, waves = ak.read_candidates()
    assert waves == orc.waves() == min(max_waves, waves)
    kps, desc = ak.describe()
    odesc = orc.describe()
    assert n == on and kps.tobytes() == orc.keypoints().tobytes()
    assert np.array_equal(desc, odesc)
