"""More GPU parity cases: the other diffusivities, a one-octave image, the
keypoint cap and its overflow report, and the C++ drop-in API run against
the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("diffusivity", [0, 1, 2, 3])
def test_diffusivities(gpu_lib, diffusivity):
    """PM g2 and Charbonnier use only +, *, / and sqrt: bit-exact.  PM g1 and
    Weickert call exp, whose device and host libraries differ in the last
    bits, so their levels are held to the 1e-4 relative tolerance."""
    A = gpu_lib
    W, H = 320, 240
    img = A.to_float(A.synth_image(7, W, H))
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H, diffusivity=diffusivity))
    ak.Create_Nonlinear_Scale_Space(img)
    orc = O.Pipeline(W, H, diffusivity=diffusivity)
    orc.scale_space(img)
    for lvl in range(ak.nlev):
        g, c = ak.read_level(lvl, 0), orc.level(lvl, 0)
        err = float(np.max(np.abs(g.astype(np.float64) - c)) / max(np.max(np.abs(c)), 1e-30))
        assert err <= 1e-4, (lvl, err)
        if diffusivity in (1, 3):
            assert np.array_equal(g, c), (lvl, err)


@pytest.mark.parametrize("nbins,perc", [(1, 0.7), (301, 0.7), (512, 0.7), (512, 1.0), (300, 0.05)])
def test_contrast_histogram_sizes(gpu_lib, nbins, perc):
    """Histogram sizes across the whole accepted range 1..512, including
    those whose bins reach past 300 (the histogram's second LDS array), and
    the percentile at its ends: every level's contrast factor and levels
    equal the oracle's."""
    A = gpu_lib
    W, H = 640, 480
    img = A.to_float(A.synth_image(21, W, H))
    kw = dict(kcontrast_nbins=nbins, kcontrast_percentile=perc)
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H, **kw))
    ak.Create_Nonlinear_Scale_Space(img)
    orc = O.Pipeline(W, H, **kw)
    orc.scale_space(img)
    for lvl in range(ak.nlev):
        assert ak.read_kcontrast(lvl) == orc.kcontrast(lvl), (lvl, nbins, perc)
        assert np.array_equal(ak.read_level(lvl, 0), orc.level(lvl, 0)), (lvl, nbins, perc)


@pytest.mark.parametrize("W,H", [(37, 19), (40, 20), (70, 31), (100, 47), (160, 17)])
def test_tiny_images(gpu_lib, W, H):
    """Images whose levels are narrower or shorter than a detector tile plus
    its halo: the detector's border tiles take the general path or the
    reflected one only where one reflection stays inside the image.  Levels,
    first derivatives and determinants equal the oracle's (these sizes leave
    no pixel inside the detector border, so there are no keypoints)."""
    A = gpu_lib
    img = A.to_float(A.synth_image(31, W, H))
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H))
    ak.Create_Nonlinear_Scale_Space(img)
    ak.detect()
    orc = O.Pipeline(W, H)
    orc.scale_space(img)
    orc.detect()
    for lvl in range(ak.nlev):
        for which in (0, 2, 3, 4):
            assert np.array_equal(ak.read_level(lvl, which), orc.level(lvl, which)), (lvl, which)
    kps, desc = ak.describe()
    assert kps.tobytes() == orc.keypoints().tobytes() and np.array_equal(desc, orc.describe())


def test_one_octave_image(gpu_lib):
    """96x80: the second octave (48x40) falls under the 80x40 cut, leaving
    four levels."""
    A = gpu_lib
    W, H = 96, 80
    img = A.to_float(A.synth_image(8, W, H))
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H))
    assert ak.nlev == 4
    kps, desc = ak.run(img)
    okp, odesc = O.Pipeline(W, H).run(img)
    assert kps.tobytes() == okp.tobytes() and np.array_equal(desc, odesc)


def test_keypoint_cap_reports_overflow(gpu_lib):
    """maxkeypoints = 100 on an image with 217 candidates: the first 100
    candidates in (level, y, x) order are suppressed and described, and the
    overflow is reported, never silent."""
    A = gpu_lib
    W, H, cap = 640, 480, 100
    img = A.to_float(A.synth_image(9, W, H))
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H, maxkeypoints=cap))
    ak.Create_Nonlinear_Scale_Space(img)
    n = ak.detect()
    assert ak.overflow
    orc = O.Pipeline(W, H)
    orc.scale_space(img)
    orc.detect()
    cand = orc.candidates()
    assert len(cand) > cap
    ls = np.minimum(orc.level_start(), cap).astype(np.int32)
    okp, _ = O.nms(cand[:cap], ls)
    okp, odesc = O.orient_describe(orc, okp)
    kps, desc = ak.describe()
    assert n == len(okp) == len(kps)
    assert kps.tobytes() == okp.tobytes() and np.array_equal(desc, odesc)


def test_cpp_api_runs_on_the_gpu(gpu_lib, tmp_path):
    """The reference's usage pattern through libAKAZEMI (namespace change
    only) on the GPU: keypoint count equal to the oracle's, self-matches at
    distance 0."""
    A = gpu_lib
    exe = tmp_path / "api_check"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(REPO, "include"),
                        os.path.join(REPO, "tests", "cpp", "api_check.cpp"), "-L", os.path.join(REPO, "akaze_amd"),
                        "-lakaze_amd", "-Wl,-rpath," + os.path.join(REPO, "akaze_amd"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    tag, nk, nm = r.stdout.strip().split("\n")[-1].split()
    img = A.to_float(A.synth_image(3, 640, 480))
    okp, _ = O.Pipeline(640, 480).run(img)
    assert tag == "ok" and int(nk) == len(okp) and int(nm) == len(okp)


@pytest.mark.parametrize("factor", [2.5, 3.0])
def test_large_derivative_steps_use_the_unfused_detector(gpu_lib, factor):
    """derivative_factor 2.5 / 3.0 give Hessian steps of 4..8 pixels, beyond
    the fused detector's LDS halo: the separate derivative, determinant and
    extrema kernels run, and match the oracle level by level."""
    A = gpu_lib
    W, H = 640, 480
    img = A.to_float(A.synth_image(12, W, H))
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H, derivative_factor=factor))
    ak.Create_Nonlinear_Scale_Space(img)
    n = ak.detect()
    orc = O.Pipeline(W, H, derivative_factor=factor)
    orc.scale_space(img)
    on = orc.detect()
    for lvl in range(ak.nlev):
        for which in (2, 3, 4):
            assert np.array_equal(ak.read_level(lvl, which), orc.level(lvl, which)), (lvl, which)
    cand, ls, waves = ak.read_candidates()
    assert cand.tobytes() == orc.candidates().tobytes() and waves == orc.waves()
    kps, desc = ak.describe()
    okp = orc.keypoints()
    odesc = orc.describe()
    assert n == on and kps.tobytes() == orc.keypoints().tobytes() and np.array_equal(desc, odesc)


def test_proposal_in_graph_batches_equals_eager(gpu_lib):
    """A learned-proposal batch context replayed as a HIP graph gives, image
    by image, the eager single-image result with the same network."""
    A = gpu_lib
    from akaze_amd.proposal import ProposalNet
    here = os.path.join(REPO, "akaze_amd")
    net = ProposalNet()
    net.load(os.path.join(here, "proposal_1080p.f32"))
    W, H, thr = 640, 480, 0.3
    imgs = [A.to_float(A.synth_image(40 + i, W, H)) for i in range(3)]
    ref = []
    one = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 1)
    one.set_proposal(net, thr)
    for im in imgs:
        one.upload([im])
        one.process_async(1)
        c, _ = one.counts(1)
        ref.append(one.read_results(0, int(c[0]))[:2])
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 3)
    bk.set_proposal(net, thr)
    bk.set_graph_mode(True)
    bk.upload(imgs)
    for rep in range(3):
        bk.process_async(3)
        c, _ = bk.counts(3)
        for z in range(3):
            kps, desc, n = bk.read_results(z, int(c[z]))
            assert kps.tobytes() == ref[z][0].tobytes() and np.array_equal(desc, ref[z][1]), (rep, z)


@pytest.mark.parametrize("patsize", [8, 9])
def test_mldb_pattern_sizes(gpu_lib, patsize):
    """Other descriptor pattern sizes (smaller grids and cells than the
    default 10, and for 8 lanes whose samples leave a cell and come back to
    it) give the oracle's descriptor bits."""
    A = gpu_lib
    W, H = 640, 480
    img = A.to_float(A.synth_image(23, W, H))
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H, descriptor_pattern_size=patsize))
    ak.Create_Nonlinear_Scale_Space(img)
    ak.detect()
    kps, desc = ak.describe()
    orc = O.Pipeline(W, H, descriptor_pattern_size=patsize)
    okp, odesc = orc.run(img)
    assert len(kps) == len(okp) > 100
    assert np.array_equal(kps["angle"], okp["angle"])
    assert np.array_equal(desc, odesc), int(np.sum(desc != odesc))


def test_conductance_reciprocal_equals_division(gpu_lib):
    """The conductance's 1/x (x >= 1) uses the hardware reciprocal and one FMA
    Newton step; on the device it equals IEEE 1.0f / x for every float in
    [1, +inf] (about 1.07e9 values), so the scale space stays bit-exact."""
    import ctypes
    A = gpu_lib
    m = ctypes.c_uint64(12345)
    A._check(A.lib().akz_selftest_reciprocal(ctypes.byref(m)), "akz_selftest_reciprocal")
    assert m.value == 0, m.value
