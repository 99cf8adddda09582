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
    n = ak.Feature_Detection()
    assert ak.overflow
    orc = O.Pipeline(W, H)
    orc.scale_space(img)
    orc.detect()
    cand = orc.candidates()
    assert len(cand) > cap
    ls = np.minimum(orc.level_start(), cap).astype(np.int32)
    okp, _ = O.nms(cand[:cap], ls)
    okp, odesc = O.orient_describe(orc, okp)
    kps, desc = ak.Compute_Descriptors()
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
