"""C-ABI library checks that need no GPU: exports, argument validation,
the C++ API header, and the synthetic input generator."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import akaze_amd as A

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for root, _, files in os.walk(os.path.join(REPO, "include")):
        for f in files:
            if f.endswith(".h"):
                txt = open(os.path.join(root, f)).read()
                names |= set(re.findall(r"^\s*(?:akz_status|const char\*)\s+(akz_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = A.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for n in sorted(names):
        assert hasattr(L, n), n
    assert names == set(A.SIGNATURES), names ^ set(A.SIGNATURES)


def test_status_strings_and_defaults():
    L = A.lib()
    assert L.akz_status_string(0) == b"ok"
    o = A.AKAZEOptions()
    assert (o.omax, o.nsublevels, o.descriptor, o.kcontrast_nbins) == (4, 4, 5, 300)
    assert abs(o.soffset - 1.6) < 1e-6 and abs(o.dthreshold - 0.001) < 1e-9


def test_create_validates_arguments_before_touching_the_device():
    L = A.lib()
    p = ctypes.c_void_p()
    bad = [dict(img_width=8, img_height=480), dict(img_width=640, img_height=480, diffusivity=7),
           dict(img_width=640, img_height=480, descriptor=3)]
    for kw in bad:
        o = A.AKAZEOptions(**kw)
        st = L.akz_create(ctypes.byref(o._c), ctypes.byref(p))
        assert st in (1, 6), (kw, st)
    assert L.akz_create(None, ctypes.byref(p)) == 1


def test_create_rejects_options_that_would_index_out_of_bounds():
    """Options whose kernels would read or write outside their buffers are
    refused with AKZ_ERR_ARG (1) before any device call: a derivative step at
    or beyond a level's width or height (reflect-101 indices of Derivate,
    cuda_akaze.cu:651-671), histogram bins outside 1..512 (ContrastPercentile
    caps them at 512, cuda_akaze.cu:619-624), a percentile outside (0, 1], and
    non-finite or non-positive scales."""
    L = A.lib()
    p = ctypes.c_void_p()
    nan, inf = float("nan"), float("inf")
    bad = [
        # 16x16: level 0's step fRound(1.6 * 6 * 2^(3/4)) = 16 reaches the width
        dict(img_width=16, img_height=16, omax=1, derivative_factor=6.0),
        dict(img_width=640, img_height=480, derivative_factor=0.0),
        dict(img_width=640, img_height=480, derivative_factor=nan),
        dict(img_width=640, img_height=480, soffset=-1.6),
        dict(img_width=640, img_height=480, soffset=inf),
        dict(img_width=640, img_height=480, kcontrast_nbins=0),
        dict(img_width=640, img_height=480, kcontrast_nbins=-3),
        dict(img_width=640, img_height=480, kcontrast_nbins=513),
        dict(img_width=640, img_height=480, kcontrast_percentile=0.0),
        dict(img_width=640, img_height=480, kcontrast_percentile=1.5),
        dict(img_width=640, img_height=480, kcontrast_percentile=nan),
        dict(img_width=640, img_height=480, dthreshold=nan),
    ]
    for kw in bad:
        o = A.AKAZEOptions(**kw)
        st = L.akz_create(ctypes.byref(o._c), ctypes.byref(p))
        assert st == 1, (kw, st)
    # the limits themselves are accepted (the call then fails only for want
    # of a device here, or succeeds on a GPU)
    ok = [dict(img_width=640, img_height=480, kcontrast_nbins=1),
          dict(img_width=640, img_height=480, kcontrast_nbins=512),
          dict(img_width=640, img_height=480, kcontrast_percentile=1.0),
          dict(img_width=16, img_height=16, omax=1, derivative_factor=5.0)]
    for kw in ok:
        o = A.AKAZEOptions(**kw)
        st = L.akz_create(ctypes.byref(o._c), ctypes.byref(p))
        assert st in (0, 7), (kw, st)
        if st == 0:
            assert L.akz_destroy(p) == 0


@pytest.mark.skipif(A.device_count() > 0, reason="checks the no-device path")
def test_no_device_is_an_error_not_a_fallback():
    o = A.AKAZEOptions(img_width=640, img_height=480)
    with pytest.raises(A.AkazeError) as e:
        A.AKAZE(o)
    assert e.value.status == 7
    with pytest.raises(A.AkazeError):
        A.Matcher()


def test_cpp_api_compiles_and_links(tmp_path):
    exe = tmp_path / "api_check"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(REPO, "include"),
                        os.path.join(REPO, "tests", "cpp", "api_check.cpp"), "-L", os.path.join(REPO, "akaze_amd"),
                        "-lakaze_amd", "-Wl,-rpath," + os.path.join(REPO, "akaze_amd"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if A.device_count() == 0:
        r = subprocess.run([str(exe)], capture_output=True, text=True)
        assert r.returncode == 10, r.stdout + r.stderr


def test_synthetic_generator_is_seeded():
    a = A.synth_image(11, 640, 480)
    b = A.synth_image(11, 640, 480)
    c = A.synth_image(12, 640, 480)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.dtype == np.uint8 and 40 < a.mean() < 220 and a.std() > 10
    H = A.synth_homography(3, 640, 480)
    assert np.all(np.abs(H[2, :2]) <= 1e-4)  # perspective terms; H[2,2] absorbs the centring
    s = np.sqrt(abs(np.linalg.det(H[:2, :2])))
    assert 0.7 <= s <= 1.3
    Hid = np.eye(3)
    assert np.array_equal(A.synth_warp(a, Hid), a)
    Hr = A.synth_homography(4, 640, 480, max_persp=0.0)
    c0 = Hr @ np.array([320.0, 240.0, 1.0])
    assert np.allclose(c0[:2] / c0[2], [320, 240])


def test_batch_ragged_and_proposal_entry_points_validate_arguments():
    """Argument checks that run before any device work."""
    L = A.lib()
    p = ctypes.c_void_p()
    o = A.AKAZEOptions(img_width=640, img_height=480)
    assert L.akz_create_batch(ctypes.byref(o._c), 0, ctypes.byref(p)) == 1
    assert L.akz_create_batch(ctypes.byref(o._c), 65, ctypes.byref(p)) == 1
    w = (ctypes.c_int * 2)(640, 640)
    h = (ctypes.c_int * 2)(480, 480)
    cap = (ctypes.c_int * 2)(1, 1)
    # duplicate size classes are rejected
    assert L.akz_ragged_create(ctypes.byref(o._c), 2, w, h, cap, ctypes.byref(p)) == 1
    assert L.akz_ragged_create(ctypes.byref(o._c), 0, w, h, cap, ctypes.byref(p)) == 1
    n = ctypes.c_int()
    assert L.akz_prop_num_params(ctypes.byref(n)) == 0 and n.value == 14513
    assert L.akz_prop_create(ctypes.c_uint64(1), None) == 1
    assert L.akz_set_proposal(None, None, ctypes.c_float(0.5)) == 1
    assert L.akz_synth_device_batch(ctypes.c_uint64(1), 0, 640, 480, None, 2560, 0) == 1
    assert L.akz_autotune(None, 1, 1) == 1
    assert L.akz_set_suppression_pool(None, 0) == 1


def test_bench_refuses_more_than_one_gpu():
    import sys
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_committed_proposal_weights_are_well_formed():
    import json
    here = os.path.join(REPO, "akaze_amd")
    w = np.fromfile(os.path.join(here, "proposal_1080p.f32"), "<f4")
    assert w.shape == (14513,) and np.all(np.isfinite(w))
    rec = json.load(open(os.path.join(here, "proposal_1080p.json")))
    assert 0.0 < rec["threshold"] < 1.0
    assert rec["heldout_kp_recall"] >= 0.98 and rec["heldout_images"] >= 2000 and rec["train_images"] >= 20000


def test_unknown_option_names_are_rejected():
    with pytest.raises(AttributeError):
        A.AKAZEOptions(max_keypoints=10)
    o = A.AKAZEOptions(maxkeypoints=10)
    assert o.maxkeypoints == 10
