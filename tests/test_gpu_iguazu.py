"""The reference's own fixtures, datasets/iguazu/img1..6.pgm with H1toNp
(900x675, identity homographies), through the GPU path:

  * every image bit-equal to the oracle (keypoints, descriptors);
  * python/test.py's call sequence through akaze_amd.libakaze_pybindings
    (no cv2: the PGM reader of the library), checked against the oracle;
  * akaze_match's evaluation (src/akaze_match.cpp:106-189): 2-NN of image N
    against image 1, NNDR 0.8, inliers at 2.5 px under H1toNp and under the
    library's RANSAC; the match pairs equal the oracle's, and the inlier
    ratios are reported;
  * the command-line programs akaze_features and akaze_match.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
IGUAZU = os.path.join(HERE, "golden", "iguazu")
REPO = os.path.dirname(HERE)
BIN = os.path.join(REPO, "akaze_amd", "bin")


def load(A, k):
    return A.read_pgm(os.path.join(IGUAZU, "img%d.pgm" % k))


def oracle_run(img_f):
    P = O.Pipeline(img_f.shape[1], img_f.shape[0])
    return P.run(img_f)


@pytest.fixture(scope="module")
def iguazu(gpu_lib):
    A = gpu_lib
    out = {}
    for k in range(1, 7):
        img = A.to_float(load(A, k))
        ak = A.AKAZE(A.AKAZEOptions(img_width=img.shape[1], img_height=img.shape[0]))
        kps, desc = ak.run(img)
        okp, odesc = oracle_run(img)
        out[k] = (img, kps, desc, okp, odesc)
    return out


@pytest.mark.parametrize("k", range(1, 7))
def test_iguazu_images_match_oracle(iguazu, k):
    img, kps, desc, okp, odesc = iguazu[k]
    print("iguazu img%d: %d keypoints" % (k, len(kps)), flush=True)
    assert len(kps) == len(okp) and len(kps) > 100
    assert kps.tobytes() == okp.tobytes()
    assert np.array_equal(desc, odesc)


def test_python_test_py_sequence(gpu_lib):
    """python/test.py with akaze_amd.libakaze_pybindings in place of the
    Boost module and the library's PGM reader in place of cv2.imread."""
    A = gpu_lib
    import akaze_amd.libakaze_pybindings as akaze
    img1, img2 = load(A, 1), load(A, 2)
    img1_32 = np.float32(img1) * (1. / 255.)
    img2_32 = np.float32(img2) * (1. / 255.)
    options = akaze.AKAZEOptions()
    height, width = np.shape(img1)
    options.setWidth(width)
    options.setHeight(height)
    evolution1 = akaze.AKAZE(options)
    height, width = np.shape(img2)
    options.setWidth(width)
    options.setHeight(height)
    evolution2 = akaze.AKAZE(options)
    assert evolution1.Create_Nonlinear_Scale_Space(img1_32) == 0
    desc1, kpts1 = evolution1.Compute_Descriptors()
    assert evolution2.Create_Nonlinear_Scale_Space(img2_32) == 0
    desc2, kpts2 = evolution2.Compute_Descriptors()
    matcher = akaze.Matcher()
    dmatch = matcher.BFMatch(desc1, desc2)
    # shapes and types of the reference module
    assert desc1.dtype == np.uint8 and desc1.shape == (len(kpts1), 61)
    assert kpts1.dtype == np.float32 and kpts1.shape[1] == 7
    assert dmatch.dtype == np.float32 and dmatch.shape == (len(desc1), 8)
    # against the oracle on the same float input
    for img_32, desc, kpts in ((img1_32, desc1, kpts1), (img2_32, desc2, kpts2)):
        okp, odesc = oracle_run(np.ascontiguousarray(img_32, np.float32))
        assert np.array_equal(desc, odesc)
        assert np.array_equal(kpts, A.kpts_to_mat(okp))
    om = O.match2nn(desc1, desc2)
    assert np.array_equal(dmatch[:, 1], om[:, 0]["trainIdx"].astype(np.float32))
    assert np.array_equal(dmatch[:, 3], om[:, 0]["distance"])
    assert np.array_equal(dmatch[:, 5], om[:, 1]["trainIdx"].astype(np.float32))
    assert np.array_equal(dmatch[:, 7], om[:, 1]["distance"])
    assert np.array_equal(dmatch[:, 0], np.arange(len(desc1), dtype=np.float32))
    assert not dmatch[:, [2, 6]].any()
    # Feature_Detection gives the same keypoints before description (angle 0)
    fd = evolution1.Feature_Detection()
    assert np.array_equal(fd[:, [0, 1, 2, 4, 5, 6]], kpts1[:, [0, 1, 2, 4, 5, 6]])


def evaluate(A, kps1, desc1, kpsN, descN, H, matches=None):
    """akaze_match.cpp:134-175: query image N against image 1, NNDR, inliers."""
    m = A.Matcher().match2nn(descN, desc1) if matches is None else matches
    pairs = A.matches2points_nndr(kpsN, kps1, m, 0.8)
    mask_h = A.inliers_homography(pairs, H, 2.5)
    _, mask_r = A.inliers_ransac(pairs, 2.5)
    return m, pairs, mask_h, mask_r


@pytest.mark.parametrize("k", range(2, 7))
def test_iguazu_matching_quality(gpu_lib, iguazu, k):
    A = gpu_lib
    _, kps1, desc1, okp1, odesc1 = iguazu[1]
    _, kpsN, descN, okpN, odescN = iguazu[k]
    H = A.read_homography(os.path.join(IGUAZU, "H1to%dp" % k))
    m, pairs, mask_h, mask_r = evaluate(A, kps1, desc1, kpsN, descN, H)
    om = O.match2nn(odescN, odesc1)
    assert np.array_equal(m["trainIdx"], om["trainIdx"]) and np.array_equal(m["distance"], om["distance"])
    _, opairs, omask_h, _ = evaluate(A, okp1, odesc1, okpN, odescN, H, om)
    assert np.array_equal(pairs, opairs) and np.array_equal(mask_h, omask_h)
    ratio_h = 100.0 * mask_h.sum() / max(len(pairs), 1)
    ratio_r = 100.0 * mask_r.sum() / max(len(pairs), 1)
    print("iguazu 1->%d: %d/%d keypoints, %d NNDR matches, inliers %d (%.1f%%) under H1to%dp, %d (%.1f%%) by RANSAC"
          % (k, len(kps1), len(kpsN), len(pairs), mask_h.sum(), ratio_h, k, mask_r.sum(), ratio_r), flush=True)
    # the set varies noise only: most NNDR matches are true correspondences
    assert len(pairs) > 50 and ratio_h > 50.0
    # RANSAC on identity data recovers at least the ground-truth consensus
    assert mask_r.sum() >= 0.95 * mask_h.sum()


def test_command_line_programs(gpu_lib, iguazu, tmp_path):
    A = gpu_lib
    img1p = os.path.join(IGUAZU, "img1.pgm")
    out = tmp_path / "kp.txt"
    r = subprocess.run([os.path.join(BIN, "akaze_features"), img1p, "--output", str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _, kps, desc, _, _ = iguazu[1]
    assert "Number of points: %d" % len(kps) in r.stdout
    lines = out.read_text().splitlines()
    assert lines[0] == "61" and int(lines[1]) == len(kps)
    rows = np.array([ln.split() for ln in lines[2:]], dtype=np.float64)
    assert np.allclose(rows[:, 0], kps["x"], rtol=1e-5) and np.allclose(rows[:, 1], kps["y"], rtol=1e-5)
    assert np.array_equal(rows[:, 5:].astype(np.uint8), desc)
    # akaze_match with the ground-truth homography and without it (RANSAC)
    img2p = os.path.join(IGUAZU, "img2.pgm")
    _, kps2, desc2, _, _ = iguazu[2]
    H = A.read_homography(os.path.join(IGUAZU, "H1to2p"))
    m, pairs, mask_h, mask_r = evaluate(A, kps, desc, kps2, desc2, H)
    for hargs, mask in (([os.path.join(IGUAZU, "H1to2p")], mask_h), ([], mask_r)):
        r = subprocess.run([os.path.join(BIN, "akaze_match"), img1p, img2p] + hargs, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        stats = dict(re.findall(r"^(Number of [A-Za-z0-9 ]+|Inliers Ratio): ([-0-9.e+]+)$", r.stdout, re.M))
        assert int(stats["Number of Keypoints Image 1"]) == len(kps)
        assert int(stats["Number of Keypoints Image 2"]) == len(kps2)
        assert int(stats["Number of Matches"]) == len(pairs)
        assert int(stats["Number of Inliers"]) == int(mask.sum())
