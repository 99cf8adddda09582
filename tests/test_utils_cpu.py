"""The evaluation and I/O helpers of the reference programs
(src/lib/utils.cpp) as the library provides them, on the CPU: PGM input,
the Mikolajczyk keypoint file, homography files, NNDR filtering, homography
inliers and the library's own RANSAC.  Each is checked against an
independent reading of the format or a closed form."""
import math
import os

import numpy as np
import pytest

import akaze_amd as A

HERE = os.path.dirname(os.path.abspath(__file__))
IGUAZU = os.path.join(HERE, "golden", "iguazu")


def numpy_pgm(path):
    """Independent P5 parse: header tokens, then the raster after one whitespace byte."""
    raw = open(path, "rb").read()
    toks, i = [], 0
    while len(toks) < 4:
        while raw[i:i + 1].isspace():
            i += 1
        if raw[i:i + 1] == b"#":
            while raw[i:i + 1] != b"\n":
                i += 1
            continue
        j = i
        while not raw[j:j + 1].isspace():
            j += 1
        toks.append(raw[i:j])
        i = j
    w, h = int(toks[1]), int(toks[2])
    return np.frombuffer(raw[i + 1:i + 1 + w * h], np.uint8).reshape(h, w)


@pytest.mark.parametrize("k", range(1, 7))
def test_read_iguazu_pgm(k):
    p = os.path.join(IGUAZU, "img%d.pgm" % k)
    img = A.read_pgm(p)
    assert img.shape == (675, 900)
    assert np.array_equal(img, numpy_pgm(p))


def test_read_pgm_variants(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, (7, 11), dtype=np.uint8)
    p5 = tmp_path / "a.pgm"
    p5.write_bytes(b"P5\n# a comment\n11 7\n# another\n255\n" + a.tobytes())
    assert np.array_equal(A.read_pgm(p5), a)
    p2 = tmp_path / "b.pgm"
    p2.write_text("P2\n11 7\n255\n" + "\n".join(" ".join(str(v) for v in row) for row in a) + "\n")
    assert np.array_equal(A.read_pgm(p2), a)
    # 16-bit samples: the high byte
    b16 = rng.integers(0, 65536, (7, 11)).astype(">u2")
    p16 = tmp_path / "c.pgm"
    p16.write_bytes(b"P5\n11 7\n65535\n" + b16.tobytes())
    assert np.array_equal(A.read_pgm(p16), (b16.astype(np.uint32) >> 8).astype(np.uint8))
    # truncated raster, wrong magic, missing file
    bad = tmp_path / "d.pgm"
    bad.write_bytes(b"P5\n11 7\n255\n" + a.tobytes()[:20])
    for q in (bad, tmp_path / "missing.pgm"):
        with pytest.raises(A.AkazeError):
            A.read_pgm(q)
    ppm = tmp_path / "e.ppm"
    ppm.write_bytes(b"P6\n1 1\n255\n\x00\x00\x00")
    with pytest.raises(A.AkazeError) as e:
        A.read_pgm(ppm)
    assert e.value.status == 6


def test_save_keypoints_format(tmp_path):
    kps = np.zeros(3, A.KP_DTYPE)
    kps["x"] = [10.5, 200.25, 3.0]
    kps["y"] = [20.0, 7.125, 640.5]
    kps["size"] = [4.8, 12.0, -9.6]
    desc = np.arange(3 * 61, dtype=np.uint32).reshape(3, 61).astype(np.uint8)
    p = tmp_path / "k.txt"
    A.save_keypoints(p, kps, desc, True)
    lines = p.read_text().splitlines()
    assert lines[0] == "61" and lines[1] == "3" and len(lines) == 5
    for i, ln in enumerate(lines[2:]):
        f = ln.split()
        assert len(f) == 5 + 61
        x, y, a, b, c = map(float, f[:5])
        r2 = float(np.float32(kps["size"][i]) * np.float32(kps["size"][i]))
        assert (x, y) == (float(kps["x"][i]), float(kps["y"][i]))
        assert b == 0.0 and a == c and abs(a - 1.0 / r2) < 1e-5 / r2
        assert [int(v) for v in f[5:]] == desc[i].tolist()
    # without descriptors the first line is 1 and rows have 5 fields
    A.save_keypoints(p, A.kpts_to_mat(kps), None, False)
    lines = p.read_text().splitlines()
    assert lines[:2] == ["1", "3"] and all(len(ln.split()) == 5 for ln in lines[2:])


def test_read_homography(tmp_path):
    for k in range(1, 7):
        H = A.read_homography(os.path.join(IGUAZU, "H1to%dp" % k))
        assert np.array_equal(H, np.eye(3))
    p = tmp_path / "h"
    p.write_text("2 4 6\n0.5 1 2\n0 0 2\n")
    assert np.allclose(A.read_homography(p), [[1, 2, 3], [0.25, 0.5, 1], [0, 0, 1]])
    assert A.read_homography(tmp_path / "none") is None


def dm_rows(rows):
    m = np.zeros((len(rows), 2), A.DM_DTYPE)
    for i, (q, t1, d1, t2, d2) in enumerate(rows):
        m[i, 0] = (q, t1, 0, d1)
        m[i, 1] = (q, t2, 0, d2)
    return m


def test_nndr_pairs():
    train = np.zeros(3, A.KP_DTYPE)
    train["x"], train["y"] = [1, 2, 3], [10, 20, 30]
    query = np.zeros(4, A.KP_DTYPE)
    query["x"], query["y"] = [5, 6, 7, 8], [50, 60, 70, 80]
    # rows: kept (40 < 0.8 * 60 = 48), dropped (48 == 0.8 * 60: strict), kept, dropped (sentinel 0 vs 0)
    m = dm_rows([(0, 3, 40, 1, 60), (1, 0, 48, 2, 60), (2, 1, 10, 0, 100), (0, 0, 0, 0, 0)])
    p = A.matches2points_nndr(train, query, m, 0.8)
    assert p.tolist() == [[1, 10, 8, 80], [3, 30, 6, 60]]


def test_inliers_homography_closed_form():
    H = np.array([[1.1, 0.05, 12.0], [-0.04, 0.95, -7.0], [1e-4, -2e-4, 1.0]])
    rng = np.random.default_rng(3)
    p1 = rng.uniform(0, 800, (60, 2))
    h = np.c_[p1, np.ones(60)] @ H.T
    p2 = h[:, :2] / h[:, 2:]
    # displacements 0..5 px in random directions: inliers are those <= 2.5
    d = np.linspace(0.0, 5.0, 60)
    ang = rng.uniform(0, 2 * math.pi, 60)
    moved = p2 + d[:, None] * np.c_[np.cos(ang), np.sin(ang)]
    pairs = np.c_[p1, moved].astype(np.float32)
    mask = A.inliers_homography(pairs, H, 2.5)
    far = np.abs(d - 2.5) > 0.02   # float rounding decides only right at the threshold
    assert np.array_equal(mask[far], (d <= 2.5)[far])
    assert np.array_equal(A.inliers_homography(pairs, np.eye(3), 2.5),
                          np.hypot(*(pairs[:, 2:] - pairs[:, :2]).T) <= 2.5)


def planted(seed, n_in, n_out, H, noise=0.3):
    rng = np.random.default_rng(seed)
    p1 = rng.uniform(0, 900, (n_in + n_out, 2))
    h = np.c_[p1, np.ones(len(p1))] @ H.T
    p2 = h[:, :2] / h[:, 2:] + rng.normal(0, noise, (len(p1), 2))
    # outliers: moved 20..300 px
    k = rng.uniform(20, 300, n_out)
    a = rng.uniform(0, 2 * math.pi, n_out)
    p2[n_in:] += k[:, None] * np.c_[np.cos(a), np.sin(a)]
    truth = np.r_[np.ones(n_in, bool), np.zeros(n_out, bool)]
    perm = rng.permutation(len(p1))
    return np.c_[p1, p2].astype(np.float32)[perm], truth[perm]


@pytest.mark.parametrize("outlier_frac", [0.0, 0.3, 0.6])
def test_ransac_recovers_planted_homography(outlier_frac):
    H = np.array([[0.9, -0.2, 40.0], [0.18, 0.92, -25.0], [2e-5, -4e-5, 1.0]])
    n = 300
    n_out = int(n * outlier_frac)
    pairs, truth = planted(11, n - n_out, n_out, H)
    Hr, mask = A.inliers_ransac(pairs, 2.5)
    assert np.array_equal(mask, truth)
    # the estimate maps the inliers to within the noise
    q = np.c_[pairs[truth, :2], np.ones(truth.sum())] @ Hr.T
    err = np.hypot(*(q[:, :2] / q[:, 2:] - pairs[truth, 2:]).T)
    assert err.max() < 2.5 and np.median(err) < 0.6
    assert np.allclose(Hr / Hr[2, 2], H, rtol=2e-2, atol=2e-3 * np.abs(H).max())
    # the same seed gives the same answer
    H2, m2 = A.inliers_ransac(pairs, 2.5)
    assert np.array_equal(H2, Hr) and np.array_equal(m2, mask)


def test_ransac_needs_more_than_eight_pairs():
    pairs, _ = planted(2, 8, 0, np.eye(3))
    H, mask = A.inliers_ransac(pairs, 2.5)
    assert not mask.any() and np.array_equal(H, np.eye(3))
    pairs, _ = planted(2, 9, 0, np.eye(3))
    H, mask = A.inliers_ransac(pairs, 2.5)
    assert mask.all()


def test_keypoint_matrix_round_trip():
    kps = np.zeros(4, A.KP_DTYPE)
    kps["x"], kps["y"], kps["size"], kps["angle"] = [1, 2, 3, 4], [5, 6, 7, 8], [9, -10, 11, 12], [0.5, 1, 2, 3]
    kps["response"], kps["octave"], kps["class_id"] = [0.1, 0.2, 0.3, 0.4], [0, 1, 2, 3], [0, 5, 10, 15]
    m = A.kpts_to_mat(kps)
    assert m.shape == (4, 7) and m.dtype == np.float32
    assert m[1].tolist() == [2, 6, -10, 1, np.float32(0.2), 1, 5]
    assert np.array_equal(A.mat_to_kpts(m), kps)
