"""Pins the oracle's M-LDB variants (no GPU): the subsampled descriptor of
descriptor_size > 0, one and two channels, and MLDB_UPRIGHT.

* The bit selection (generateDescriptorSubsample, AKAZE.cpp:498-603) is
  restated here from the reference text, driven by the C library's own
  srand(1024) / rand() through ctypes, and must give the oracle's sample
  list and comparisons exactly.
* Every subset bit is a comparison of two cells of one grid and one channel,
  so on real images it must equal the bit of the same comparison in the full
  486-bit descriptor (whose oracle is pinned by closed forms elsewhere); the
  first six picks are forced to the 2x2 grid.
* One channel keeps the intensity comparisons of the full descriptor; the
  second channel of the two-channel form is the cell sum of the gradient
  magnitude, checked in closed form on a ramp.
* MLDB_UPRIGHT leaves every keypoint at angle 0 and describes it with the
  unrotated pattern.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib as O

LO, HI = [0, 4, 13], [4, 13, 29]


def reference_subsample(nbits, P, nch):
    """generateDescriptorSubsample (AKAZE.cpp:498-603) with libc rand()."""
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    fullM = []
    for i in range(3):
        gdiv = i + 2
        psz = math.ceil(2.0 * P / float(np.float32(gdiv)))
        for j in range(gdiv * gdiv):
            for k in range(j + 1, gdiv * gdiv):
                fullM.append([i, psz * (j % gdiv) - P, psz * (j // gdiv) - P, psz * (k % gdiv) - P,
                              psz * (k // gdiv) - P])
    rows = len(fullM)
    assert rows == 162 and nbits <= rows * nch
    libc.srand(1024)
    npicks = math.ceil(nbits / float(nch))
    comps = [[1000, 1000] for _ in range(nch * npicks)]
    samples = []
    full = [r[:] for r in fullM]
    for i in range(npicks):
        k = libc.rand() % (rows - i)
        if i < 6:
            k = i   # the coarser grid's comparisons are always used
        for end in (0, 1):
            key = (full[k][0], full[k][1 + 2 * end], full[k][2 + 2 * end])
            if key in samples:
                j = samples.index(key)
            else:
                samples.append(key)
                j = len(samples) - 1
            for ch in range(nch):
                comps[i * nch + ch][end] = nch * j + ch
        full[k] = full[rows - i - 1][:]
    return np.array(samples, np.int32), np.array(comps[:nbits], np.int32)


@pytest.mark.parametrize("nbits,nch,P", [(256, 3, 10), (486, 3, 10), (128, 3, 10), (64, 1, 10), (162, 1, 10),
                                         (100, 2, 10), (250, 2, 8), (200, 3, 9)])
def test_subsample_matches_the_reference_procedure(nbits, nch, P):
    smp, cmp = O.descriptor_subsample(nbits, P, nch)
    rs, rc = reference_subsample(nbits, P, nch)
    assert np.array_equal(smp, rs)
    assert np.array_equal(cmp, rc)
    # the first six picks are the 2x2 grid's six pairs, in order
    cells2 = [tuple(s) for s in smp[:4]]
    assert all(c[0] == 0 for c in cells2)
    assert O.descriptor_subsample(162 * nch + 1, P, nch) is None


def full_position(nch_full=3):
    a, b = O.mldb_table(0, 10, nch_full)
    return {(int(x), int(y)): i for i, (x, y) in enumerate(zip(a, b))}


@pytest.mark.parametrize("nbits", [64, 256, 400])
def test_subset_table_is_a_selection_of_full_comparisons(nbits):
    a, b = O.mldb_table(nbits, 10, 3)
    pos = full_position()
    assert len(a) == nbits
    for x, y in zip(a, b):
        assert x % 3 == y % 3 and x < y            # one channel, ordered cells
        cx, cy = x // 3, y // 3
        g = [k for k in range(3) if LO[k] <= cx < HI[k]][0]
        assert LO[g] <= cy < HI[g]                # one grid
        assert (int(x), int(y)) in pos
    # the 18 forced bits: pick i, channel c -> full bit c * 6 + i
    for i in range(6):
        for c in range(3):
            assert pos[(int(a[3 * i + c]), int(b[3 * i + c]))] == c * 6 + i


def run_pipeline(img, **kw):
    P = O.Pipeline(img.shape[1], img.shape[0], **kw)
    kps, desc = P.run(img)
    return P, kps, desc


@pytest.fixture(scope="module")
def image():
    import akaze_amd as A
    return A.to_float(A.synth_image(4, 320, 240))


def bits(desc):
    return np.unpackbits(desc, axis=1, bitorder="little")


def test_subset_bits_equal_the_full_descriptor_bits(image):
    _, kf, df = run_pipeline(image)
    pos = full_position()
    for nbits in (64, 256, 485):
        _, ks, ds = run_pipeline(image, descriptor_size=nbits)
        assert ks.tobytes() == kf.tobytes() and ds.shape[1] == (nbits + 7) // 8
        a, b = O.mldb_table(nbits, 10, 3)
        sel = [pos[(int(x), int(y))] for x, y in zip(a, b)]
        assert np.array_equal(bits(ds)[:, :nbits], bits(df)[:, sel])
        assert not bits(ds)[:, nbits:].any()


def test_one_and_two_channel_intensity_bits(image):
    _, kf, df = run_pipeline(image)
    pos = full_position()
    for nch in (1, 2):
        _, k, d = run_pipeline(image, descriptor_channels=nch)
        assert k.tobytes() == kf.tobytes() and d.shape[1] == (162 * nch + 7) // 8
        a, b = O.mldb_table(0, 10, nch)
        ch0 = [i for i in range(len(a)) if a[i] % 3 == 0]
        assert len(ch0) == 162
        sel = [pos[(int(a[i]), int(b[i]))] for i in ch0]
        assert np.array_equal(bits(d)[:, ch0], bits(df)[:, sel])


def test_two_channel_magnitude_closed_form():
    """Lx a positive ramp, Ly = 0: each sample's magnitude is Lx itself, so
    channel 1's cell sums are count x the ramp at the cell centroid."""
    h, w = 120, 130
    yy, xx = np.mgrid[0:h, 0:w]
    Lt = (0.5 * xx - 0.25 * yy + 3.0).astype(np.float32)
    Lx = (0.75 * xx + 2.0625 * yy + 10.0).astype(np.float32)
    Ly = np.zeros((h, w), np.float32)
    kp = np.zeros(1, O.KP_DTYPE)
    kp["x"], kp["y"], kp["size"] = 60, 55, 5.4   # scale 3, angle 0, octave 0
    table = O.mldb_table(0, 10, 2)
    desc, vals = O.mldb_ex(kp, Lt, Lx, Ly, 10, 2, table)
    size = [10, 7, 5]
    for g in range(3):
        n = g + 2
        for c in range(n * n):
            cx, cy = c % n, c // n
            l = (2 * cx * size[g] + size[g] - 1) / 2.0 - 10   # centroid grid column -> image y
            k = (2 * cy * size[g] + size[g] - 1) / 2.0 - 10   # centroid grid row -> image x
            px, py = 60 + 3 * k, 55 + 3 * l
            cnt = size[g] * size[g]
            assert abs(vals[3 * (LO[g] + c) + 1] - cnt * (0.75 * px + 2.0625 * py + 10.0)) < 1e-3
            assert abs(vals[3 * (LO[g] + c)] - cnt * (0.5 * px - 0.25 * py + 3.0)) < 1e-3
    b = np.unpackbits(desc, bitorder="little")[:324]
    a1, a2 = table
    want = (vals[a1] > vals[a2]).astype(np.uint8)
    assert np.array_equal(b, want)


def test_upright_keypoints_keep_angle_zero(image):
    P5, kf, df = run_pipeline(image)
    P4, ku, du = run_pipeline(image, descriptor=4)
    assert np.array_equal(ku[["x", "y", "size", "response", "octave", "class_id"]],
                          kf[["x", "y", "size", "response", "octave", "class_id"]])
    assert not ku["angle"].any() and kf["angle"].any()
    # each upright descriptor is the full descriptor of the same keypoint at angle 0
    table = O.mldb_table(0, 10, 3)
    for i in range(0, len(ku), max(1, len(ku) // 25)):
        l = int(ku[i]["class_id"])
        d, _ = O.mldb_ex(ku[i:i + 1], P4.level(l, 0), P4.level(l, 2), P4.level(l, 3), 10, 3, table)
        assert np.array_equal(d, du[i])


def test_invalid_variants_are_refused():
    for kw in (dict(descriptor=3), dict(descriptor_channels=0), dict(descriptor_channels=4),
               dict(descriptor_size=487), dict(descriptor_size=163, descriptor_channels=1)):
        with pytest.raises(ValueError):
            O.Pipeline(320, 240, **kw)
