"""Pins the oracle's orientation (FindOrientation, cuda_akaze.cu:2114-2167)
and MLDB (ExtractDescriptors + BuildDescriptor, cuda_akaze.cu:1420-1533,
1786-1807) against closed forms, no GPU.

Orientation: the gradient field has two directions, one inside a disc about
the keypoint and one in the ring around it.  Every sample then holds one of
two vectors, so the histogram reduces to two weighted vectors, and the window
arithmetic fixes the answer:
  * directions at most 6 bins apart share one 7-bin window, which wins, and
    the angle is that of W_in * v_in + W_out * v_out;
  * directions 7 bins apart never share a window, and the angle is that of
    the larger of the two;
with W the summed Gaussian weights exp(-r^2 / (2 * 2.5^2)) of the lattice
samples (i, j), i^2 + j^2 < 36, that land in the disc or the ring.  The test
computes W from that definition and the answer from the vector sum; it also
computes what a wrong sigma, window width, lattice step, keypoint centre,
radius test or bin rule would give, and checks each lies outside the
tolerance, so each of those mistakes in the oracle would fail here.

MLDB: on linear ramps f(px, py) = a px + b py + c every cell sum is the cell's
sample count times f at the cell centroid, so the 87 values and the 486 bits
follow from the sample-position map alone:
  sample (grid x, grid y) -> keypoint + scale * (k cos - l sin, k sin + l cos),
  k = y - size2, l = x - size2 (the grid row k moves image x at angle 0),
  channels Lt, rx = -Lx sin + Ly cos, ry = Lx cos + Ly sin,
  bits per grid (2x2, 3x3, 4x4), per channel, pairs j < i: vals[j] > vals[i].
Checked at angle 0 and pi/2, octave 0 and 1, pattern sizes 10 and 8, with
the same mistake checks (rx and ry swapped, sin sign flipped, grid not
transposed, wrong scale).
"""
import math

import numpy as np
import pytest

import oracle_lib as O

BIN = math.pi / 21.0


# ---------------------------------------------------------------- orientation
def two_direction_field(h, w, cx, cy, r_in2, r_out2, v_in, v_out):
    """Lx, Ly equal to v_in at pixels whose squared distance from (cx, cy)
    is below r_in2, v_out from there up to r_out2 (inclusive), 0 beyond."""
    yy, xx = np.mgrid[0:h, 0:w]
    d2 = (xx - cx) ** 2 + (yy - cy) ** 2
    inner = d2 < r_in2
    outer = (~inner) & (d2 <= r_out2)
    Lx = np.where(inner, v_in[0], np.where(outer, v_out[0], 0.0)).astype(np.float32)
    Ly = np.where(inner, v_in[1], np.where(outer, v_out[1], 0.0)).astype(np.float32)
    return Lx, Ly


def lattice_weights(step, r_in2, r_out2, sigma=2.5, shift=(0, 0), r2_inclusive=False):
    """Summed sample weights in the disc and the ring for a lattice of the
    given step whose centre is `shift` pixels off the field's centre."""
    w_in = w_out = 0.0
    for j in range(-6, 7):
        for i in range(-6, 7):
            r2 = i * i + j * j
            if r2 > 36 or (r2 == 36 and not r2_inclusive):
                continue
            wt = math.exp(-r2 / (2.0 * sigma * sigma))
            d2 = (shift[0] + step * i) ** 2 + (shift[1] + step * j) ** 2
            if d2 < r_in2:
                w_in += wt
            elif d2 <= r_out2:
                w_out += wt
    return w_in, w_out


def bin_of(theta, rule="trunc"):
    q = theta * 21.0 / math.pi
    return (int(q) if rule == "trunc" else math.floor(q)) + 21


def window_answer(w_in, w_out, v_in, v_out, window=7, rule="trunc"):
    """The orientation the 42-bin, `window`-wide circular sum yields when all
    samples carry one of two vectors."""
    a_in = (w_in * v_in[0], w_in * v_in[1])
    a_out = (w_out * v_out[0], w_out * v_out[1])
    b_in = bin_of(math.atan2(v_in[1], v_in[0]), rule)
    b_out = bin_of(math.atan2(v_out[1], v_out[0]), rule)
    best, best_r = None, -1.0
    for start in range(42):
        cover = {(start + t) % 42 for t in range(window)}
        sx = (a_in[0] if b_in in cover else 0.0) + (a_out[0] if b_out in cover else 0.0)
        sy = (a_in[1] if b_in in cover else 0.0) + (a_out[1] if b_out in cover else 0.0)
        r = sx * sx + sy * sy
        if r > best_r * (1 + 1e-12):
            best, best_r = (sx, sy), r
    ang = math.atan2(best[1], best[0])
    return ang + 2 * math.pi if ang < 0 else ang


def run_orientation(Lx, Ly, x, y, size, octave):
    k = np.zeros(1, O.KP_DTYPE)
    k["x"], k["y"], k["size"], k["octave"], k["class_id"] = x, y, size, octave, 0
    h, w = Lx.shape
    O.lib().orc_orientation(k.ctypes.data, O.fptr(np.ascontiguousarray(Lx)), O.fptr(np.ascontiguousarray(Ly)), w, h)
    return float(k[0]["angle"])


def unit(theta, mag):
    return (mag * math.cos(theta), mag * math.sin(theta))


# keypoint at octave 1, off the pixel grid: centre (int)(pt + 0.5) >> 1 and
# step (int)(0.5 size + 0.5) >> 1 give (60, 50) and 3; rounding pt / 2
# instead would give x = 61, and 0.5 size / 2 rounded would give step 4
PT = (2 * 60 + 1.3, 2 * 50 + 0.4)
SIZE = 14.2
CX, CY, STEP = 60, 50, 3
R_IN2 = STEP * STEP * 10.5     # samples with i^2 + j^2 <= 10 are in the disc
R_OUT2 = STEP * STEP * 40.0    # the ring reaches past the r^2 = 36 samples
TOL = 3e-6


def mistakes(v_in, v_out, truth):
    """Answers of plausible mistakes; each must differ from the truth."""
    alts = {
        "sigma 2.0": window_answer(*lattice_weights(STEP, R_IN2, R_OUT2, sigma=2.0), v_in, v_out),
        "sigma 3.0": window_answer(*lattice_weights(STEP, R_IN2, R_OUT2, sigma=3.0), v_in, v_out),
        "window 5": window_answer(*lattice_weights(STEP, R_IN2, R_OUT2), v_in, v_out, window=5),
        "window 9": window_answer(*lattice_weights(STEP, R_IN2, R_OUT2), v_in, v_out, window=9),
        "step 4": window_answer(*lattice_weights(4, R_IN2, R_OUT2), v_in, v_out),
        "unit lattice": window_answer(*lattice_weights(1, R_IN2, R_OUT2), v_in, v_out),
        "centre x + 1": window_answer(*lattice_weights(STEP, R_IN2, R_OUT2, shift=(1, 0)), v_in, v_out),
        "r^2 <= 36": window_answer(*lattice_weights(STEP, R_IN2, R_OUT2, r2_inclusive=True), v_in, v_out),
        "floor bins": window_answer(*lattice_weights(STEP, R_IN2, R_OUT2), v_in, v_out, rule="floor"),
    }
    return {k: v for k, v in alts.items() if abs(v - truth) <= 100 * TOL}


@pytest.mark.parametrize("case", ["six_bins_apart", "seven_bins_apart", "negative_angle_truncation"])
def test_orientation_two_direction_field(case):
    h, w = 110, 130
    if case == "six_bins_apart":
        # bins 23 and 29: one 7-bin window holds both, the answer is their sum
        v_in, v_out = unit(2.5 * BIN, 1.0), unit(8.5 * BIN, 0.6)
    elif case == "seven_bins_apart":
        # bins 23 and 30: no window holds both; the ring's weighted vector is
        # the longer one with sigma 2.5 but not with sigma 2.0
        v_in, v_out = unit(2.5 * BIN, 1.0), unit(9.5 * BIN, 2.2)
    else:
        # -6.5 bins truncates to bin 15 (floor would give 14), +0.5 is bin 21:
        # six apart under the reference's (int) cast, seven under floor
        v_in, v_out = unit(-6.5 * BIN, 1.0), unit(0.5 * BIN, 0.9)
    w_in, w_out = lattice_weights(STEP, R_IN2, R_OUT2)
    truth = window_answer(w_in, w_out, v_in, v_out)
    if case == "seven_bins_apart":
        assert w_out * 2.2 > w_in * 1.0             # the ring wins...
        assert abs(truth - 9.5 * BIN) < 1e-12       # ...so the answer is its direction
    else:
        sx = w_in * v_in[0] + w_out * v_out[0]
        sy = w_in * v_in[1] + w_out * v_out[1]
        assert abs(truth - (math.atan2(sy, sx) % (2 * math.pi))) < 1e-12
    Lx, Ly = two_direction_field(h, w, CX, CY, R_IN2, R_OUT2, v_in, v_out)
    got = run_orientation(Lx, Ly, PT[0], PT[1], SIZE, 1)
    assert abs(got - truth) < TOL, (got, truth)


def test_orientation_mistake_coverage():
    """Every mistake of the list changes at least one of the three answers by
    far more than the tolerance; the summed-window field alone catches every
    mistake that moves the weights (sigma, lattice, centre, radius test)."""
    fields = [(unit(2.5 * BIN, 1.0), unit(8.5 * BIN, 0.6)), (unit(2.5 * BIN, 1.0), unit(9.5 * BIN, 2.2)),
              (unit(-6.5 * BIN, 1.0), unit(0.5 * BIN, 0.9))]
    names = {"sigma 2.0", "sigma 3.0", "window 5", "window 9", "step 4", "unit lattice", "centre x + 1",
             "r^2 <= 36", "floor bins"}
    caught = set()
    for n, (v_in, v_out) in enumerate(fields):
        truth = window_answer(*lattice_weights(STEP, R_IN2, R_OUT2), v_in, v_out)
        missed = set(mistakes(v_in, v_out, truth))
        if n == 0:
            assert not missed & {"sigma 2.0", "sigma 3.0", "step 4", "unit lattice", "centre x + 1", "r^2 <= 36",
                                 "window 5"}, missed
        caught |= names - missed
    assert caught == names, names - caught


# ---------------------------------------------------------------- MLDB
def grid_sizes(patsize):
    size2 = patsize
    size3 = math.ceil(2.0 * patsize / 3.0)
    size4 = math.ceil(0.5 * patsize)
    return size2, size3, size4, max(3 * size3, 4 * size4)


def cells(patsize):
    """(value index base, list of cells as (x range, y range)) per grid."""
    size2, size3, size4, _ = grid_sizes(patsize)
    out = []
    for n, sz, base in ((2, size2, 0), (3, size3, 4), (4, size4, 13)):
        grid = []
        for cy in range(n):
            for cx in range(n):
                grid.append(((cx * sz, (cx + 1) * sz), (cy * sz, (cy + 1) * sz)))
        out.append((base, grid))
    return out


def ramp(a, b, c, h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    return (a * xx + b * yy + c).astype(np.float32)


def closed_form_vals(kp, patsize, coef, angle_cs, transpose=True, swap=False, scale_delta=0):
    """87 cell sums from centroids.  coef = ((a, b, c) of Lt, Lx, Ly);
    angle_cs = (cos, sin) used for positions and channels."""
    size2 = patsize
    co, si = angle_cs
    o = kp["octave"]
    scale = int(0.5 * kp["size"] / (1 << o) + 0.5) + scale_delta
    xf, yf = kp["x"] / (1 << o), kp["y"] / (1 << o)

    def f(abc, px, py):
        return abc[0] * px + abc[1] * py + abc[2]

    vals = np.zeros(87)
    for base, grid in cells(patsize):
        for ci, ((x0, x1), (y0, y1)) in enumerate(grid):
            n = (x1 - x0) * (y1 - y0)
            l = (x0 + x1 - 1) / 2.0 - size2   # centroid grid column
            k = (y0 + y1 - 1) / 2.0 - size2   # centroid grid row
            if not transpose:
                k, l = l, k
            px = xf + scale * (k * co - l * si)
            py = yf + scale * (k * si + l * co)
            im, dx, dy = (f(coef[t], px, py) for t in range(3))
            rx = -dx * si + dy * co
            ry = dx * co + dy * si
            if swap:
                rx, ry = ry, rx
            vals[3 * (base + ci):3 * (base + ci) + 3] = (n * im, n * rx, n * ry)
    return vals


def bits_from_vals(vals, patsize):
    bits = []
    for base, grid in cells(patsize):
        for ch in range(3):
            for j in range(len(grid)):
                for i in range(j + 1, len(grid)):
                    bits.append(1 if vals[3 * (base + j) + ch] > vals[3 * (base + i) + ch] else 0)
    return np.array(bits, np.uint8)


def min_gap(vals, patsize):
    g = np.inf
    for base, grid in cells(patsize):
        for ch in range(3):
            v = [vals[3 * (base + c) + ch] for c in range(len(grid))]
            for j in range(len(v)):
                for i in range(j + 1, len(v)):
                    g = min(g, abs(v[j] - v[i]))
    return g


def run_mldb(kp, Lt, Lx, Ly, patsize):
    h, w = Lt.shape
    desc = np.zeros(61, np.uint8)
    vals = np.zeros(87, np.float32)
    O.lib().orc_mldb(kp.ctypes.data, O.fptr(Lt), O.fptr(Lx), O.fptr(Ly), w, h, patsize,
                     desc.ctypes.data_as(O.c_u8_p), O.fptr(vals))
    return desc, vals


# slopes whose ratios match no ratio of cell offsets, so no two compared
# cells tie; dyadic values keep every sample exact in float32
COEF = ((1.5, -0.625, 40.0), (0.75, 2.0625, -30.0), (-2.125, 0.375, 60.0))


@pytest.mark.parametrize("patsize", [10, 8])
@pytest.mark.parametrize("angle", [0.0, "pi/2"])
@pytest.mark.parametrize("octave", [0, 1])
def test_mldb_ramps_closed_form(patsize, angle, octave):
    h, w = 140, 150
    kp = np.zeros(1, O.KP_DTYPE)
    # scale = (int)(0.5 * size / 2^o + 0.5) = 3; centre (70, 64) at the level
    kp["x"], kp["y"], kp["size"], kp["octave"], kp["class_id"] = 70 * (1 << octave), 64 * (1 << octave), \
        5.4 * (1 << octave), octave, 0
    if angle == "pi/2":
        kp["angle"] = np.float32(math.pi / 2)
        s, c = O.sincosf(float(kp["angle"][0]))
        cs_true = (float(c), float(s))
        assert abs(cs_true[0]) < 1e-7 and cs_true[1] == 1.0
        cs = (0.0, 1.0)
    else:
        kp["angle"] = 0.0
        cs = (1.0, 0.0)
    Lt, Lx, Ly = (ramp(*COEF[t], h, w) for t in range(3))
    desc, vals = run_mldb(kp[0:1], Lt, Lx, Ly, patsize)
    truth = closed_form_vals(kp[0], patsize, COEF, cs)
    # the centroid sums (cos pi/2 is -4.4e-8 in float: its terms stay far
    # below the gaps between compared cells)
    scale_ref = np.abs(truth).max()
    assert np.max(np.abs(vals - truth)) < 1e-5 * scale_ref + 1e-3, np.max(np.abs(vals - truth))
    assert min_gap(truth, patsize) > 10.0
    want = bits_from_vals(truth, patsize)
    got = np.unpackbits(desc, bitorder="little")
    assert np.array_equal(got[:486], want)
    assert not got[486:].any()
    # each mistake changes many bits
    alts = {
        "rx and ry swapped": closed_form_vals(kp[0], patsize, COEF, cs, swap=True),
        "grid not transposed": closed_form_vals(kp[0], patsize, COEF, cs, transpose=False),
    }
    # on ramps the scale moves every value but no comparison: the value
    # check above catches a wrong scale
    wrong_scale = closed_form_vals(kp[0], patsize, COEF, cs, scale_delta=1)
    assert np.max(np.abs(wrong_scale - truth)) > 1e-2 * scale_ref
    if angle == "pi/2":
        alts["sin sign flipped"] = closed_form_vals(kp[0], patsize, COEF, (cs[0], -cs[1]))
        alts["angle ignored"] = closed_form_vals(kp[0], patsize, COEF, (1.0, 0.0))
    for name, v in alts.items():
        differ = int(np.count_nonzero(bits_from_vals(v, patsize) != want))
        assert differ >= 10, (name, differ)
