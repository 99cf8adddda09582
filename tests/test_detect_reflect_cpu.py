"""Bounds of the detector's reflected border path, checked on the host.

A border tile of k_detect_fused (akz_detect.hip) takes the constant-offset
path when its LDS grids can hold reflect-101 images: Lsmooth is loaded from
reflected coordinates, and the Lx / Ly positions outside the image are then
copied from their reflections inside the tile's own derivative grid.  This
test walks every tile of every level of many image sizes (small ones
included, where a level is narrower than a tile plus its halo) through the
kernel's condition and checks by brute force that each reflected load lies
inside the image and each copy source inside both the image and the grid.
"""
import numpy as np

import oracle_lib as O

DF_TH = 30  # tile rows of k_detect_fused


def refl(v, n):
    return -v if v < 0 else (2 * n - 2 - v if v >= n else v)


def takes_reflected_path(w, h, s, x0, y0):
    H1 = 2 * s + 1
    interior = x0 >= H1 and x0 + 64 + H1 <= w and y0 >= H1 and y0 + DF_TH + H1 <= h
    return (not interior and 2 <= s <= 4 and H1 <= w - 1 and H1 <= h - 1 and x0 + 63 + H1 <= 2 * w - 2
            and y0 + DF_TH - 1 + H1 <= 2 * h - 2 and (x0 + 64 <= w or w - x0 >= 33)
            and (y0 + DF_TH <= h or h - y0 >= 16))


def check_tile(w, h, s, x0, y0):
    H1, H2 = 2 * s + 1, s + 1
    # loads: rows [y0-H1, y0+TH+H1), columns [x0-H1, x0+64+H1)
    for y in range(y0 - H1, y0 + DF_TH + H1):
        assert 0 <= refl(y, h) < h, (w, h, s, x0, y0, y)
    for x in range(x0 - H1, x0 + 64 + H1):
        assert 0 <= refl(x, w) < w, (w, h, s, x0, y0, x)
    # copies: every derivative-grid position outside the image from its
    # reflection, which must be inside the image and the grid
    for y in range(y0 - H2, y0 + DF_TH + H2):
        for x in (range(x0 - H2, x0 + 64 + H2) if (y < 0 or y >= h) else
                  [x for x in range(x0 - H2, x0 + 64 + H2) if x < 0 or x >= w]):
            ry, rx = refl(y, h), refl(x, w)
            assert 0 <= ry < h and 0 <= rx < w
            assert y0 - H2 <= ry < y0 + DF_TH + H2 and x0 - H2 <= rx < x0 + 64 + H2, (w, h, s, x0, y0, x, y)


def test_reflected_border_tiles_stay_in_bounds():
    sizes = [(W, H) for W in (96, 130, 144, 200, 288, 320, 500, 640, 900, 1280, 1920)
             for H in (80, 96, 120, 136, 144, 152, 240, 270, 380, 480, 675, 1080)]
    # images whose only level is narrower than a tile plus its halo
    sizes += [(W, H) for W in (33, 34, 35, 36, 37, 40, 70, 100) for H in (16, 17, 18, 19, 20, 24, 31, 47)]
    used = 0
    for W, H in sizes:
        g = O.geometry(W, H)
        for lvl in range(g.nlev):
            w, h, s = g.w[lvl], g.h[lvl], g.deriv_step[lvl]
            for y0 in range(0, h, DF_TH):
                for x0 in range(0, w, 64):
                    if takes_reflected_path(w, h, s, x0, y0):
                        check_tile(w, h, s, x0, y0)
                        used += 1
    assert used > 1000


def test_condition_refuses_the_narrow_levels():
    """A 33..36-pixel wide level (or 16..19 rows) cannot reflect a full tile
    plus its halo once: the condition sends those tiles to the general path."""
    for s in (2, 3, 4):
        H1 = 2 * s + 1
        for w in range(20, 40):
            ok = takes_reflected_path(w, 200, s, 0, 60)
            assert ok == (H1 <= w - 1 and 63 + H1 <= 2 * w - 2 and w >= 33), (s, w)
        for h in range(12, 24):
            ok = takes_reflected_path(640, h, s, 64, 0)
            assert ok == (H1 <= h - 1 and DF_TH - 1 + H1 <= 2 * h - 2 and h >= 16), (s, h)
    assert not takes_reflected_path(34, 200, 4, 0, 60)
    assert not takes_reflected_path(640, 17, 4, 64, 0)
    assert np.all([takes_reflected_path(1920, 1080, 2, 0, 0), takes_reflected_path(1920, 1080, 4, 1856, 1050)])
