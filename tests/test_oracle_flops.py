"""The oracle's FP64 operation count (the numerator of the clip kernels'
roofline figures): closed counts for three cases of Sutherland-Hodgman with a
4-vertex clip polygon, and the build's total equal to the per-pair counts
summed over its candidates.

Per clip edge: the edge normal and its norm (18), one side test per input
vertex plus the first (8 + 8 m), 21 per crossing; the area of an m-gon fan:
40 per triangle + m - 2 adds."""
import math

import numpy as np

from oracle import oracle as o


def _cell(lon0, lon1, lat0, lat1):
    pts = [(lon0, lat0), (lon1, lat0), (lon1, lat1), (lon0, lat1)]
    return np.array([[math.cos(math.radians(b)) * math.cos(math.radians(a)),
                      math.cos(math.radians(b)) * math.sin(math.radians(a)), math.sin(math.radians(b))]
                     for a, b in pts])


CLIP = _cell(0, 3, 0, 3)


def test_subject_inside_every_edge():
    # 4 edges x (18 + 8 + 32), then the 4-gon's area: 2 x 40 + 2
    assert o.clip_area_flops(_cell(1, 2, 1, 2), CLIP) == 4 * 58 + 82


def test_subject_outside_the_second_edge():
    # bottom edge keeps all 4 vertices (58); the right edge drops them all (58); nothing is left
    assert o.clip_area_flops(_cell(10, 11, 1, 2), CLIP) == 2 * 58


def test_subject_crossing_one_edge():
    # the right edge (lon 3) cuts the subject twice: 58 + 2 x 21; the result keeps 4 vertices
    assert o.clip_area_flops(_cell(1, 5, 1, 2), CLIP) == 4 * 58 + 2 * 21 + 82


def test_build_total_is_the_sum_over_candidates():
    src = np.stack([_cell(i, i + 1, j, j + 1) for j in range(-2, 2) for i in range(0, 4)])
    dst = np.stack([_cell(0.5 + 1.5 * i, 2.0 + 1.5 * i, -1.5 + 1.5 * j, 1.5 * j) for j in range(2)
                    for i in range(2)])
    R = o.build(dst, src, clip_is_dst=True, nthreads=2, count_flops=True)
    # the candidates: pairs whose bounding caps (normalised vertex mean, largest
    # chord to a vertex) meet; each is clipped once, source cell by destination cell
    def cap(c):
        m = c.mean(axis=0)
        m /= np.linalg.norm(m)
        return m, np.linalg.norm(c - m, axis=1).max()
    want = kept = 0
    a_src = np.array([o.polygon_area(c) for c in src])
    a_dst = np.array([o.polygon_area(c) for c in dst])
    for i, d in enumerate(dst):
        cd, rd = cap(d)
        for j, s_ in enumerate(src):
            cs, rs = cap(s_)
            if np.sum((cs - cd) ** 2) <= (rs + rd) ** 2:
                f = o.clip_area_flops(s_, d)
                want += f
                if o.clip_area(s_, d) > 1e-14 * min(a_src[j], a_dst[i]):
                    kept += f
    assert R.clip_flops == want > 0
    assert R.clip_flops_kept == kept > 0 and kept < want
    assert o.build(dst, src, clip_is_dst=True, nthreads=1, count_flops=True).clip_flops == want
    assert o.build(dst, src, clip_is_dst=True, nthreads=1).clip_flops == 0  # off by default
