"""Generate tests/golden/clip_areas.json: spherical clip areas by brute force.

Independent of Sutherland-Hodgman and of the Van Oosterom-Strackee fan: the
intersection of two convex spherical polygons is found by vertex enumeration
(vertices of each polygon inside the other + all edge/edge great-circle
crossings), ordered by angle around their centroid, and its area is computed
with Girard's theorem (sum of interior angles - (n-2) pi) in 50-digit mpmath.

Inputs are the double-precision unit vectors the oracle and the device code
see (written to the fixture as exact hex floats); the mpmath computation starts
from those exact doubles.  Run:  python tests/golden/gen_clip_fixtures.py
"""
from __future__ import annotations

import json
import os
import random

import mpmath as mp

mp.mp.dps = 50
EPS = mp.mpf("1e-35")


def sincosd(deg):
    """Exact-degree sin/cos rounded to double (multiples of 90 are exact)."""
    r = mp.radians(mp.mpf(deg))
    s, c = mp.sin(r), mp.cos(r)
    if deg % 90 == 0:
        s, c = mp.mpf(round(float(s))), mp.mpf(round(float(c)))
    return float(s), float(c)


def lonlat(lon, lat):
    sl, cl = sincosd(lon)
    sp, cp = sincosd(lat)
    return (cp * cl, cp * sl, sp)


def cell(lon0, lon1, lat0, lat1):
    """Lon-lat cell, CCW from outside: SW, SE, NE, NW."""
    return [lonlat(lon0, lat0), lonlat(lon1, lat0), lonlat(lon1, lat1), lonlat(lon0, lat1)]


def mv(p):
    """Exact double input projected onto the unit sphere in 50 digits."""
    return norm([mp.mpf(x) for x in p])


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def norm(a):
    n = mp.sqrt(dot(a, a))
    return [x / n for x in a]


def sub(a, b):
    return [a[i] - b[i] for i in range(3)]


def edges(P):
    n = len(P)
    out = []
    for i in range(n):
        a, b = P[i], P[(i + 1) % n]
        nrm = cross(a, b)
        if mp.sqrt(dot(nrm, nrm)) < mp.mpf("1e-30"):
            continue  # zero-length edge (pole)
        out.append((a, b, nrm))
    return out


def inside(p, P):
    return all(dot(n, p) >= -EPS for _, _, n in edges(P))


def on_arc(x, a, b, n):
    # x on the minor arc a->b of the great circle with normal n
    return dot(cross(a, x), n) >= -EPS and dot(cross(x, b), n) >= -EPS


def intersection_vertices(P, Q):
    pts = [p for p in P if inside(p, Q)] + [q for q in Q if inside(q, P)]
    for a, b, n1 in edges(P):
        for c, d, n2 in edges(Q):
            x = cross(n1, n2)
            if mp.sqrt(dot(x, x)) < mp.mpf("1e-30"):
                continue
            x = norm(x)
            for y in (x, [-v for v in x]):
                if on_arc(y, a, b, n1) and on_arc(y, c, d, n2):
                    pts.append(y)
    return dedupe(pts)


def dedupe(V):
    uniq = []
    for p in V:
        if all(mp.sqrt(dot(sub(p, u), sub(p, u))) > mp.mpf("1e-28") for u in uniq):
            uniq.append(p)
    return uniq


def girard_area(V):
    """Area of the convex spherical polygon with vertex set V (any order)."""
    V = dedupe(V)
    if len(V) < 3:
        return mp.mpf(0)
    c = norm([sum(v[i] for v in V) for i in range(3)])
    # tangent basis at c
    t = [1, 0, 0] if abs(c[0]) < mp.mpf("0.9") else [0, 1, 0]
    e1 = norm(cross(c, [mp.mpf(x) for x in t]))
    e2 = cross(c, e1)
    V = sorted(V, key=lambda v: mp.atan2(dot(v, e2), dot(v, e1)))
    n = len(V)
    ang = mp.mpf(0)
    for i in range(n):
        a, b, d = V[i - 1], V[i], V[(i + 1) % n]
        # interior angle at b between arcs b->a and b->d
        ta = norm(sub(a, [dot(a, b) * x for x in b]))
        td = norm(sub(d, [dot(d, b) * x for x in b]))
        ang += mp.acos(max(-1, min(1, dot(ta, td))))
    return ang - (n - 2) * mp.pi


def clip_area_mp(subject, clip):
    return girard_area(intersection_vertices([mv(p) for p in subject], [mv(p) for p in clip]))


def hexs(P):
    return [[float(x).hex() for x in p] for p in P]


def cases():
    out = []
    # named cases from the configurations' degenerate geometry
    out.append(("shared_meridian_corner", cell(20, 24, 32, 36), cell(20, 25, 30, 35)))
    out.append(("pole_triangles", cell(0, 4, 86, 90), cell(0, 5, 85, 90)))
    out.append(("south_pole_triangles", cell(356, 360, -90, -86), cell(355, 360, -90, -85)))
    out.append(("sliver_20N_src_below", cell(4, 8, 16, 20), cell(5, 10, 20, 25)))
    out.append(("sliver_20N_dst_below", cell(0, 4, 20, 24), cell(0, 5, 15, 20)))
    out.append(("sliver_80N", cell(4, 8, 76, 80), cell(5, 10, 80, 85)))
    out.append(("sliver_40S", cell(12, 16, -40, -36), cell(10, 15, -45, -40)))
    out.append(("vertex_touch", cell(0, 4, 0, 4), cell(4, 8, 4, 8)))
    out.append(("edge_touch", cell(0, 4, 0, 4), cell(4, 8, 0, 4)))
    out.append(("identical", cell(30, 35, 10, 15), cell(30, 35, 10, 15)))
    out.append(("contained", cell(11, 12, 11, 12), cell(10, 20, 10, 20)))
    out.append(("octant_vs_cell", cell(10, 20, 10, 20), [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)]))
    out.append(("cell_vs_octant_partial", cell(80, 100, 60, 80), [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)]))
    out.append(("large_cells", cell(0, 90, 0, 45), cell(45, 135, -10, 30)))
    out.append(("dateline", cell(356, 360, -4, 0), cell(355, 365, -5, 0)))
    out.append(("quarter_offset", cell(0, 0.25, 0, 0.25), cell(0.1, 0.3, 0.05, 0.2)))
    # random convex lon-lat cells (deterministic seed), small and medium
    rng = random.Random(20261202)
    for k in range(24):
        size = [0.25, 1.0, 4.0, 10.0][k % 4]
        lon0 = rng.uniform(0, 350)
        lat0 = rng.uniform(-80, 80 - size)
        a = cell(lon0, lon0 + size, lat0, lat0 + size)
        lon1 = lon0 + rng.uniform(-0.6, 0.6) * size
        lat1 = max(-89.0, min(89.0 - 1.3 * size, lat0 + rng.uniform(-0.6, 0.6) * size))
        b = cell(lon1, lon1 + 1.3 * size, lat1, lat1 + 1.3 * size)
        out.append((f"random_{k}", a, b))
    return out


def main():
    data = []
    for name, subj, clip in cases():
        area = clip_area_mp(subj, clip)
        a_s = girard_area([mv(p) for p in subj]) if len(subj) >= 3 else mp.mpf(0)
        a_c = girard_area([mv(p) for p in clip])
        data.append({
            "name": name,
            "subject": hexs(subj),
            "clip": hexs(clip),
            "area": mp.nstr(area, 40),
            "subject_area": mp.nstr(a_s, 40),
            "clip_area": mp.nstr(a_c, 40),
        })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "clip_areas.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/gen_clip_fixtures.py", "dps": 50, "cases": data}, f, indent=1)
    print(f"wrote {len(data)} cases to {path}")


if __name__ == "__main__":
    main()
