"""Ragged-polygon test grids (imported by the generic-adapter tests).

``geodesic_dual(level)``: the hexagon/pentagon grid dual to an icosahedron
subdivided ``level`` times (10*4^level + 2 cells: 12 pentagons, the rest
hexagons), the shape of ISEA-type grids.  Dual vertices are the normalised
centroids of the mesh triangles, each shared by three cells; a cell's ring
runs counter-clockwise seen from outside.  Returned as a ragged polygon list
``(vertices (nverts, 3), offsets (ncells+1,), indices)``.
"""
import numpy as np


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(level):
    v, f = _icosahedron()
    verts = list(v)
    for _ in range(level):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.array(nf, dtype=np.int64)
    return np.array(verts), f


def geodesic_dual(level, rotate=None):
    """Dual hexagon/pentagon grid; ``rotate`` (3x3) turns the whole grid."""
    v, f = icosphere(level)
    if rotate is not None:
        v = v @ np.asarray(rotate).T
    cen = v[f].sum(axis=1)
    cen /= np.linalg.norm(cen, axis=1, keepdims=True)
    adj = [[] for _ in range(v.shape[0])]
    for k, tri in enumerate(f):
        for a in tri:
            adj[a].append(k)
    offsets = [0]
    indices = []
    for p, faces in enumerate(adj):
        n = v[p]
        e1 = np.cross(n, [0.0, 0.0, 1.0] if abs(n[2]) < 0.9 else [1.0, 0.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        d = cen[faces] - n
        ang = np.arctan2(d @ e2, d @ e1)
        indices += [faces[i] for i in np.argsort(ang)]
        offsets.append(len(indices))
    return cen, np.array(offsets, dtype=np.int64), np.array(indices, dtype=np.int32)


def rotation(seed):
    """A seeded random rotation matrix."""
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def folded_points(cr, nx, ny, fold_lat):
    """Point matrix ``(nx+1, ny+1, 3)`` of a grid with a RightCenterFolded top
    row: vertex rows 0..ny-1 are parallels from -90 to ``fold_lat`` at
    longitudes i*360/nx, and the top vertex row is row ny-1 reversed
    (P[k, ny] = P[nx-k, ny-1]), so fold-row cells i and nx-1-i are the same
    strip of the polar cap between the chords (lon_i, -lon_i) and
    (lon_i+1, -lon_i+1).  The cells tile the sphere once when each strip is
    counted once."""
    lon = np.arange(nx + 1) * (360.0 / nx)
    lat = -90.0 + np.arange(ny) * ((fold_lat + 90.0) / (ny - 1))
    L, P = np.meshgrid(lon, lat, indexing="ij")
    pts = np.empty((nx + 1, ny + 1, 3))
    pts[:, :ny] = cr.lonlat_to_unit(L, P)
    pts[:, ny] = pts[::-1, ny - 1]
    return pts


def fold_primary(i, nx):
    """Fold-row cells that keep their polygon (ext/ConservativeRegriddingOceananigansExt.jl:146-150)."""
    nh, nq = nx // 2, nx // 4
    return (i < nq) | ((i >= nh) & (i < nh + nq))


def folded_cells(cr, pts):
    """Oracle input of a folded grid: CellBasedGrid cells, with every non-primary
    fold-row cell replaced by the degenerate ghost polygon at P[0, ny-1]."""
    nx, ny = pts.shape[0] - 1, pts.shape[1] - 1
    cells = cr.points_to_cells(pts)
    i = np.arange(nx)
    ghost = (ny - 1) * nx + i[~fold_primary(i, nx)]
    cells[ghost] = pts[0, ny - 1]
    return cells, ghost
