"""Planar grids shared by the planar tests (the Planar manifold).

* :func:`simple_polys` -- test/usecases/simple.jl:9-26: a 2 x 2 grid of unit
  squares (destination) and a diamond plus four corner triangles (source)
* :func:`unit_square_grid` -- test/regridding.jl:299-310's ``make_grid(nx, ny)``
* :func:`jittered_quads`, :func:`triangles` -- two tilings of one rectangle
"""
import numpy as np


def simple_polys():
    """(squares, sources) as lists of open rings in the reference's order and
    orientation (the squares are clockwise there: gridpoints[i,j] ->
    gridpoints[i,j+1] -> gridpoints[i+1,j+1] -> gridpoints[i+1,j])."""
    g = lambda i, j: (i - 1.0, j - 1.0)  # noqa: E731 - gridpoints[i, j] = (i-1, j-1), 1-based
    squares = []
    for j in (1, 2):  # vec of a 2x2 matrix: i fastest
        for i in (1, 2):
            squares.append(np.array([g(i, j), g(i, j + 1), g(i + 1, j + 1), g(i + 1, j)]))
    sources = [np.array([(0, 1), (1, 2), (2, 1), (1, 0)], float),
               np.array([(0, 0), (1, 0), (0, 1)], float),
               np.array([(0, 1), (0, 2), (1, 2)], float),
               np.array([(1, 2), (2, 1), (2, 2)], float),
               np.array([(2, 1), (2, 0), (1, 0)], float)]
    return squares, sources


# A of Regridder(Planar(), squares, sources): square i meets half the diamond
# and its own corner triangle (rows: squares [0,1]^2, [1,2]x[0,1], [0,1]x[1,2],
# [1,2]^2; columns: diamond, triangles at (0,0), (0,2), (2,2), (2,0)).
SIMPLE_A = np.array([[0.5, 0.5, 0.0, 0.0, 0.0],
                     [0.5, 0.0, 0.0, 0.0, 0.5],
                     [0.5, 0.0, 0.5, 0.0, 0.0],
                     [0.5, 0.0, 0.0, 0.5, 0.0]])


def unit_square_grid(nx, ny):
    """Rings of the rectangles ((i-1)/nx, (j-1)/ny) .. (i/nx, j/ny), i fastest."""
    rings = []
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            x0, x1, y0, y1 = (i - 1) / nx, i / nx, (j - 1) / ny, j / ny
            rings.append(np.array([(x0, y0), (x1, y0), (x1, y1), (x0, y1)]))
    return rings


def ccw(ring):
    """The ring counter-clockwise (the oracle's convention)."""
    r = np.asarray(ring, float)
    x, y = r[:, 0], r[:, 1]
    s = np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
    return r if s > 0 else r[::-1]


def table(rings, nv=None):
    """Fixed-width (ncells, nv, 2) table of counter-clockwise rings (narrower
    rings repeat their last point)."""
    nv = nv or max(len(r) for r in rings)
    out = np.empty((len(rings), nv, 2))
    for c, r in enumerate(rings):
        r = ccw(r)
        out[c, :len(r)] = r
        out[c, len(r):] = r[-1]
    return out


def jittered_points(nx, ny, w, h, seed, frac=0.2):
    """(nx+1, ny+1, 2) points of a jittered nx x ny grid of [0,w] x [0,h];
    boundary points move only along their boundary."""
    rng = np.random.default_rng(seed)
    dx, dy = w / nx, h / ny
    X, Y = np.meshgrid(np.arange(nx + 1) * dx, np.arange(ny + 1) * dy, indexing="ij")
    jx = rng.uniform(-frac, frac, X.shape) * dx
    jy = rng.uniform(-frac, frac, Y.shape) * dy
    jx[0, :] = jx[-1, :] = 0.0
    jy[:, 0] = jy[:, -1] = 0.0
    return np.stack([X + jx, Y + jy], axis=-1)


def jittered_quads(nx, ny, w, h, seed, frac=0.2):
    """Rings of the jittered grid's quads, cell i + j*nx."""
    P = jittered_points(nx, ny, w, h, seed, frac)
    return [np.array([P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]]) for j in range(ny) for i in range(nx)]


def triangles(nx, ny, w, h):
    """Each rectangle of a regular nx x ny grid of [0,w] x [0,h] split into two
    counter-clockwise triangles along its rising diagonal."""
    xs, ys = np.linspace(0.0, w, nx + 1), np.linspace(0.0, h, ny + 1)
    rings = []
    for j in range(ny):
        for i in range(nx):
            p00, p10 = (xs[i], ys[j]), (xs[i + 1], ys[j])
            p11, p01 = (xs[i + 1], ys[j + 1]), (xs[i], ys[j + 1])
            rings.append(np.array([p00, p10, p11]))
            rings.append(np.array([p00, p11, p01]))
    return rings
