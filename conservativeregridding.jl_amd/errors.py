"""Exception types named after the Julia exceptions the reference raises.

* ``ArgumentError``     -- bad ``dims`` (src/regridder/regrid.jl:242-245)
* ``DimensionMismatch`` -- rank / non-spatial axes / length mismatch (regrid.jl:285-298)
* ``MethodError``       -- unsupported field types or options (test/regridding.jl:273-276)
* ``ErrorException``    -- ``error(...)``, e.g. mismatched manifolds (src/regridder/regridder.jl:116)
* ``DefaultIntersectionFailureError`` -- a pair whose intersection failed (regridder.jl:64-73)
"""


class ArgumentError(ValueError):
    pass


class DimensionMismatch(ValueError):
    pass


class MethodError(TypeError):
    pass


class ErrorException(RuntimeError):
    """Julia's ``error(msg)`` (e.g. mismatched manifolds, regridder.jl:116)."""


class DefaultIntersectionFailureError(RuntimeError):
    """A pair whose intersection could not be computed
    (src/regridder/regridder.jl:64-73): ``p1`` is the source cell, ``p2`` the
    destination cell (each a ``(nv, 3)`` ring, or None for a failure of one
    cell's own area), ``e`` the underlying error."""

    def __init__(self, p1, p2, e):
        super().__init__("Intersection failed with the following error.  Capture this error object and access "
                         f"`err.p1` and `err.p2` to access the polygons that failed to intersect.  {e}")
        self.p1, self.p2, self.e = p1, p2, e
