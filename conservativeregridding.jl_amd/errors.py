"""Exception types named after the Julia exceptions the reference raises.

* ``ArgumentError``     -- bad ``dims`` (src/regridder/regrid.jl:242-245)
* ``DimensionMismatch`` -- rank / non-spatial axes / length mismatch (regrid.jl:285-298)
* ``MethodError``       -- unsupported field types or options (test/regridding.jl:273-276)
"""


class ArgumentError(ValueError):
    pass


class DimensionMismatch(ValueError):
    pass


class MethodError(TypeError):
    pass
