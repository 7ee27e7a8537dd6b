# ConservativeRegriddingMI355XExt -- the reference-side binding of libcrhip.so.
#
# A package extension of ConservativeRegridding.jl (weak dependencies AMDGPU and
# ChainRulesCore; Project.toml: [extensions] ConservativeRegriddingMI355XExt =
# ["AMDGPU", "ChainRulesCore"]) that puts the two hot stages behind the C ABI of
# include/crhip.h, with `ccall`:
#
#   * Regridder(manifold, dst, src) for device grids -> cr_regridder_build
#     (src/regridder/regridder.jl:125-163: treeify, intersection_areas, areas,
#     normalize!), on the Spherical or the Planar manifold;
#   * regrid!(dst, R, src; dims, normalize) for ROCArray fields of any rank:
#     extract_source_arraylike / extract_dest_arraylike return a DeviceSlicer (an
#     AbstractDimensionalSlicer, src/regridder/regrid.jl:157-318) that carries
#     the field's cell and field strides, and perform_regridding! on two slicers
#     makes ONE cr_regrid call for all fields whose non-spatial axes flatten into
#     one stride (every field of a dims=1 array), where the reference's
#     NDSliceLoop makes one 1-D mul! per slice;
#   * transpose(R) stays zero-copy (regridder.jl:49-50): a flag on the handle;
#   * ChainRulesCore rules: regrid(R, s) w.r.t. the field, and
#     regrid_on_grids(dst_xyz, src_xyz, s) w.r.t. the field and both grids'
#     vertex coordinates, through cr_regrid_backward.
#
# Only dense ROCArrays take the device path; host arrays keep the reference's
# own methods.  Paths are relative to the ConservativeRegridding.jl repository.
module ConservativeRegriddingMI355XExt

using ConservativeRegridding
import ConservativeRegridding as CR
import ConservativeRegridding: Regridder, AbstractDimensionalSlicer
import ConservativeRegridding.Trees as Trees
import GeometryOps as GO
import GeometryOpsCore as GOCore
import GeoInterface as GI
import SparseArrays
using GeometryOps.UnitSpherical: UnitSphericalPoint
using AMDGPU: ROCArray
import AMDGPU
using ChainRulesCore
using LinearAlgebra

# The library; LIBCRHIP overrides the default path.
const libcrhip = get(ENV, "LIBCRHIP", "libcrhip.so")

const CR_OK = Cint(0)
const CR_F64 = Int32(0)
const CR_F32 = Int32(1)

"""A failing library call whose status has no Julia exception of its own."""
struct CRError <: Exception
    status::Cint
    msg::String
end
Base.showerror(io::IO, e::CRError) = print(io, "libcrhip status ", e.status, ": ", e.msg)

last_error() = unsafe_string(ccall((:cr_last_error, libcrhip), Cstring, ()))

# Status codes -> the exceptions the reference raises (include/crhip.h).
function check(st::Cint)
    st == CR_OK && return nothing
    msg = last_error()
    st == 1 && throw(ArgumentError(msg))
    st == 2 && throw(DimensionMismatch(msg))
    st == 3 && throw(ErrorException(msg))
    throw(CRError(st, msg))
end

# A device array's address for the C side (no data moves).
device_pointer(x::ROCArray) = Ptr{Cvoid}(pointer(x))
# The task-local stream AMDGPU queues the arrays' own work on: launches on it
# stay ordered with that work.
current_stream() = Ptr{Cvoid}(AMDGPU.stream().stream)
elem_code(::Type{Float64}) = CR_F64
elem_code(::Type{Float32}) = CR_F32

# --- grids (what Trees.treeify makes for the device) ---------------------------
mutable struct DeviceGrid
    h::Ptr{Cvoid}
    ncells::Int
    nverts::Int        # caller vertices: vertex gradients are indexed by them
    function DeviceGrid(h::Ptr{Cvoid}, ncells::Integer)
        nv = Ref{Int64}(0)
        check(ccall((:cr_grid_vertex_count, libcrhip), Cint, (Ptr{Cvoid}, Ref{Int64}), h, nv))
        g = new(h, ncells, nv[])
        finalizer(x -> ccall((:cr_grid_destroy, libcrhip), Cvoid, (Ptr{Cvoid},), x.h), g)
        return g
    end
end
GOCore.best_manifold(::DeviceGrid) = GO.Spherical()

# nblocks point blocks of (ni+1) x (nj+1) points, stacked along the last axis
function grid_from_points(X::Array{Float64,3}, nblocks::Integer)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    ni, nj = size(X, 2) - 1, size(X, 3) ÷ nblocks - 1
    check(ccall((:cr_grid_from_points, libcrhip), Cint,
                (Ref{Ptr{Cvoid}}, Ptr{Float64}, Int64, Int64, Int64, Ptr{Cvoid}),
                h, X, nblocks, ni, nj, C_NULL))
    return DeviceGrid(h[], nblocks * ni * nj)
end

# A Matrix{UnitSphericalPoint} of (ni+1) x (nj+1) points (a CellBasedGrid,
# src/trees/interfaces.jl:125-126): its column-major memory is the C layout.
function DeviceGrid(pts::AbstractMatrix{<:UnitSphericalPoint})
    P = Matrix{UnitSphericalPoint{Float64}}(pts)
    X = Array{Float64,3}(undef, 3, size(P)...)
    for j in axes(P, 2), i in axes(P, 1), k in 1:3
        X[k, i, j] = P[i, j][k]
    end
    return grid_from_points(X, 1)
end

# The same from a 3 x (ni+1) x nblocks(nj+1) array of coordinates (the form
# vertex gradients come back in); a cubed sphere is nblocks = 6.
DeviceGrid(xyz::AbstractArray{Float64,3}; nblocks::Integer = 1) = grid_from_points(Array{Float64,3}(xyz), nblocks)

# Any tree with the Trees interface (Trees.getcell(tree) iterates the cells in
# the order the regridder numbers them, regridder.jl:165-178): the polygons'
# open rings become shared vertices + ragged index lists (cr_grid_from_polygons;
# vertex identity is bitwise equality, so vertex gradients sum over the cells
# sharing a vertex).
function DeviceGrid(manifold::GO.Spherical, grid; convex::Bool = true)
    tree = Trees.treeify(manifold, grid)
    id = Dict{NTuple{3,Float64},Int32}()
    xyz = Float64[]
    offsets = Int64[0]
    verts = Int32[]
    for cell in Trees.getcell(tree)
        ring = collect(GI.getpoint(GI.getexterior(cell)))
        n = length(ring)
        if n > 1 && Tuple(ring[1]) == Tuple(ring[end])   # closed ring: drop the repeat
            n -= 1
        end
        for p in view(ring, 1:n)
            key = (Float64(p[1]), Float64(p[2]), Float64(p[3]))
            k = get!(id, key) do
                append!(xyz, key)
                Int32(length(xyz) ÷ 3 - 1)
            end
            push!(verts, k)
        end
        push!(offsets, length(verts))
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    ncell = length(offsets) - 1
    check(ccall((:cr_grid_from_polygons, libcrhip), Cint,
                (Ref{Ptr{Cvoid}}, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Int32}, Int64, Int32, Ptr{Cvoid}),
                h, xyz, length(xyz) ÷ 3, offsets, verts, ncell, convex, C_NULL))
    return DeviceGrid(h[], ncell)
end

# The same on the Planar manifold: the polygons of any planar tree or polygon
# collection (GO.SpatialTreeInterface.FlatNoTree(polys), test/usecases/simple.jl;
# matrices of planar rectangles, test/regridding.jl:299-376) as 2-D points
# (cr_grid_from_planar_polygons): regridders between two such grids clip in the
# plane, the DefaultIntersectionOperator{Planar} of regridder.jl:87-94.
function DeviceGrid(manifold::GO.Planar, grid; convex::Bool = true)
    tree = Trees.treeify(manifold, grid)
    id = Dict{NTuple{2,Float64},Int32}()
    xy = Float64[]
    offsets = Int64[0]
    verts = Int32[]
    for cell in Trees.getcell(tree)
        ring = collect(GI.getpoint(GI.getexterior(cell)))
        n = length(ring)
        if n > 1 && (GI.x(ring[1]), GI.y(ring[1])) == (GI.x(ring[end]), GI.y(ring[end]))
            n -= 1
        end
        for p in view(ring, 1:n)
            key = (Float64(GI.x(p)), Float64(GI.y(p)))
            k = get!(id, key) do
                append!(xy, key)
                Int32(length(xy) ÷ 2 - 1)
            end
            push!(verts, k)
        end
        push!(offsets, length(verts))
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    ncell = length(offsets) - 1
    check(ccall((:cr_grid_from_planar_polygons, libcrhip), Cint,
                (Ref{Ptr{Cvoid}}, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Int32}, Int64, Int32, Ptr{Cvoid}),
                h, xy, length(xy) ÷ 2, offsets, verts, ncell, convex, C_NULL))
    return DeviceGrid(h[], ncell)
end

# HEALPix (HealpixRootNode{O}(nside), ext/ConservativeRegriddingHealpixExt.jl):
# nested or ring order, `step` boundary points per side (1: the reference's 4-gons).
function healpix_grid(nside::Integer; nested::Bool = true, step::Integer = 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cr_grid_healpix, libcrhip), Cint, (Ref{Ptr{Cvoid}}, Int64, Int32, Int32, Ptr{Cvoid}),
                h, nside, nested ? 0 : 1, step, C_NULL))
    return DeviceGrid(h[], 12 * Int(nside)^2)
end

# A RightCenterFolded tripolar grid from its 3 x (nx+1) x (ny+1) points
# (ext/ConservativeRegriddingOceananigansExt.jl:119-187, with the fold's ghosts).
function tripolar_grid(xyz::AbstractArray{Float64,3})
    X = Array{Float64,3}(xyz)
    nx, ny = size(X, 2) - 1, size(X, 3) - 1
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cr_grid_tripolar, libcrhip), Cint, (Ref{Ptr{Cvoid}}, Ptr{Float64}, Int64, Int64, Ptr{Cvoid}),
                h, X, nx, ny, C_NULL))
    return DeviceGrid(h[], nx * ny)
end

# --- the intersection matrix: a device handle (the Regridder's W parameter) ----
mutable struct DeviceHandle
    h::Ptr{Cvoid}
    function DeviceHandle(h::Ptr{Cvoid})
        r = new(h)
        finalizer(x -> ccall((:cr_regridder_destroy, libcrhip), Cvoid, (Ptr{Cvoid},), x.h), r)
        return r
    end
end

struct DeviceIntersections <: AbstractMatrix{Float64}
    handle::DeviceHandle
    n_dst::Int
    n_src::Int
    transposed::Bool
end
Base.size(A::DeviceIntersections) = A.transposed ? (A.n_src, A.n_dst) : (A.n_dst, A.n_src)
Base.getindex(::DeviceIntersections, ::Int, ::Int) =
    error("DeviceIntersections lives on the device; copy it with host_csr(A)")
LinearAlgebra.transpose(A::DeviceIntersections) =
    DeviceIntersections(A.handle, A.n_dst, A.n_src, !A.transposed)   # zero copy

# A host SparseMatrixCSC copy (inspection, save_esmf_weights, tests).
function host_csr(A::DeviceIntersections)
    nd, ns, nnz = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:cr_regridder_info, libcrhip), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}),
                A.handle.h, nd, ns, nnz))
    nrows = A.transposed ? ns[] : nd[]
    rp = Vector{Int64}(undef, nrows + 1)
    col = Vector{Int32}(undef, nnz[])
    val = Vector{Float64}(undef, nnz[])
    check(ccall((:cr_regridder_copy_csr, libcrhip), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}),
                A.handle.h, A.transposed, rp, col, val))
    # the CSR of A is the CSC of its transpose
    At = SparseArrays.SparseMatrixCSC(size(A, 2), size(A, 1), rp .+ 1, Int.(col) .+ 1, val)
    return SparseArrays.sparse(transpose(At))
end

struct BuildOpts
    threshold_rel::Float64
    radius::Float64
    normalize::Int32
    clip_side::Int32
    reserved0::Int32
    reserved1::Int32
end

# Regridder(manifold, dst, src; normalize, ...) for device grids.
function Regridder(manifold::GO.Spherical, dst::DeviceGrid, src::DeviceGrid;
                   normalize::Bool = false, threshold_rel::Real = 1e-14, intersection_operator = nothing,
                   kwargs...)
    intersection_operator === nothing ||
        throw(ArgumentError("a per-pair intersection_operator cannot run inside the device build"))
    opts = Ref(BuildOpts(threshold_rel, manifold.radius, normalize, 0, 0, 0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cr_regridder_build, libcrhip), Cint,
                (Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{BuildOpts}, Ptr{Cvoid}),
                h, dst.h, src.h, opts, C_NULL))
    return wrap_regridder(DeviceHandle(h[]), dst.ncells, src.ncells)
end

# Regridder(GO.Planar(), dst, src; ...) for two planar device grids: planar
# areas, no radius (a spherical grid on either side fails in cr_regridder_build
# with the reference's "manifolds must be the same").
function Regridder(::GO.Planar, dst::DeviceGrid, src::DeviceGrid;
                   normalize::Bool = false, threshold_rel::Real = 1e-14, intersection_operator = nothing,
                   kwargs...)
    intersection_operator === nothing ||
        throw(ArgumentError("a per-pair intersection_operator cannot run inside the device build"))
    opts = Ref(BuildOpts(threshold_rel, 1.0, normalize, 0, 0, 0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cr_regridder_build, libcrhip), Cint,
                (Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{BuildOpts}, Ptr{Cvoid}),
                h, dst.h, src.h, opts, C_NULL))
    return wrap_regridder(DeviceHandle(h[]), dst.ncells, src.ncells)
end

# A Regridder around a built handle: the areas are copied to the host once
# (the reference keeps them as Vector{Float64}).
function wrap_regridder(handle::DeviceHandle, n_dst::Integer, n_src::Integer)
    a_dst, a_src = zeros(n_dst), zeros(n_src)
    check(ccall((:cr_regridder_copy_areas, libcrhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                handle.h, a_dst, a_src))
    A = DeviceIntersections(handle, n_dst, n_src, false)
    return Regridder(A, a_dst, a_src, zeros(n_dst), zeros(n_src))
end

# An ESMF weight file (save_esmf_weights' format) read back into a device regridder.
function regridder_from_esmf(path::AbstractString)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cr_regridder_from_esmf, libcrhip), Cint, (Ref{Ptr{Cvoid}}, Cstring, Ptr{Cvoid}), h, path, C_NULL))
    handle = DeviceHandle(h[])
    nd, ns, nnz = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:cr_regridder_info, libcrhip), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}),
                handle.h, nd, ns, nnz))
    return wrap_regridder(handle, nd[], ns[])
end

# normalize!(R) (regridder.jl:54-62): A, a_src and a_dst divided by maximum(A)
# on the device; the host area copies follow.
function LinearAlgebra.normalize!(R::Regridder{<:DeviceIntersections})
    A = R.intersections
    check(ccall((:cr_regridder_normalize, libcrhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), A.handle.h, C_NULL))
    a_dst = A.transposed ? R.src_areas : R.dst_areas
    a_src = A.transposed ? R.dst_areas : R.src_areas
    check(ccall((:cr_regridder_copy_areas, libcrhip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                A.handle.h, a_dst, a_src))
    return R
end

# --- regrid!: the device slicer -----------------------------------------------
"""
    DeviceSlicer

The AbstractDimensionalSlicer of a ROCArray field: the spatial axis `dims`
(any stride) and the other axes, whose leading run that flattens into one
stride is the field count of one launch; axes left over are looped
(column-major), one launch each.
"""
struct DeviceSlicer{T,N} <: AbstractDimensionalSlicer
    array::ROCArray{T,N}
    dims::Int
end
Base.parent(s::DeviceSlicer) = s.array
ncells(s::DeviceSlicer) = size(s.array, s.dims)
cell_stride(s::DeviceSlicer) = stride(s.array, s.dims)
other_axes(s::DeviceSlicer{T,N}) where {T,N} = Tuple(d for d in 1:N if d != s.dims)

# The interface's 1-D views, column-major over the non-spatial axes (regrid.jl:205-210);
# perform_regridding! below does not iterate them.
function CR.slice_views(s::DeviceSlicer{T,N}) where {T,N}
    A, D = s.array, s.dims
    oth = Tuple(axes(A, d) for d in other_axes(s))
    return (view(A, ntuple(d -> d == D ? Colon() : I[d < D ? d : d - 1], N)...) for I in CartesianIndices(oth))
end

# (fields per launch, their stride, the looped axes) of the non-spatial axes
function flatten_fields(s::DeviceSlicer)
    ax = other_axes(s)
    isempty(ax) && return (1, 1, ())
    nf, fs = size(s.array, ax[1]), stride(s.array, ax[1])
    k = 2
    while k <= length(ax) && stride(s.array, ax[k]) == fs * nf
        nf *= size(s.array, ax[k])
        k += 1
    end
    return (nf, fs, ax[k:end])
end

function CR.extract_source_arraylike(src::ROCArray{T,N}, ::Regridder{<:DeviceIntersections};
                                     dims::Int = 1, kwargs...) where {T<:Union{Float32,Float64},N}
    return DeviceSlicer{T,N}(src, N == 1 ? 1 : dims)
end
function CR.extract_dest_arraylike(dst::ROCArray{T,N}, ::Regridder{<:DeviceIntersections};
                                   dims::Int = 1, kwargs...) where {T<:Union{Float32,Float64},N}
    return DeviceSlicer{T,N}(dst, N == 1 ? 1 : dims)
end

# The reference's checks for its built-in slicer (regrid.jl:216-220, 270-283),
# plus the regridder's shape.
function check_slicers(dst::DeviceSlicer{T,N}, R, src::DeviceSlicer{T,M}) where {T,N,M}
    1 <= src.dims <= M || throw(ArgumentError("dims=$(src.dims) is out of range for a $M-dimensional array"))
    N == M || throw(DimensionMismatch("source and destination ranks must match; got source rank $M and destination rank $N"))
    dst.dims == src.dims ||
        throw(DimensionMismatch("source and destination regrid dimensions must match; got source dims=$(src.dims) and destination dims=$(dst.dims)"))
    sa = Tuple(axes(src.array, d) for d in other_axes(src))
    da = Tuple(axes(dst.array, d) for d in other_axes(dst))
    sa == da || throw(DimensionMismatch("source and destination non-spatial axes must match; got source axes $sa and destination axes $da"))
    n_out, n_in = size(R.intersections)
    (ncells(src) == n_in && ncells(dst) == n_out) ||
        throw(DimensionMismatch("regridder of size $((n_out, n_in)) cannot map $(ncells(src)) cells to $(ncells(dst))"))
    return nothing
end

# One cr_regrid per looped index, every flattened field in that launch.
# normalize (default true) divides by the destination areas inside the launch:
# finalize_regridding! of a slicer is the reference's no-op (regrid.jl:266-267).
function CR.perform_regridding!(dst::DeviceSlicer{T}, R::Regridder{<:DeviceIntersections}, src::DeviceSlicer{T};
                                normalize::Bool = true, stream::Ptr{Cvoid} = current_stream(),
                                kwargs...) where {T}
    check_slicers(dst, R, src)
    A = R.intersections
    nf, fs, loop_ax = flatten_fields(src)
    nfd, fd, _ = flatten_fields(dst)
    nfd == nf || throw(DimensionMismatch("source and destination fields do not flatten alike"))
    ps, pd = device_pointer(src.array), device_pointer(dst.array)
    nloop = length(loop_ax)
    for I in CartesianIndices(Tuple(size(src.array, d) for d in loop_ax))
        os = sum((I[k] - 1) * stride(src.array, loop_ax[k]) for k in 1:nloop; init = 0)
        od = sum((I[k] - 1) * stride(dst.array, loop_ax[k]) for k in 1:nloop; init = 0)
        check(ccall((:cr_regrid, libcrhip), Cint,
                    (Ptr{Cvoid}, Int32, Int32, Int64, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64, Int64, Int32,
                     Ptr{Cvoid}),
                    A.handle.h, A.transposed, elem_code(T), nf, ps + os * sizeof(T), cell_stride(src), fs,
                    pd + od * sizeof(T), cell_stride(dst), fd, normalize, stream))
    end
    return dst
end

# regrid(R, src) for device fields (the reference's regrid, regrid.jl:322-330,
# allocates a host vector); the destination is a new device array.
function CR.regrid(R::Regridder{<:DeviceIntersections}, src::ROCArray{T,N}; dims::Int = 1, kwargs...) where {T,N}
    sz = collect(size(src))
    sz[N == 1 ? 1 : dims] = size(R.intersections, 1)
    dst = AMDGPU.zeros(T, sz...)
    return CR.regrid!(dst, R, src; dims, kwargs...)
end

# A RightCenterFolded destination: copy each primary's value to its partner
# after the apply (mirror_fold_partners!, ext/ConservativeRegriddingOceananigansExt.jl:207-240);
# an Oceananigans field's finalize_regridding! calls this before copying to the field.
function mirror_fold_partners!(dst::ROCArray{T}, R::Regridder{<:DeviceIntersections}) where {T}
    A = R.intersections
    nf = length(dst) ÷ size(dst, 1)
    check(ccall((:cr_mirror_fold_partners, libcrhip), Cint,
                (Ptr{Cvoid}, Int32, Int32, Int64, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}),
                A.handle.h, A.transposed, elem_code(T), nf, device_pointer(dst), 1, size(dst, 1), current_stream()))
    return dst
end

# --- differentiation ---------------------------------------------------------
field_stride(x::ROCArray) = ndims(x) == 1 ? 1 : stride(x, 2)

# The pullback kernels: gx = A^T (ȳ ./ a_dst), and vertex cotangents
# accumulated into g_dst / g_src (3 doubles per caller vertex) when given.
function regrid_backward!(gx, g_dst, g_src, R::Regridder{<:DeviceIntersections}, dst_grid, src_grid,
                          s::ROCArray{Float64}, y::ROCArray{Float64}, ȳ::ROCArray{Float64}, normalize::Bool)
    A = R.intersections
    nf = length(s) ÷ size(s, 1)
    ptr(x) = x === nothing ? C_NULL : device_pointer(x)
    hnd(g) = g === nothing ? C_NULL : g.h
    check(ccall((:cr_regrid_backward, libcrhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Int64,
                 Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64, Int64,
                 Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                A.handle.h, hnd(dst_grid), hnd(src_grid), A.transposed, normalize, nf,
                device_pointer(s), 1, field_stride(s), device_pointer(y), 1, field_stride(y),
                device_pointer(ȳ), 1, field_stride(ȳ), ptr(gx), 1, gx === nothing ? 1 : field_stride(gx),
                ptr(g_dst), ptr(g_src), current_stream()))
    return gx
end

# regrid(R, s) w.r.t. the field s (dims = 1).  The regridder is a constant here:
# its device matrix is no Julia array a cotangent could live in; differentiate
# w.r.t. the grids with regrid_on_grids.
function ChainRulesCore.rrule(::typeof(CR.regrid), R::Regridder{<:DeviceIntersections}, s::ROCArray{Float64};
                              normalize::Bool = true)
    y = CR.regrid(R, s; normalize)
    function regrid_pullback(ȳ)
        ȳd = ROCArray{Float64}(unthunk(ȳ))
        s̄ = similar(s)
        regrid_backward!(s̄, nothing, nothing, R, nothing, nothing, s, y, ȳd, normalize)
        return NoTangent(), NoTangent(), s̄
    end
    return y, regrid_pullback
end

"""
    regrid_on_grids(dst_xyz, src_xyz, s; radius = 1.0, dst_blocks = 1, src_blocks = 1)

The regridded field regrid(Regridder(dst, src), s) as a function of the source
field s (a device vector or n x k matrix, Float64) and of both grids' vertex
coordinates, each a 3 x (ni+1) x nblocks(nj+1) array of unit vectors (the memory
of a Matrix{UnitSphericalPoint}; a cubed sphere has 6 blocks).  Its rrule
returns the cotangents of all three arrays.
"""
function regrid_on_grids(dst_xyz::AbstractArray{Float64,3}, src_xyz::AbstractArray{Float64,3},
                         s::ROCArray{Float64}; radius::Real = 1.0, dst_blocks::Integer = 1,
                         src_blocks::Integer = 1)
    dg = DeviceGrid(dst_xyz; nblocks = dst_blocks)
    sg = DeviceGrid(src_xyz; nblocks = src_blocks)
    return CR.regrid(Regridder(GO.Spherical(; radius), dg, sg), s)
end

function ChainRulesCore.rrule(::typeof(regrid_on_grids), dst_xyz::AbstractArray{Float64,3},
                              src_xyz::AbstractArray{Float64,3}, s::ROCArray{Float64}; radius::Real = 1.0,
                              dst_blocks::Integer = 1, src_blocks::Integer = 1)
    dg = DeviceGrid(dst_xyz; nblocks = dst_blocks)
    sg = DeviceGrid(src_xyz; nblocks = src_blocks)
    R = Regridder(GO.Spherical(; radius), dg, sg)
    y = CR.regrid(R, s)
    function regrid_on_grids_pullback(ȳ)
        ȳd = ROCArray{Float64}(unthunk(ȳ))
        s̄ = similar(s)
        g_dst = AMDGPU.zeros(Float64, 3, dg.nverts)   # accumulated into
        g_src = AMDGPU.zeros(Float64, 3, sg.nverts)
        regrid_backward!(s̄, g_dst, g_src, R, dg, sg, s, y, ȳd, true)
        # caller vertex v of a point grid is point v of the coordinate array
        # (column-major, i fastest): the gradients have the arrays' own shape
        d̄ = reshape(Array(g_dst), size(dst_xyz))
        s̄xyz = reshape(Array(g_src), size(src_xyz))
        return NoTangent(), d̄, s̄xyz, s̄
    end
    return y, regrid_on_grids_pullback
end

end # module
