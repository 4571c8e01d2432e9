# Partia package extension binding libpartia_mi355x (include/partia_mi355x.h) with ccall.
# Activated by the weak dependency PartiaMI355X_jll, which ships the shared library.
# Adds MI355XComputeBackend methods to Partia.PointSamples_interpolation and
# Partia.LineSamples_interpolation, the two methods the CUDA extension adds
# (Partia.jl/ext/CUDAExt/KernelInterpolation/grid_interpolation/PointSamples_interpolation.jl:12,
#  LineSamples_interpolation.jl:11); StructuredGrid_interpolation dispatches through them, except
# for Cartesian grids, whose samples the library flattens on the device from the axes.
module MI355XExt

using Partia
using Partia.MI355XExtDummy: MI355XComputeBackend
import PartiaMI355X_jll: libpartia_mi355x

const PMX_F32, PMX_F64 = Int32(0), Int32(1)
const PMX_MEM_HOST = Int32(0)
const PMX_PREC_F64, PMX_PREC_F32C = Int32(0), Int32(1)
const PMX_BUILD_WRITEBACK = Int32(1)
const PMX_LINE_INV_H2 = Int32(0x100)   # line weight Inorm / h^2 (an extension; default Inorm / h)

struct PmxParticles
    dtype::Int32; mem::Int32; n::Int64
    x::Ptr{Cvoid}; y::Ptr{Cvoid}; z::Ptr{Cvoid}
    m::Ptr{Cvoid}; h::Ptr{Cvoid}; rho::Ptr{Cvoid}
    ncol::Int32; quant::Ptr{Ptr{Cvoid}}; kernel::Int32
end
struct PmxCatalog
    n_scalar::Int32; scalar_slots::Ptr{Int32}; scalar_shepard::Ptr{UInt8}
    n_grad::Int32; grad_slots::Ptr{Int32}
    n_div::Int32; div_slots::Ptr{Int32}
    n_curl::Int32; curl_slots::Ptr{Int32}
end
struct PmxPoints
    mem::Int32; n::Int64; x::Ptr{Float64}; y::Ptr{Float64}; z::Ptr{Float64}
end
struct PmxLines
    mem::Int32; n::Int64
    ox::Ptr{Float64}; oy::Ptr{Float64}; oz::Ptr{Float64}
    dx::Ptr{Float64}; dy::Ptr{Float64}; dz::Ptr{Float64}
end
struct PmxAxes
    mem::Int32; n::NTuple{3,Int64}; a::NTuple{3,Ptr{Float64}}
end
struct PmxOutput
    dtype::Int32; mem::Int32; cols::Ptr{Ptr{Cvoid}}
end

_kernel_id(::Partia.M4_spline) = Int32(0)
_kernel_id(::Partia.M5_spline) = Int32(1)
_kernel_id(::Partia.M6_spline) = Int32(2)
_kernel_id(::Partia.C2_Wendland) = Int32(3)
_kernel_id(::Partia.C4_Wendland) = Int32(4)
_kernel_id(::Partia.C6_Wendland) = Int32(5)
_strategy(::Type{itpGather}) = Int32(0)
_strategy(::Type{itpScatter}) = Int32(1)
_strategy(::Type{itpSymmetric}) = Int32(2)
_dtype(::Type{Float32}) = PMX_F32
_dtype(::Type{Float64}) = PMX_F64
_precision(b::MI355XComputeBackend) = b.precision === :f64 ? PMX_PREC_F64 : PMX_PREC_F32C

# one context per (device) for the session; contexts are not re-entrant, so calls are serialised
const _CTX = Dict{Int, Ptr{Cvoid}}()
const _LOCK = ReentrantLock()
function _ctx(dev::Int)
    get!(_CTX, dev) do
        out = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:pmx_create, libpartia_mi355x), Cint, (Ref{Ptr{Cvoid}}, Int32), out, Int32(dev))
        rc == 0 || error("pmx_create failed with status $rc")
        out[]
    end
end

function _check(ctx, rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:pmx_last_error, libpartia_mi355x), Cstring, (Ptr{Cvoid},), ctx))
    rc == -1 && throw(ArgumentError(msg))
    rc == -2 && throw(DimensionMismatch(msg))
    error(msg)
end

# Flatten a catalog to 0-based slots (density-gradient slot 0 -> -1).
function _catalog(cat::InterpolationCatalog)
    ss = Int32[s - 1 for s in cat.scalar_slots]
    sh = UInt8[f ? 1 : 0 for f in cat.scalar_snormalization]
    gs = Int32[s - 1 for s in cat.grad_slots]
    ds = Int32[s - 1 for t in cat.div_slots for s in t]
    cs = Int32[s - 1 for t in cat.curl_slots for s in t]
    return ss, sh, gs, ds, cs
end

function _particles(input::InterpolationInput{3,TF}, qptrs) where {TF}
    x, y, z = input.coord
    PmxParticles(_dtype(TF), PMX_MEM_HOST, Int64(input.Npart), pointer(x), pointer(y), pointer(z),
                 pointer(input.m), pointer(input.h), pointer(input.ρ), Int32(length(input.quant)),
                 pointer(qptrs), _kernel_id(input.smoothed_kernel))
end

function Partia.PointSamples_interpolation(backend::MI355XComputeBackend, grid_template::PointSamples{3,TF},
        input::InterpolationInput{3,TF}, catalog::InterpolationCatalog{3,N,G,Div,C,L},
        itp_strategy::Type{ITP} = itpSymmetric) where {N,G,Div,C,L,TF<:AbstractFloat,ITP<:AbstractInterpolationStrategy}
    grids = ntuple(_ -> similar(grid_template), Val(L))
    ss, sh, gs, ds, cs = _catalog(catalog)
    qptrs = Ptr{Cvoid}[pointer(q) for q in input.quant]
    outp = Ptr{Cvoid}[pointer(g.grid) for g in grids]
    qx, qy, qz = (Float64.(c) for c in grid_template.coor)
    lock(_LOCK) do
        ctx = _ctx(backend.device)
        GC.@preserve input qptrs outp grids ss sh gs ds cs qx qy qz begin
            p = Ref(_particles(input, qptrs))
            cat = Ref(PmxCatalog(length(ss), pointer(ss), pointer(sh), length(gs), pointer(gs),
                                 length(ds) ÷ 3, pointer(ds), length(cs) ÷ 3, pointer(cs)))
            pts = Ref(PmxPoints(PMX_MEM_HOST, length(qx), pointer(qx), pointer(qy), pointer(qz)))
            out = Ref(PmxOutput(_dtype(TF), PMX_MEM_HOST, pointer(outp)))
            rc = ccall((:pmx_point_samples_interpolation, libpartia_mi355x), Cint,
                       (Ptr{Cvoid}, Ref{PmxParticles}, Int32, Int32, Ref{PmxCatalog}, Ref{PmxPoints}, Ref{PmxOutput}),
                       ctx, p, _strategy(itp_strategy), _precision(backend), cat, pts, out)
            _check(ctx, rc)
        end
    end
    return GridBundle(grids, catalog.ordered_names)
end

function Partia.LineSamples_interpolation(backend::MI355XComputeBackend, grid_template::LineSamples{3,TF},
        input::InterpolationInput{3,TF}, catalog::InterpolationCatalog{3,N,0,0,0,N},
        itp_strategy::Type{ITP} = itpScatter; inv_h2::Bool = false) where {N,TF<:AbstractFloat,ITP<:AbstractInterpolationStrategy}
    itp_strategy === itpScatter || throw(ArgumentError(
        "LineSamples_interpolation only supports itpScatter. " *
        "Line-integrated samples do not have a well-defined query smoothing length ha, " *
        "so itpGather and itpSymmetric are not supported."))
    grids = ntuple(_ -> similar(grid_template), Val(N))
    ss, sh, gs, ds, cs = _catalog(catalog)
    qptrs = Ptr{Cvoid}[pointer(q) for q in input.quant]
    outp = Ptr{Cvoid}[pointer(g.grid) for g in grids]
    o = ntuple(d -> Float64.(grid_template.origin[d]), 3)
    dv = ntuple(d -> Float64.(grid_template.direction[d]), 3)
    lock(_LOCK) do
        ctx = _ctx(backend.device)
        GC.@preserve input qptrs outp grids ss sh o dv begin
            p = Ref(_particles(input, qptrs))
            cat = Ref(PmxCatalog(length(ss), pointer(ss), pointer(sh), 0, C_NULL, 0, C_NULL, 0, C_NULL))
            ln = Ref(PmxLines(PMX_MEM_HOST, length(o[1]), pointer(o[1]), pointer(o[2]), pointer(o[3]),
                              pointer(dv[1]), pointer(dv[2]), pointer(dv[3])))
            out = Ref(PmxOutput(_dtype(TF), PMX_MEM_HOST, pointer(outp)))
            rc = ccall((:pmx_line_samples_interpolation, libpartia_mi355x), Cint,
                       (Ptr{Cvoid}, Ref{PmxParticles}, Int32, Int32, Ref{PmxCatalog}, Ref{PmxLines}, Ref{PmxOutput}),
                       ctx, p, _strategy(itp_strategy), _precision(backend) | (inv_h2 ? PMX_LINE_INV_H2 : Int32(0)),
                       cat, ln, out)
            _check(ctx, rc)
        end
    end
    return GridBundle(grids, catalog.ordered_names)
end

# StructuredGrid_interpolation (StructuredGrid_interpolation.jl:46-64) for Cartesian grids: the
# build (LinearBVH!, permuting the input like the reference), then the samples generated on the
# device from the axes in coordinate_grid's column-major order (StructuredGrid.jl:126-141).
function Partia.StructuredGrid_interpolation(backend::MI355XComputeBackend, ::Type{Cartesian},
        grid_template::StructuredGrid{3,TF}, input::InterpolationInput{3,TF},
        catalog::InterpolationCatalog{3,N,G,Div,C,L},
        itp_strategy::Type{ITP} = itpSymmetric) where {N,G,Div,C,L,TF<:AbstractFloat,ITP<:AbstractInterpolationStrategy}
    vals = ntuple(_ -> similar(grid_template.grid), Val(L))
    ss, sh, gs, ds, cs = _catalog(catalog)
    qptrs = Ptr{Cvoid}[pointer(q) for q in input.quant]
    outp = Ptr{Cvoid}[pointer(v) for v in vals]
    ax = ntuple(d -> Float64.(grid_template.axes[d]), 3)
    lock(_LOCK) do
        ctx = _ctx(backend.device)
        GC.@preserve input qptrs outp vals ss sh gs ds cs ax begin
            p = Ref(_particles(input, qptrs))
            rc = ccall((:pmx_build_lbvh, libpartia_mi355x), Cint,
                       (Ptr{Cvoid}, Ref{PmxParticles}, Int32, Ptr{UInt64}, Ptr{Int64}),
                       ctx, p, PMX_BUILD_WRITEBACK, C_NULL, C_NULL)
            _check(ctx, rc)
            cat = Ref(PmxCatalog(length(ss), pointer(ss), pointer(sh), length(gs), pointer(gs),
                                 length(ds) ÷ 3, pointer(ds), length(cs) ÷ 3, pointer(cs)))
            axes = Ref(PmxAxes(PMX_MEM_HOST, ntuple(d -> Int64(length(ax[d])), 3), ntuple(d -> pointer(ax[d]), 3)))
            out = Ref(PmxOutput(_dtype(TF), PMX_MEM_HOST, pointer(outp)))
            rc = ccall((:pmx_interpolate_cartesian_grid, libpartia_mi355x), Cint,
                       (Ptr{Cvoid}, Int32, Int32, Ref{PmxCatalog}, Ref{PmxAxes}, Ref{PmxOutput}),
                       ctx, _strategy(itp_strategy), _precision(backend), cat, axes, out)
            _check(ctx, rc)
        end
    end
    grids = ntuple(i -> StructuredGrid(vals[i], grid_template.axes, grid_template.size), Val(L))
    return GridBundle(grids, catalog.ordered_names)
end

end
