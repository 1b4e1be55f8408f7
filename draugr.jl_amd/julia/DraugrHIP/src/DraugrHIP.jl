"""
    DraugrHIP

Julia side of the gfx950 backend: methods of Draugr.jl's solve / resetup API
(`amg_cycle!` src/cycle.jl:12, `amg_solve!` src/cycle.jl:110, `amg_resetup!`
src/resetup.jl:20, `ldiv!` src/preconditioner.jl:76-102) on a new hierarchy
type, `HIPHierarchy`, in the way ext/DraugrAMDGPUExt.jl:28-52 adds methods
for ROCSparseMatrixCSR.  Every call goes through `ccall` into
`libdraugr_hip.so` (declarations in include/draugr_hip.h); device vectors
(`ROCVector{Float64}`) are passed as raw HBM pointers, host vectors are
copied by the library.  A `ChainRulesCore.rrule` gives the adjoint of the
solve (lambda = A^-1 dL/dx, dL/db = lambda, dL/dA_ij = -lambda_i x_j).

Not exercised in this repository (no Julia toolchain); the same C-ABI is
tested through its Python binding.
"""
module DraugrHIP

using Draugr
using Draugr: AMGConfig
using AMDGPU
using ChainRulesCore
using LinearAlgebra
using SparseArrays

export HIPHierarchy, HIPPreconditioner, hip_setup, pcg_solve!, solve_7pt

const libdraugr_hip = get(ENV, "DRAUGR_HIP_LIB",
                          normpath(joinpath(@__DIR__, "..", "..", "..", "lib", "libdraugr_hip.so")))

struct DraugrHIPError <: Exception
    msg::String
end
Base.showerror(io::IO, e::DraugrHIPError) = print(io, "DraugrHIPError: ", e.msg)

last_error() = unsafe_string(ccall((:draugr_amg_last_error, libdraugr_hip), Cstring, ()))
check(r::Integer) = r < 0 ? throw(DraugrHIPError(last_error())) : r

# ── config: Draugr's AMGConfig -> the JSON of draugr_amg_config_from_json ──
_coarsening_name(c) = c isa Draugr.AggregationCoarsening ? "aggregation" :
                      c isa Draugr.PMISCoarsening ? "pmis" :
                      c isa Draugr.HMISCoarsening ? "hmis" :
                      c isa Draugr.RSCoarsening ? "rs" :
                      c isa Draugr.SmoothedAggregationCoarsening ? "smoothed_aggregation" :
                      c isa Draugr.AggressiveCoarsening ? "aggressive_" * String(c.base) :
                      throw(ArgumentError("unsupported coarsening $(typeof(c))"))

function _interp_json(ip)
    ip isa Draugr.DirectInterpolation && return "{\"type\":\"direct\",\"trunc_factor\":$(ip.trunc_factor)}"
    ip isa Draugr.StandardInterpolation && return "{\"type\":\"standard\",\"trunc_factor\":$(ip.trunc_factor)}"
    return "{\"type\":\"extended_i\",\"trunc_factor\":$(ip.trunc_factor),\"max_elements\":$(ip.max_elements)," *
           "\"norm_p\":$(ip.norm_p),\"rescale\":$(ip.rescale)}"
end

const _SMOOTHERS = Dict(Draugr.JacobiSmootherType => "jacobi", Draugr.L1JacobiSmootherType => "l1_jacobi",
                        Draugr.ColoredGaussSeidelType => "colored_gs",
                        Draugr.L1ColoredGaussSeidelType => "l1_colored_gs")

function config_json(c::AMGConfig; precision::Symbol=:fp64)
    co = c.coarsening
    parts = ["\"coarsening\":\"$(_coarsening_name(co))\"", "\"theta\":$(co.θ)",
             "\"smoother\":\"$(_SMOOTHERS[typeof(c.smoother)])\"", "\"jacobi_omega\":$(c.jacobi_omega)",
             "\"cycle\":\"$(lowercase(String(c.cycle_type)))\"", "\"max_levels\":$(c.max_levels)",
             "\"max_coarse_size\":$(c.max_coarse_size)", "\"pre_smoothing_steps\":$(c.pre_smoothing_steps)",
             "\"post_smoothing_steps\":$(c.post_smoothing_steps)", "\"max_row_sum\":$(c.max_row_sum)",
             "\"initial_coarsening_levels\":$(c.initial_coarsening_levels)",
             "\"precision\":\"$(precision)\""]
    hasproperty(co, :interpolation) && push!(parts, "\"interpolation\":" * _interp_json(co.interpolation))
    ic = c.initial_coarsening
    push!(parts, "\"initial_coarsening\":\"$(_coarsening_name(ic))\"", "\"initial_theta\":$(ic.θ)")
    hasproperty(ic, :interpolation) && push!(parts, "\"initial_interpolation\":" * _interp_json(ic.interpolation))
    return "{" * join(parts, ",") * "}"
end

mutable struct ConfigHandle
    h::Int32
    function ConfigHandle(c::AMGConfig; precision::Symbol=:fp64)
        h = check(ccall((:draugr_amg_config_from_json, libdraugr_hip), Int32, (Cstring,),
                        config_json(c; precision)))
        obj = new(Int32(h))
        finalizer(o -> ccall((:draugr_amg_config_free, libdraugr_hip), Int32, (Int32,), o.h), obj)
    end
end

# ── hierarchy ────────────────────────────────────────────────────────────────
"""
    HIPHierarchy

Handle of a hierarchy set up on the host (the reference's setup, restated in
the library) and resident in HBM: level operators, P / Pt, Galerkin triple
maps, smoother diagonals and the factored coarsest matrix.
"""
mutable struct HIPHierarchy
    handle::Int32
    n::Int
    nnz::Int
    config::AMGConfig
    precision::Symbol
end

function _free!(h::HIPHierarchy)
    h.handle > 0 && ccall((:draugr_amg_free, libdraugr_hip), Int32, (Int32,), h.handle)
    h.handle = -1
    return nothing
end

_csr(A::SparseMatrixCSC) = (At = sparse(transpose(A)); (Int32.(At.colptr), Int32.(At.rowval), Float64.(At.nzval)))

"""
    hip_setup(A, config; precision=:fp64, allow_partial_resetup=true) -> HIPHierarchy

`amg_setup` (src/setup.jl:47-87) into HBM.  A SparseMatrixCSC is converted to
1-based CSR with sorted columns (its transpose's CSC arrays).
"""
function hip_setup(A::SparseMatrixCSC, config::AMGConfig=AMGConfig(); precision::Symbol=:fp64,
                   allow_partial_resetup::Bool=true)
    rp, cv, nz = _csr(A)
    c = ConfigHandle(config; precision)
    h = check(ccall((:draugr_amg_setup, libdraugr_hip), Int32,
                    (Int32, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Int32, Int32, Int32),
                    size(A, 1), length(nz), rp, cv, nz, c.h, 1, allow_partial_resetup))
    obj = HIPHierarchy(Int32(h), size(A, 1), length(nz), config, precision)
    finalizer(_free!, obj)
    return obj
end

_stream() = AMDGPU.stream().stream  # the task-local HIP stream (hipStream_t)

function Draugr.amg_cycle!(x::ROCVector{Float64}, b::ROCVector{Float64}, h::HIPHierarchy,
                           config::AMGConfig=h.config)
    c = ConfigHandle(config; precision=h.precision)
    check(ccall((:draugr_hip_cycle_device, libdraugr_hip), Int32,
                (Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32, Int32, Ptr{Cvoid}),
                h.handle, h.n, pointer(x), pointer(b), c.h, 0, _stream()))
    return x
end

function Draugr.amg_cycle!(x::Vector{Float64}, b::Vector{Float64}, h::HIPHierarchy, config::AMGConfig=h.config)
    c = ConfigHandle(config; precision=h.precision)
    check(ccall((:draugr_amg_cycle, libdraugr_hip), Int32, (Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32),
                h.handle, h.n, x, b, c.h))
    return x
end

function Draugr.amg_solve!(x::ROCVector{Float64}, b::ROCVector{Float64}, h::HIPHierarchy,
                           config::AMGConfig=h.config; tol::Float64=1e-10, maxiter::Int=100)
    c = ConfigHandle(config; precision=h.precision)
    rel = Ref{Float64}(0.0)
    it = check(ccall((:draugr_hip_solve_device, libdraugr_hip), Int32,
                     (Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32, Float64, Int32, Ptr{Cvoid}, Ref{Float64}),
                     h.handle, h.n, pointer(x), pointer(b), c.h, tol, maxiter, _stream(), rel))
    return x, Int(it)
end

function Draugr.amg_solve!(x::Vector{Float64}, b::Vector{Float64}, h::HIPHierarchy, config::AMGConfig=h.config;
                           tol::Float64=1e-10, maxiter::Int=100)
    c = ConfigHandle(config; precision=h.precision)
    it = check(ccall((:draugr_amg_solve, libdraugr_hip), Int32,
                     (Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32, Float64, Int32),
                     h.handle, h.n, x, b, c.h, tol, maxiter))
    return x, Int(it)
end

"""
    pcg_solve!(x, b, h; tol=1e-8, maxiter=200) -> (x, iterations, relres)

FP64 PCG with one cycle (in the hierarchy's precision) as preconditioner.
"""
function pcg_solve!(x::ROCVector{Float64}, b::ROCVector{Float64}, h::HIPHierarchy; tol::Float64=1e-8,
                    maxiter::Int=200)
    c = ConfigHandle(h.config; precision=h.precision)
    rel = Ref{Float64}(0.0)
    it = check(ccall((:draugr_hip_pcg_device, libdraugr_hip), Int32,
                     (Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32, Float64, Int32, Ptr{Cvoid}, Ref{Float64}),
                     h.handle, h.n, pointer(x), pointer(b), c.h, tol, maxiter, _stream(), rel))
    return x, Int(it), rel[]
end

function Draugr.amg_resetup!(h::HIPHierarchy, A_new::SparseMatrixCSC, config::AMGConfig=h.config;
                             partial::Bool=true, allow_partial_resetup::Bool=true)
    rp, cv, nz = _csr(A_new)
    c = ConfigHandle(config; precision=h.precision)
    check(ccall((:draugr_amg_resetup, libdraugr_hip), Int32,
                (Int32, Int32, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Int32, Int32, Int32, Int32),
                h.handle, size(A_new, 1), length(nz), rp, cv, nz, c.h, 1, partial, allow_partial_resetup))
    return h
end

"""Partial resetup from new values already in HBM (same pattern, CSR order)."""
function Draugr.amg_resetup!(h::HIPHierarchy, nzval::ROCVector{Float64}, config::AMGConfig=h.config)
    check(ccall((:draugr_hip_resetup_device, libdraugr_hip), Int32, (Int32, Int64, Ptr{Float64}, Ptr{Cvoid}),
                h.handle, length(nzval), pointer(nzval), _stream()))
    return h
end

# preconditioner interface (src/preconditioner.jl:76-102): z = 0, one cycle
function LinearAlgebra.ldiv!(z::ROCVector{Float64}, h::HIPHierarchy, r::ROCVector{Float64})
    c = ConfigHandle(h.config; precision=h.precision)
    check(ccall((:draugr_hip_cycle_device, libdraugr_hip), Int32,
                (Int32, Int32, Ptr{Float64}, Ptr{Float64}, Int32, Int32, Ptr{Cvoid}),
                h.handle, h.n, pointer(z), pointer(r), c.h, 1, _stream()))
    return z
end

# ── preconditioner adapters (ext/DraugrLinearSolveExt.jl:18-62,
#    ext/DraugrJutulExt.jl:146-185) on the HIP hierarchy ─────────────────────
"""
    HIPPreconditioner(config; precision=:fp32)

AMG preconditioner on the GPU for LinearSolve.jl (`ldiv!`) and Jutul
(`update_preconditioner!` / `apply!` / `operator_nrows`): the first update
sets the hierarchy up, later updates reset it up (partial), apply is z = 0
and one cycle, zero-copy on `ROCVector`s.
"""
mutable struct HIPPreconditioner
    config::AMGConfig
    precision::Symbol
    hierarchy::Union{Nothing, HIPHierarchy}
end
HIPPreconditioner(config::AMGConfig=AMGConfig(); precision::Symbol=:fp32) =
    HIPPreconditioner(config, precision, nothing)

function update!(P::HIPPreconditioner, A::SparseMatrixCSC)
    if P.hierarchy === nothing
        P.hierarchy = hip_setup(A, P.config; precision=P.precision)
    else
        Draugr.amg_resetup!(P.hierarchy, A, P.config)
    end
    return P
end

LinearAlgebra.ldiv!(z::ROCVector{Float64}, P::HIPPreconditioner, r::ROCVector{Float64}) =
    LinearAlgebra.ldiv!(z, P.hierarchy, r)

# Jutul's interface, when Jutul is loaded alongside (its methods are added
# by the package extension that wires this type in):
#   Jutul.update_preconditioner!(P::HIPPreconditioner, A, b, context, executor) = update!(P, A)
#   Jutul.apply!(x, P::HIPPreconditioner, y) = ldiv!(x, P, y)
#   Jutul.operator_nrows(P::HIPPreconditioner) = P.hierarchy === nothing ? 0 : P.hierarchy.n

# ── differentiation of x = A(k)^-1 b for the 7-point operator ──────────────
function _assemble!(nz::ROCVector{Float64}, h::HIPHierarchy, m::Int, k::ROCVector{Float64})
    check(ccall((:draugr_hip_assemble_7pt, libdraugr_hip), Int32,
                (Int32, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}), h.handle, m, pointer(k), pointer(nz),
                _stream()))
    return nz
end

"""
    solve_7pt(logk, b, h, m; tol=1e-10) -> x

x = A(exp(logk))^-1 b for the cell-centred 7-point operator on m^3 cells
(the hierarchy was set up on one of them; its values are replaced).
"""
function solve_7pt(logk::ROCVector{Float64}, b::ROCVector{Float64}, h::HIPHierarchy, m::Int; tol=1e-10,
                   maxiter=1000)
    k = exp.(logk)
    nz = _assemble!(similar(logk, h.nnz), h, m, k)
    Draugr.amg_resetup!(h, nz)
    x = zero(b)
    _, _, rel = pcg_solve!(x, b, h; tol, maxiter)
    rel < tol || throw(DraugrHIPError("PCG did not converge (relres $rel)"))
    return x
end

function ChainRulesCore.rrule(::typeof(solve_7pt), logk::ROCVector{Float64}, b::ROCVector{Float64},
                             h::HIPHierarchy, m::Int; tol=1e-10, maxiter=1000)
    k = exp.(logk)
    nz = _assemble!(similar(logk, h.nnz), h, m, k)
    x = solve_7pt(logk, b, h, m; tol, maxiter)
    function solve_7pt_pullback(x̄)
        gx = ROCVector{Float64}(unthunk(x̄))
        Draugr.amg_resetup!(h, nz)  # the operator of this forward
        lam = zero(gx)
        pcg_solve!(lam, gx, h; tol, maxiter)  # A symmetric: A^-T = A^-1
        glogk = similar(logk)
        check(ccall((:draugr_hip_grad_logk_7pt, libdraugr_hip), Int32,
                    (Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                    m, pointer(k), pointer(x), pointer(lam), pointer(glogk), _stream()))
        return NoTangent(), glogk, lam, NoTangent(), NoTangent()
    end
    return x, solve_7pt_pullback
end

end # module
