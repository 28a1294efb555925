#
# SpectralElementsHIP -- methods of SpectralElements.jl's hot-path functions
# on device fields, forwarding to libsemhip.so (include/sem_abi.h) via ccall.
#
# Loaded next to SpectralElements (which stays unchanged): `lapl(u, msh)`,
# `gatherScatter(u, msh)`, `pcg(b, opA; ...)` etc. keep their signatures and
# dispatch here when `u`/`b` is a `HIPField`. The host `Mesh` (mesh.jl:34-62)
# is mirrored on the device once per mesh object (a WeakKeyDict cache).
#
# Not executed in this repository's tests (no Julia toolchain here); the same
# entry points are exercised through Python ctypes (semhip/_lib.py) with the
# argument meanings written here.
#
module SpectralElementsHIP

using SpectralElements
import SpectralElements: ABu, lapl, mass, hlmz, gatherScatter, mask, pcg, pcg!, linsolve,
                         grad, advect
using ChainRulesCore

const libsemhip = get(ENV, "SEMHIP_LIB", joinpath(@__DIR__, "..", "lib", "libsemhip.so"))

const SEM_OK = Cint(0)
const SEM_WARN_MAXITER = Cint(1)
const SEM_WARN_STALLED = Cint(2)
const SEM_F64 = Cint(0)
const SEM_F32 = Cint(1)
const SEM_F16_MFMA = Cint(2)
const SEM_BF16_MFMA = Cint(3)
const SEM_PRECOND_IDENTITY = Cint(0)
const SEM_PRECOND_DIAG = Cint(1)
const SEM_PRECOND_DIAG_LOCAL = Cint(2)
const SEM_DEFORM_NONE = Cint(0)
const SEM_DEFORM_SINSIN = Cint(1)
const SEM_B = Cint(8)
const SEM_MULT = Cint(13)

lasterror() = unsafe_string(ccall((:sem_last_error, libsemhip), Cstring, ()))

function check(st::Cint)
    st == SEM_OK && return st
    st == SEM_WARN_MAXITER && return st   # pcg.jl:39 warns and returns
    st == SEM_WARN_STALLED && return st   # mixed solve: no progress, best iterate kept
    error("libsemhip: ", lasterror())
end

# ------------------------------------------------------------------ context
mutable struct Ctx
    h::Ptr{Cvoid}
end
const CTX = Ref{Union{Nothing,Ctx}}(nothing)
function ctx()
    if CTX[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:sem_ctx_create, libsemhip), Cint, (Cint, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                    0, C_NULL, h))
        c = Ctx(h[])
        finalizer(x -> ccall((:sem_ctx_destroy, libsemhip), Cint, (Ptr{Cvoid},), x.h), c)
        CTX[] = c
    end
    CTX[].h
end

# ---------------------------------------------------------- device fields
# A column-major (nr*Ex, ns*Ey) device array: the reference's local layout.
mutable struct HIPField{T} <: AbstractMatrix{T}
    ptr::Ptr{T}
    dims::NTuple{2,Int}
end
Base.size(u::HIPField) = u.dims
Base.pointer(u::HIPField) = u.ptr

function HIPField{T}(::UndefInitializer, m::Int, n::Int) where {T}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    r = ccall((:hipMalloc, "libamdhip64"), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, sizeof(T) * m * n)
    r == 0 || error("hipMalloc failed ($r)")
    u = HIPField{T}(Ptr{T}(p[]), (m, n))
    finalizer(x -> ccall((:hipFree, "libamdhip64"), Cint, (Ptr{Cvoid},), x.ptr), u)
    u
end
Base.similar(u::HIPField{T}) where {T} = HIPField{T}(undef, size(u)...)

function HIPField(a::Array{T,2}) where {T}
    u = HIPField{T}(undef, size(a)...)
    ccall((:hipMemcpy, "libamdhip64"), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
          u.ptr, a, sizeof(a), 1)
    u
end
function Base.Array(u::HIPField{T}) where {T}
    a = Array{T}(undef, size(u))
    ccall((:sem_ctx_synchronize, libsemhip), Cint, (Ptr{Cvoid},), ctx())
    ccall((:hipMemcpy, "libamdhip64"), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
          a, u.ptr, sizeof(a), 2)
    a
end

# -------------------------------------------------------- device mesh mirror
mutable struct DevMesh
    h::Ptr{Cvoid}
end
const MESHES = WeakKeyDict{Any,DevMesh}()

# deform_kind / param: the mesh was built with the config map of the
# benchmarks (sin-sin, amplitude a) or fixU; any other map uploads x, y.
function devmesh(msh::Mesh; deform_kind = SEM_DEFORM_NONE, deform_param = 0.0)
    get!(MESHES, msh) do
        nr, ns = length(msh.zr), length(msh.zs)
        nx, ny = size(msh.x)
        Ex, Ey = nx ÷ nr, ny ÷ ns
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:sem_mesh_create, libsemhip), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Cint, Cint, Cdouble, Ref{Ptr{Cvoid}}),
                    ctx(), nr, ns, Ex, Ey, msh.ifperiodic[1], msh.ifperiodic[2],
                    deform_kind, deform_param, h))
        if deform_kind == SEM_DEFORM_NONE
            x, y = HIPField(msh.x), HIPField(msh.y)   # user geometry: GPU recomputes jac/G
            check(ccall((:sem_mesh_set_coords, libsemhip), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), ctx(), h[], x.ptr, y.ptr))
        end
        d = DevMesh(h[])
        finalizer(x -> ccall((:sem_mesh_destroy, libsemhip), Cint, (Ptr{Cvoid},), x.h), d)
        d
    end
end

function maskbits(bc::Vector{Char}, msh::Mesh)
    b = Ref{Cint}(0)
    check(ccall((:sem_mask_bits, libsemhip), Cint, (Ptr{Cvoid}, Cstring, Ref{Cint}),
                devmesh(msh).h, String(bc), b))
    b[]
end

# ---------------------------------------------------------- L3 operators
# lapl(u, msh), lapl.jl:26-36
function lapl(u::HIPField{Float64}, msh::Mesh)
    y = similar(u)
    check(ccall((:sem_lapl, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, u.ptr, y.ptr))
    y
end

# mass(u, msh), mass.jl:12-22
function mass(u::HIPField{Float64}, msh::Mesh)
    y = similar(u)
    check(ccall((:sem_mass, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, u.ptr, y.ptr))
    y
end

# hlmz(u, ν, k, msh), hlmz.jl:12-19 (ν, k scalars or device fields)
function hlmz(u::HIPField{Float64}, ν, k, msh::Mesh)
    y = similar(u)
    νs, νf = ν isa Number ? (Float64(ν), Ptr{Cdouble}(C_NULL)) : (0.0, ν.ptr)
    ks, kf = k isa Number ? (Float64(k), Ptr{Cdouble}(C_NULL)) : (0.0, k.ptr)
    check(ccall((:sem_hlmz, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Cdouble, Ptr{Cdouble}, Cdouble, Ptr{Cdouble},
                 Ptr{Cdouble}), ctx(), devmesh(msh).h, u.ptr, νs, νf, ks, kf, y.ptr))
    y
end

# gatherScatter(u, msh), gatherScatter.jl:18-21
function gatherScatter(u::HIPField{Float64}, msh::Mesh)
    y = similar(u)
    check(ccall((:sem_gather_scatter, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), ctx(), devmesh(msh).h, u.ptr, y.ptr))
    y
end

# mask(u, M), mask.jl:10-18 (M a device 0/1 field)
function mask(u::HIPField{Float64}, M::HIPField{Float64})
    y = similar(u)
    check(ccall((:sem_mask, libsemhip), Cint, (Ptr{Cvoid}, Int64, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), length(u), M.ptr, u.ptr, y.ptr))
    y
end

# ABu(As, Br, u), ABu.jl:9-37 (As/Br host matrices, [] = identity)
function ABu(As, Br, u::HIPField{Float64})
    m, n = size(u)
    dev(A) = isempty(A) ? (nothing, 0, 0) : (HIPField(Matrix{Float64}(A)), size(A)...)
    Ad, ma, na = dev(As)
    Bd, mb, nb = dev(Br)
    m1 = Bd === nothing ? m : (m ÷ nb) * mb
    n1 = Ad === nothing ? n : (n ÷ na) * ma
    y = HIPField{Float64}(undef, m1, n1)
    check(ccall((:sem_abu, libsemhip), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Ptr{Cdouble}, Cint,
                 Cint, Cint, Ptr{Cdouble}), ctx(), m, n, u.ptr,
                Ad === nothing ? C_NULL : Ad.ptr, ma, na, Bd === nothing ? C_NULL : Bd.ptr, mb, nb,
                0, y.ptr))
    y
end

# ------------------------------------------------------- fused operator
struct OpDesc
    nu::Cdouble
    k::Cdouble
    nu_f::Ptr{Cdouble}
    k_f::Ptr{Cdouble}
    G11::Ptr{Cdouble}
    G12::Ptr{Cdouble}
    G22::Ptr{Cdouble}
    B::Ptr{Cdouble}
    bc_bits::Cint
    assemble::Cint
end

# v -> mask(gatherScatter(hlmz(v, ν, k, msh)), M): opLHS of diffusion.jl:36-45
# as one recognised callable; pcg dispatches it to the fused solver.
mutable struct SEMOperator
    h::Ptr{Cvoid}
    msh::Mesh
    keep::Vector{Any}
end
function SEMOperator(ν, k, msh::Mesh, bc::Vector{Char}; precision = SEM_F64, G = nothing)
    g = G === nothing ? (C_NULL, C_NULL, C_NULL) : map(x -> x.ptr, G)
    d = OpDesc(Float64(ν), Float64(k), C_NULL, C_NULL, g..., C_NULL, maskbits(bc, msh), 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sem_operator_create, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ref{OpDesc}, Cint, Ref{Ptr{Cvoid}}),
                ctx(), devmesh(msh).h, d, precision, h))
    op = SEMOperator(h[], msh, Any[G])
    finalizer(x -> ccall((:sem_operator_destroy, libsemhip), Cint, (Ptr{Cvoid},), x.h), op)
    op
end
function (op::SEMOperator)(u::HIPField{Float64})
    y = similar(u)
    check(ccall((:sem_operator_apply, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                ctx(), op.h, u.ptr, y.ptr))
    y
end
function jacobi(op::SEMOperator)
    d = HIPField{Float64}(undef, size(op.msh.x)...)
    check(ccall((:sem_operator_jacobi, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}),
                ctx(), op.h, d.ptr))
    d
end

struct PcgInfo
    iters::Cint
    converged::Cint
    res::Cdouble
    elapsed_ms::Cdouble
    outer::Cint
    inner_total::Cint
end
PcgInfo() = PcgInfo(0, 0, 0.0, 0.0, 0, 0)

struct JacobiDiag
    d::HIPField{Float64}
end

# pcg(b, opA; opM, mult, ifv, tol, maxiter), pcg.jl:16-60: fused when opA is
# an SEMOperator and opM is the identity or a JacobiDiag.
function pcg(b::HIPField{Float64}, opA::SEMOperator; opM = identity, mult = nothing,
             ifv = false, tol = 1e-8, maxiter = length(b))
    x = similar(b)
    info = Ref(PcgInfo())
    pre, d = opM isa JacobiDiag ? (SEM_PRECOND_DIAG, opM.d.ptr) : (SEM_PRECOND_IDENTITY, C_NULL)
    st = ccall((:sem_pcg, libsemhip), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint,
                Ref{PcgInfo}), ctx(), opA.h, pre, d, b.ptr, x.ptr, tol, min(maxiter, typemax(Cint)), info)
    check(st)
    st == SEM_WARN_MAXITER && println("warning: res: ", info[].res)
    ifv && println("PCG iter: ", info[].iters, ", res: ", info[].res)
    x
end
function pcg!(x::HIPField{Float64}, b::HIPField{Float64}, opA::SEMOperator; kw...)
    y = pcg(b, opA; kw...)
    ccall((:hipMemcpy, "libamdhip64"), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
          x.ptr, y.ptr, sizeof(Float64) * length(x), 3)
    nothing
end

# ------------------------------------------------------------ adjoint
# The rrule of linsolve (adjoint.jl:10-25) for problem(ν) = (SEMOperator with
# ν folded into G, rhs): λ = solver(lhs, ū, true); dJ/dν = -B .* (∇λ·∇u).
function adjoint_dnu(u::HIPField{Float64}, λ::HIPField{Float64}, msh::Mesh)
    g = similar(u)
    check(ccall((:sem_adjoint_dnu, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, u.ptr, λ.ptr, g.ptr))
    g
end

function vardiff_factors(ν::HIPField{Float64}, msh::Mesh)
    G = ntuple(_ -> similar(ν), 3)
    check(ccall((:sem_vardiff_factors, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, ν.ptr, G[1].ptr, G[2].ptr, G[3].ptr))
    G
end

function ChainRulesCore.rrule(::typeof(linsolve), a::HIPField{Float64}, problem, solver)
    lhs, rhs = problem(a)
    u = solver(lhs, rhs, false)
    function linsolve_pullback(ū)
        λ = solver(lhs, unthunk(ū), true)
        g = adjoint_dnu(u, λ, lhs.msh)            # d<λ, lhs(a)(u)>/da
        neg = similar(g)
        check(ccall((:sem_combine, libsemhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}, Ref{Ptr{Cdouble}}, Ptr{Cint},
                     Cint, Cint, Ptr{Cdouble}),
                    ctx(), devmesh(lhs.msh).h, 1, -1.0, g.ptr, C_NULL, -1, 0, neg.ptr))
        return NoTangent(), neg, NoTangent(), NoTangent()
    end
    return u, linsolve_pullback
end

end # module
