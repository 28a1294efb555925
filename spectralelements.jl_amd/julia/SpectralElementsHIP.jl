#
# SpectralElementsHIP -- methods of SpectralElements.jl's hot-path functions
# on device fields, forwarding to libsemhip.so (include/sem_abi.h) via ccall.
#
# Loaded next to SpectralElements (which stays unchanged): `lapl(u, msh)`,
# `gatherScatter(u, msh)`, `pcg(b, opA; ...)` etc. keep their signatures and
# dispatch here when `u`/`b` is a `HIPField`. The host `Mesh` (mesh.jl:34-62)
# is mirrored on the device once per mesh object (a WeakKeyDict cache).
#
# A HIPField supports the array algebra the reference's generic code uses
# (pcg.jl:16-60, diffusion.jl:36-77, convectionDiffusion.jl:76-122): `zero`,
# `copy`, `+`, `-`, scalar `*` and `/`, `norm(u, Inf)`, `norm(u)`, `sum`, and
# broadcasts of `+ - * /` over device fields, scalars and constant host arrays
# (`.=`, `.+=`, `.-=`, `.*`, `./`). So the reference's own `pcg` runs
# unchanged with closures over device fields (its per-iteration scalars are
# read on the host), while an `SEMOperator` with the identity or a
# `JacobiDiag` goes to the fused solver. `HIPDiffusion` and
# `HIPConvectionDiffusion` hold the reference equations' state in device
# fields and carry methods of the reference's `opLHS`, `opPrecond`,
# `makeRHS!`, `solve!`, `evolve!`, `step!` and `simulate!`.
#
# Not executed in this repository's tests (no Julia toolchain here); the same
# entry points are exercised through Python ctypes (semhip/_lib.py), and
# tests/test_gpu_julia_replay.py issues, entry by entry, the calls these
# methods make for the Diffusion and ConvectionDiffusion steps.
#
module SpectralElementsHIP

using SpectralElements
import SpectralElements: ABu, lapl, mass, hlmz, gatherScatter, mask, pcg, pcg!, linsolve,
                         grad, advect, opLHS, opPrecond, makeRHS!, solve!, evolve!, step!,
                         simulate!, updateHist!
using ChainRulesCore
using LinearAlgebra
using Base.Broadcast: Broadcasted, AbstractArrayStyle, DefaultArrayStyle

const fixU! = SpectralElements.fixU!   # the no-op setter of mesh.jl:3-5 (not exported)

const libsemhip = get(ENV, "SEMHIP_LIB", joinpath(@__DIR__, "..", "lib", "libsemhip.so"))
const libhip = "libamdhip64"

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
const SEM_PW_MUL = Cint(0)
const SEM_PW_DIV = Cint(1)
const SEM_PW_SCALAR_DIV = Cint(2)
const SEM_PW_DIV_SCALAR = Cint(3)
const SEM_PW_SCALE = Cint(4)

lasterror() = unsafe_string(ccall((:sem_last_error, libsemhip), Cstring, ()))

function check(st::Cint)
    st == SEM_OK && return st
    st == SEM_WARN_MAXITER && return st   # pcg.jl:39 warns and returns
    st == SEM_WARN_STALLED && return st   # mixed solve: no progress, best iterate kept
    error("libsemhip: ", lasterror())
end

# ------------------------------------------------------------------ context
# One context on device 0 with the null stream: every call below is ordered
# with hipMemcpy and with the other calls.
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
synchronize() = check(ccall((:sem_ctx_synchronize, libsemhip), Cint, (Ptr{Cvoid},), ctx()))

# ---------------------------------------------------------- device fields
# A column-major (nr*Ex, ns*Ey) device array: the reference's local layout.
mutable struct HIPField{T} <: AbstractMatrix{T}
    ptr::Ptr{T}
    dims::NTuple{2,Int}
end
Base.size(u::HIPField) = u.dims
Base.pointer(u::HIPField) = u.ptr
Base.show(io::IO, u::HIPField{T}) where {T} = print(io, "HIPField{$T}", u.dims)
Base.show(io::IO, ::MIME"text/plain", u::HIPField) = show(io, u)

function HIPField{T}(::UndefInitializer, m::Int, n::Int) where {T}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    r = ccall((:hipMalloc, libhip), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, sizeof(T) * max(m * n, 1))
    r == 0 || error("hipMalloc failed ($r)")
    u = HIPField{T}(Ptr{T}(p[]), (m, n))
    finalizer(x -> ccall((:hipFree, libhip), Cint, (Ptr{Cvoid},), x.ptr), u)
    u
end
Base.similar(u::HIPField{T}) where {T} = HIPField{T}(undef, size(u)...)
Base.similar(u::HIPField, ::Type{T}, dims::Dims{2}) where {T} = HIPField{T}(undef, dims...)

function upload!(u::HIPField{T}, a::AbstractMatrix) where {T}
    size(a) == size(u) || throw(DimensionMismatch("upload: $(size(a)) vs $(size(u))"))
    h = Matrix{T}(a)
    synchronize()
    r = ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
              u.ptr, h, sizeof(h), 1)
    r == 0 || error("hipMemcpy failed ($r)")
    u
end
HIPField(a::AbstractMatrix{T}) where {T} = upload!(HIPField{T}(undef, size(a)...), a)
function Base.Array(u::HIPField{T}) where {T}
    a = Array{T}(undef, size(u))
    synchronize()
    r = ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
              a, u.ptr, sizeof(a), 2)
    r == 0 || error("hipMemcpy failed ($r)")
    a
end
# one entry (slow: a synchronising copy; for inspection only)
function Base.getindex(u::HIPField{T}, i::Int) where {T}
    v = Ref{T}()
    synchronize()
    ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
          v, u.ptr + sizeof(T) * (i - 1), sizeof(T), 2)
    v[]
end
Base.getindex(u::HIPField, i::Int, j::Int) = u[i + size(u, 1) * (j - 1)]
Base.IndexStyle(::Type{<:HIPField}) = IndexLinear()

# ------------------------------------------------------------ field algebra
const F64 = HIPField{Float64}
cp(p) = p === nothing ? Ptr{Cdouble}(C_NULL) : p.ptr

_copy!(y::F64, x::F64) =   # mask(u, []) is a copy (mask.jl:13)
    (check(ccall((:sem_mask, libsemhip), Cint,
                 (Ptr{Cvoid}, Int64, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                 ctx(), length(x), C_NULL, x.ptr, y.ptr)); y)
_axpby!(y::F64, α, x::F64, β, z::F64) =
    (check(ccall((:sem_axpby, libsemhip), Cint,
                 (Ptr{Cvoid}, Int64, Cdouble, Ptr{Cdouble}, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}),
                 ctx(), length(y), Float64(α), x.ptr, Float64(β), z.ptr, y.ptr)); y)
_pw!(y::F64, op::Cint, α, a::F64, b) =
    (check(ccall((:sem_pointwise, libsemhip), Cint,
                 (Ptr{Cvoid}, Cint, Int64, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                 ctx(), op, length(y), Float64(α), a.ptr, cp(b), y.ptr)); y)
function _dot(a::F64, b, w)
    r = Ref{Cdouble}(0.0)
    check(ccall((:sem_dot, libsemhip), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Cdouble}),
                ctx(), length(a), a.ptr, cp(b), cp(w), r))
    r[]
end
function Base.fill!(u::F64, v)
    if v == 0
        synchronize()
        r = ccall((:hipMemset, libhip), Cint, (Ptr{Cvoid}, Cint, Csize_t), u.ptr, 0, sizeof(Float64) * length(u))
        r == 0 || error("hipMemset failed ($r)")
    else
        upload!(u, fill(Float64(v), size(u)))
    end
    u
end
Base.zero(u::F64) = fill!(similar(u), 0)
Base.copy(u::F64) = _copy!(similar(u), u)
Base.copyto!(y::F64, x::F64) = _copy!(y, x)
Base.:+(a::F64, b::F64) = _axpby!(similar(a), 1.0, a, 1.0, b)
Base.:-(a::F64, b::F64) = _axpby!(similar(a), 1.0, a, -1.0, b)
Base.:-(a::F64) = _pw!(similar(a), SEM_PW_SCALE, -1.0, a, nothing)
Base.:*(s::Number, a::F64) = _pw!(similar(a), SEM_PW_SCALE, s, a, nothing)
Base.:*(a::F64, s::Number) = s * a
Base.:/(a::F64, s::Number) = _pw!(similar(a), SEM_PW_DIV_SCALAR, s, a, nothing)
Base.sum(a::F64) = _dot(a, nothing, nothing)
function LinearAlgebra.norm(a::F64, p::Real = 2)
    if p == Inf
        r = Ref{Cdouble}(0.0)
        check(ccall((:sem_norm_inf, libsemhip), Cint, (Ptr{Cvoid}, Int64, Ptr{Cdouble}, Ref{Cdouble}),
                    ctx(), length(a), a.ptr, r))
        return r[]
    end
    p == 2 || error("norm of a HIPField: p = 2 or Inf")
    sqrt(_dot(a, a, nothing))
end

# constant host arrays met in broadcasts (mult, M) are uploaded once per
# array object
const HOSTCOPY = WeakKeyDict{Any,F64}()
devcopy(a::Array{Float64,2}) = get!(() -> HIPField(a), HOSTCOPY, a)

# ------------------------------------------------------------ broadcasting
# `y .= expr` over device fields: each node of the broadcast tree becomes
# one library call (nested nodes into temporaries, left to right); a sum
# term `s .* x` folds into the coefficient of one axpby.
struct HIPStyle <: AbstractArrayStyle{2} end
HIPStyle(::Val{2}) = HIPStyle()
HIPStyle(::Val{N}) where {N} = DefaultArrayStyle{N}()
Base.BroadcastStyle(::Type{<:HIPField}) = HIPStyle()
Base.BroadcastStyle(::HIPStyle, ::DefaultArrayStyle{N}) where {N} = HIPStyle()
Base.similar(bc::Broadcasted{HIPStyle}, ::Type{T}) where {T} = HIPField{Float64}(undef, size(bc)...)

_isscalar(x) = x isa Number || x isa Base.RefValue
_leaf(x::F64) = x
_leaf(x::Number) = Float64(x)
_leaf(x::Base.RefValue) = _leaf(x[])
_leaf(x::Array{Float64,2}) = devcopy(x)
_leaf(x::AbstractMatrix) = HIPField(Matrix{Float64}(x))
_leaf(bc::Broadcasted) = _eval!(HIPField{Float64}(undef, size(bc)...), bc)

# (coefficient, field) of one term of a sum
function _term(x)
    if x isa Broadcasted && x.f === (*) && length(x.args) == 2
        a, b = x.args
        _isscalar(a) && !_isscalar(b) && return (_leaf(a), _leaf(b))
        _isscalar(b) && !_isscalar(a) && return (_leaf(b), _leaf(a))
    end
    return (1.0, _leaf(x))
end

function _eval!(y::F64, bc::Broadcasted)
    f, args, n = bc.f, bc.args, length(bc.args)
    if f === identity && n == 1
        x = _leaf(args[1])
        x isa F64 ? _copy!(y, x) : fill!(y, x)
    elseif (f === (+) || f === (-)) && n == 1
        x = _leaf(args[1])
        f === (+) ? _copy!(y, x) : _pw!(y, SEM_PW_SCALE, -1.0, x, nothing)
    elseif (f === (+) || f === (-)) && n >= 2
        any(_isscalar, args) && error("HIPField broadcast: field + scalar is not supported")
        α, x = _term(args[1])
        β, z = _term(args[2])
        _axpby!(y, α, x, f === (-) ? -β : β, z)
        for a in args[3:end]                      # n-ary + only
            γ, w = _term(a)
            _axpby!(y, 1.0, y, γ, w)
        end
    elseif f === (*) && n >= 2
        vals = map(_leaf, args)
        s = prod(v for v in vals if v isa Float64; init = 1.0)
        fs = [v for v in vals if v isa F64]
        isempty(fs) && return fill!(y, s)
        length(fs) == 1 ? _pw!(y, SEM_PW_SCALE, s, fs[1], nothing) :
                          _pw!(y, SEM_PW_MUL, 0.0, fs[1], fs[2])
        for v in fs[3:end]
            _pw!(y, SEM_PW_MUL, 0.0, y, v)
        end
        length(fs) > 1 && s != 1.0 && _pw!(y, SEM_PW_SCALE, s, y, nothing)
    elseif f === (/) && n == 2
        a, b = _leaf(args[1]), _leaf(args[2])
        if a isa F64 && b isa F64
            _pw!(y, SEM_PW_DIV, 0.0, a, b)
        elseif a isa F64
            _pw!(y, SEM_PW_DIV_SCALAR, b, a, nothing)
        elseif b isa F64
            _pw!(y, SEM_PW_SCALAR_DIV, a, b, nothing)
        else
            fill!(y, a / b)
        end
    else
        error("HIPField broadcast: unsupported operation $f with $n arguments")
    end
    y
end
Base.copyto!(y::F64, bc::Broadcasted{HIPStyle}) = _eval!(y, bc)
Base.copyto!(y::F64, bc::Broadcasted{Nothing}) = _eval!(y, bc)

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
function _hlmz!(y::F64, u::F64, ν, k, msh::Mesh)
    νs, νf = ν isa Number ? (Float64(ν), Ptr{Cdouble}(C_NULL)) : (0.0, _leaf(ν).ptr)
    ks, kf = k isa Number ? (Float64(k), Ptr{Cdouble}(C_NULL)) : (0.0, _leaf(k).ptr)
    check(ccall((:sem_hlmz, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Cdouble, Ptr{Cdouble}, Cdouble, Ptr{Cdouble},
                 Ptr{Cdouble}), ctx(), devmesh(msh).h, u.ptr, νs, νf, ks, kf, y.ptr))
    y
end

# lapl(u, msh), lapl.jl:26-36
function lapl(u::F64, msh::Mesh)
    y = similar(u)
    check(ccall((:sem_lapl, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, u.ptr, y.ptr))
    y
end
# lapl(u, ν, msh) = ν .* lapl(u, msh), lapl.jl:38-45 (ν a field or a scalar):
# the field ν post-multiplies inside the fused kernel
lapl(u::F64, ν, msh::Mesh) = _hlmz!(similar(u), u, ν, 0.0, msh)
# lapl(u, msh1, msh2), lapl.jl:47-52: the reference forwards to msh1
lapl(u::F64, msh1::Mesh, msh2::Mesh) = lapl(u, msh1)

# mass(u, msh), mass.jl:12-22
function mass(u::F64, msh::Mesh)
    y = similar(u)
    check(ccall((:sem_mass, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, u.ptr, y.ptr))
    y
end
# mass(u, msh1, msh2), mass.jl:25-30: forwards to msh1
mass(u::F64, msh1::Mesh, msh2::Mesh) = mass(u, msh1)

# hlmz(u, ν, k, msh), hlmz.jl:12-19 (ν, k scalars or fields)
hlmz(u::F64, ν, k, msh::Mesh) = _hlmz!(similar(u), u, ν, k, msh)
# hlmz(u, ν, k, msh1, msh2), hlmz.jl:22-30: forwards to msh1
hlmz(u::F64, ν, k, msh1::Mesh, msh2::Mesh) = hlmz(u, ν, k, msh1)

# gatherScatter(u, msh), gatherScatter.jl:18-21
function gatherScatter(u::F64, msh::Mesh)
    y = similar(u)
    check(ccall((:sem_gather_scatter, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), ctx(), devmesh(msh).h, u.ptr, y.ptr))
    y
end

# gather-scatter maps from a local->global table
mutable struct GSMap
    h::Ptr{Cvoid}
end
function GSMap(l2g::AbstractArray{<:Integer}, nglobal::Integer)
    flat = Int32.(vec(l2g))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sem_gsmap_create, libsemhip), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Int64, Ref{Ptr{Cvoid}}),
                ctx(), length(flat), flat, nglobal, h))
    g = GSMap(h[])
    finalizer(x -> ccall((:sem_gsmap_destroy, libsemhip), Cint, (Ptr{Cvoid},), x.h), g)
    g
end
function (g::GSMap)(u::F64)
    y = similar(u)
    check(ccall((:sem_gsmap_apply, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), g.h, u.ptr, y.ptr))
    y
end

# gatherScatter(u, l2g), gatherScatter.jl:27-34 (l2g 1-based, as NNlib.scatter
# takes it); the map is built once per table object
const L2GMAPS = WeakKeyDict{Any,GSMap}()
gatherScatter(u::F64, l2g::AbstractArray{<:Integer}) =
    get!(() -> GSMap(l2g .- 1, maximum(l2g)), L2GMAPS, l2g)(u)

# gatherScatter(u, QQtx, QQty), gatherScatter.jl:8-16: the 0/1 factors of
# mesh.jl:81-82 are read once (per pair of matrix objects) into a map: local
# nodes i, j are one global node when QQtx[i, j] != 0
const QQMAPS = WeakKeyDict{Any,Tuple{Any,GSMap}}()
function _components(Q::AbstractMatrix)
    n = size(Q, 1)
    lab = fill(-1, n)
    g = 0
    for i in 1:n
        if lab[i] < 0
            for j in 1:n
                Q[i, j] != 0 && (lab[j] = g)
            end
            g += 1
        end
    end
    lab, g
end
function gatherScatter(u::F64, QQtx::AbstractMatrix, QQty::AbstractMatrix)
    e = get(QQMAPS, QQtx, nothing)
    if e === nothing || e[1] !== QQty
        gx, nx = _components(QQtx)
        gy, ny = _components(QQty)
        l2g = [gx[i] + nx * gy[j] for i in eachindex(gx), j in eachindex(gy)]
        e = (QQty, GSMap(l2g, nx * ny))
        QQMAPS[QQtx] = e
    end
    e[2](u)
end

# mask(u, M), mask.jl:10-18: M a device field, a host 0/1 array (uploaded
# once per array object) or empty (a copy)
function mask(u::F64, M::F64)
    y = similar(u)
    check(ccall((:sem_mask, libsemhip), Cint, (Ptr{Cvoid}, Int64, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), length(u), M.ptr, u.ptr, y.ptr))
    y
end
mask(u::F64, M::AbstractArray) = isempty(M) ? copy(u) : mask(u, _leaf(M))

# ABu(As, Br, u), ABu.jl:9-37 (As/Br host matrices, [] = identity); trans
# applies As', Br' (the pullback, ABu.jl:39-41)
function _abu(As, Br, u::F64, trans::Bool)
    m, n = size(u)
    dev(A) = isempty(A) ? (nothing, 0, 0) : (HIPField(Matrix{Float64}(A)), size(A)...)
    Ad, ma, na = dev(As)
    Bd, mb, nb = dev(Br)
    rnb, rmb = trans ? (mb, nb) : (nb, mb)
    rna, rma = trans ? (ma, na) : (na, ma)
    m1 = Bd === nothing ? m : (m ÷ rnb) * rmb
    n1 = Ad === nothing ? n : (n ÷ rna) * rma
    y = HIPField{Float64}(undef, m1, n1)
    check(ccall((:sem_abu, libsemhip), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Cint, Ptr{Cdouble}, Cint,
                 Cint, Cint, Ptr{Cdouble}), ctx(), m, n, u.ptr, cp(Ad), ma, na, cp(Bd), mb, nb,
                trans ? 1 : 0, y.ptr))
    y
end
ABu(As, Br, u::F64) = _abu(As, Br, u, false)

function _gprod(G11, G12, G22, a::F64, b::F64)
    wr, ws = similar(a), similar(a)
    g = map(_leaf, (G11, G12, G22))
    check(ccall((:sem_gprod, libsemhip), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                 Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), length(a), g[1].ptr, g[2].ptr, g[3].ptr, a.ptr, b.ptr, wr.ptr, ws.ptr))
    wr, ws
end

# lapl(u, M, Jr, Js, QQtx, QQty, Dr, Ds, G11, G12, G22, mult), lapl.jl:54-68:
# ur = Dr u, us = u Ds' (interpolated by Jr, Js unless empty, lapl.jl:83-103),
# the G product, the transposed contractions, QQ' and the mask. (The Zygote
# hook scaling the cotangent by mult is the identity in the forward pass.)
function lapl(u::F64, M, Jr, Js, QQtx, QQty, Dr, Ds, G11, G12, G22, mult)
    ur = ABu([], Dr, u)
    us = ABu(Ds, [], u)
    deal = !(isempty(Jr) && isempty(Js))
    if deal
        ur, us = ABu(Js, Jr, ur), ABu(Js, Jr, us)
    end
    vr, vs = _gprod(G11, G12, G22, ur, us)
    if deal
        vr, vs = _abu(Js, Jr, vr, true), _abu(Js, Jr, vs, true)
    end
    Au = _abu([], Dr, vr, true) + _abu(Ds, [], vs, true)
    mask(gatherScatter(Au, QQtx, QQty), M)
end

# mass(u, M, B, Jr, Js, QQtx, QQty, mult), mass.jl:32-50
function mass(u::F64, M, B, Jr, Js, QQtx, QQty, mult)
    Ju = ABu(Js, Jr, u)
    BJu = isempty(B) ? Ju : mask(Ju, B)
    Bu = _abu(Js, Jr, BJu, true)
    mask(gatherScatter(Bu, QQtx, QQty), M)
end

# grad(u, msh), grad.jl:15-34
function grad(u::F64, msh::Mesh)
    ux, uy = similar(u), similar(u)
    check(ccall((:sem_grad, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, u.ptr, ux.ptr, uy.ptr))
    ux, uy
end

# advect(T, vx, vy, mshV, mshD[, Jr, Js]), advect.jl:45-78: the dealiased
# convective term; Jr, Js are interpMat(mshD.zr, mshV.zr) and (zs) as the
# reference builds them, so the library forms them from the meshes
function advect(T::F64, vx, vy, mshV::Mesh, mshD::Mesh)
    y = similar(T)
    check(ccall((:sem_advect, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                 Ptr{Cdouble}), ctx(), devmesh(mshV).h, devmesh(mshD).h, T.ptr, _leaf(vx).ptr,
                _leaf(vy).ptr, y.ptr))
    y
end
advect(T::F64, vx, vy, mshV::Mesh, mshD::Mesh, Jr, Js) = advect(T, vx, vy, mshV, mshD)

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
# as one recognised callable; pcg dispatches it to the fused solver. nodal:
# the parameter of a linsolve through this operator is a continuous nodal
# field (its gradient is summed over the copies of each node).
mutable struct SEMOperator
    h::Ptr{Cvoid}
    msh::Mesh
    keep::Vector{Any}
    nodal::Bool
end
function SEMOperator(ν, k, msh::Mesh, bc::Vector{Char}; precision = SEM_F64, G = nothing,
                     nodal = true)
    νs, νf = ν isa Number ? (Float64(ν), nothing) : (0.0, _leaf(ν))
    ks, kf = k isa Number ? (Float64(k), nothing) : (0.0, _leaf(k))
    g = G === nothing ? (C_NULL, C_NULL, C_NULL) : map(x -> _leaf(x).ptr, G)
    d = OpDesc(νs, ks, cp(νf), cp(kf), g..., C_NULL, maskbits(bc, msh), 1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sem_operator_create, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ref{OpDesc}, Cint, Ref{Ptr{Cvoid}}),
                ctx(), devmesh(msh).h, d, precision, h))
    op = SEMOperator(h[], msh, Any[G, νf, kf], nodal)
    finalizer(x -> ccall((:sem_operator_destroy, libsemhip), Cint, (Ptr{Cvoid},), x.h), op)
    op
end
function (op::SEMOperator)(u::F64)
    y = similar(u)
    check(ccall((:sem_operator_apply, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                ctx(), op.h, u.ptr, y.ptr))
    y
end
Base.:*(op::SEMOperator, u::F64) = op(u)
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

# z = d .* r with the assembled Jacobi inverse (continuous), or with any
# diagonal that may differ between the copies of a node (local = true, as
# u ./ B ./ b0 of convectionDiffusion.jl:87-91)
struct JacobiDiag
    d::F64
    local_::Bool
end
JacobiDiag(d::F64) = JacobiDiag(d, false)
(p::JacobiDiag)(r::F64) = _pw!(similar(r), SEM_PW_MUL, 0.0, p.d, r)
Base.:*(p::JacobiDiag, r::F64) = p(r)

# the fused solver weights its inner products with the mesh's mult; other
# weights (the reference's default ones(size(b))) take the generic pcg
_meshmult(mult, msh::Mesh) = mult === msh.mult

# pcg(b, opA; opM, mult, ifv, tol, maxiter), pcg.jl:16-60: fused when opA is
# an SEMOperator, opM the identity (the default here) or a JacobiDiag and
# mult the mesh's own; otherwise the reference's own pcg runs over the
# HIPField algebra (with its default mult = ones, an unweighted recurrence).
function pcg(b::F64, opA::SEMOperator; opM = identity, mult = ones(size(b)), ifv = false,
             tol = 1e-8, maxiter = length(b))
    if !((opM === identity || opM isa JacobiDiag) && _meshmult(mult, opA.msh))
        return invoke(pcg, Tuple{Any,Any}, b, opA; opM = opM, mult = mult, ifv = ifv, tol = tol,
                      maxiter = maxiter)
    end
    x = similar(b)
    info = Ref(PcgInfo())
    pre, d = opM isa JacobiDiag ?
             (opM.local_ ? SEM_PRECOND_DIAG_LOCAL : SEM_PRECOND_DIAG, opM.d.ptr) :
             (SEM_PRECOND_IDENTITY, C_NULL)
    st = ccall((:sem_pcg, libsemhip), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint,
                Ref{PcgInfo}), ctx(), opA.h, pre, d, b.ptr, x.ptr, tol, min(maxiter, typemax(Cint)), info)
    check(st)
    st == SEM_WARN_MAXITER && println("warning: res:", info[].res)
    ifv && println("PCG iter: ", info[].iters, ", res: ", info[].res)
    x
end
function pcg!(x::F64, b::F64, opA::SEMOperator; kw...)
    _copy!(x, pcg(b, opA; kw...))
    nothing
end

# multigrid V-cycle as opM of the fused solver (a stronger preconditioner
# than the identity / Jacobi of diffusion.jl:47-49)
struct MGParams
    degree::Cint
    ratio::Cdouble
    degree_coarse::Cint
    npow::Cint
    safety::Cdouble
    max_levels::Cint
    precision::Cint
end
mutable struct MGPrecond
    h::Ptr{Cvoid}
    op::SEMOperator
end
function MGPrecond(op::SEMOperator; degree = 3, ratio = 20.0, degree_coarse = 8, npow = 20,
                   safety = 1.1, max_levels = 32, precision = SEM_F64)
    p = MGParams(degree, ratio, degree_coarse, npow, safety, max_levels, precision)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sem_mg_create, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ref{MGParams}, Ref{Ptr{Cvoid}}), ctx(), op.h, p, h))
    mg = MGPrecond(h[], op)
    finalizer(x -> ccall((:sem_mg_destroy, libsemhip), Cint, (Ptr{Cvoid},), x.h), mg)
    mg
end
function (mg::MGPrecond)(r::F64)
    z = similar(r)
    check(ccall((:sem_mg_vcycle, libsemhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), mg.h, r.ptr, z.ptr))
    z
end
Base.:*(mg::MGPrecond, r::F64) = mg(r)
function pcg(b::F64, opA::SEMOperator, opM::MGPrecond; mult = opA.msh.mult, tol = 1e-8,
             maxiter = length(b))
    _meshmult(mult, opA.msh) || error("multigrid PCG weights with the mesh's mult")
    x = similar(b)
    info = Ref(PcgInfo())
    st = ccall((:sem_pcg_mg, libsemhip), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint,
                Ref{PcgInfo}), ctx(), opA.h, opM.h, b.ptr, x.ptr, tol, min(maxiter, typemax(Cint)), info)
    check(st)
    st == SEM_WARN_MAXITER && println("warning: res:", info[].res)
    x
end

# FP64 defect correction around a low-precision fused PCG (SEM_F32 or an
# MFMA form of the same operator): norm(b - op64 x, Inf) <= tol
function pcg_mixed(b::F64, op64::SEMOperator, oplo::SEMOperator; d = nothing, tol = 1e-10,
                   inner_tol = 1e-2, max_outer = 200, max_inner = 1_000_000)
    x = similar(b)
    info = Ref(PcgInfo())
    st = ccall((:sem_pcg_mixed, libsemhip), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble,
                Cdouble, Cint, Cint, Ref{PcgInfo}), ctx(), op64.h, oplo.h, cp(d), b.ptr, x.ptr, tol,
               inner_tol, max_outer, max_inner, info)
    check(st)
    st == SEM_WARN_MAXITER && println("warning: res:", info[].res)
    st == SEM_WARN_STALLED && println("warning: stalled at res:", info[].res)
    x
end

# ------------------------------------------------------ equations on device
# Field (mesh.jl:179-195) with device u, uh, ub; the mask stays a host array
# (mask(u, M) uploads it once)
struct HIPFieldSet
    u::F64
    uh::Vector{F64}
    ub::F64
    M::Array{Float64,2}
    msh::Mesh
end
function HIPFieldSet(bc::Vector{Char}, msh::Mesh; k = 3)
    z() = HIPField(zero(msh.x))
    HIPFieldSet(z(), [z() for _ in 1:k], z(), generateMask(bc, msh), msh)
end
# updateHist!(fld), mesh.jl:199-215
function updateHist!(fld::HIPFieldSet)
    for i in length(fld.uh):-1:2
        fld.uh[i] .= fld.uh[i-1]
    end
    fld.uh[1] .= fld.u
    nothing
end

# a setter(array, x, y, t) of the reference writes a host array; its result
# is uploaded into the device field
function _set!(setter, u::F64, msh::Mesh, t)
    setter === fixU! && return
    a = Array(u)
    setter(a, msh.x, msh.y, t)
    upload!(u, a)
end

# Diffusion (diffusion.jl:7-34) with device fields
struct HIPDiffusion
    fld::HIPFieldSet
    ν::F64
    f::F64
    rhs::F64
    tstep::TimeStepper
    msh::Mesh
end
function HIPDiffusion(bc::Vector{Char}, msh::Mesh; Ti = 0.0, Tf = 0.0, dt = 0.0, k = 3)
    fld = HIPFieldSet(bc, msh)
    HIPDiffusion(fld, zero(fld.u), zero(fld.u), zero(fld.u), TimeStepper(Ti, Tf, dt, k), msh)
end

# opLHS(u, dfn), diffusion.jl:36-45
function opLHS(u::F64, dfn::HIPDiffusion)
    lhs = hlmz(u, dfn.ν, dfn.tstep.bdfB[1], dfn.msh)
    lhs .= gatherScatter(lhs, dfn.msh)
    lhs .= mask(lhs, dfn.fld.M)
    lhs
end
# opPrecond(u, dfn), diffusion.jl:47-49
opPrecond(u::F64, dfn::HIPDiffusion) = u

# makeRHS!(dfn), diffusion.jl:51-65
function makeRHS!(dfn::HIPDiffusion)
    fld, rhs, ν, msh, b = dfn.fld, dfn.rhs, dfn.ν, dfn.msh, dfn.tstep.bdfB
    rhs .= mass(dfn.f, msh)
    rhs .-= ν .* lapl(fld.ub, msh)
    for i in 1:length(fld.uh)
        rhs .-= b[1+i] .* mass(fld.uh[i], msh)
    end
    rhs .= mask(rhs, fld.M)
    rhs .= gatherScatter(rhs, msh)
    nothing
end

# solve!(dfn), diffusion.jl:67-77: the reference's pcg! with closures (the
# generic path); fused = true uses the fused solver on the same operator
function solve!(dfn::HIPDiffusion; fused = false)
    fld, msh = dfn.fld, dfn.msh
    if fused
        op = SEMOperator(dfn.ν, dfn.tstep.bdfB[1], msh, _bc_of(fld.M, msh))
        pcg!(fld.u, dfn.rhs, op; mult = msh.mult)
    else
        opL(u) = opLHS(u, dfn)
        opP(u) = opPrecond(u, dfn)
        pcg!(fld.u, dfn.rhs, opL; opM = opP, mult = msh.mult, ifv = false)
    end
    fld.u .+= fld.ub
    nothing
end

# evolve!(dfn, setBC!, setForcing!, setVisc!), diffusion.jl:81-106
function evolve!(dfn::HIPDiffusion, setBC! = fixU!, setForcing! = fixU!, setVisc! = fixU!;
                 fused = false)
    ts, msh = dfn.tstep, dfn.msh
    updateHist!(dfn.fld)
    updateHist!(ts)
    t = ts.time[1]
    _set!(setBC!, dfn.fld.ub, msh, t)
    _set!(setForcing!, dfn.f, msh, t)
    _set!(setVisc!, dfn.ν, msh, t)
    makeRHS!(dfn)
    solve!(dfn; fused = fused)
    nothing
end

# simulate!(dfn, callback!, setIC!, setBC!, setForcing!, setVisc!), diffusion.jl:110-137
function simulate!(dfn::HIPDiffusion, callback!::Function, setIC! = fixU!, setBC! = fixU!,
                   setForcing! = fixU!, setVisc! = fixU!; fused = false)
    ts = dfn.tstep
    _set!(setIC!, dfn.fld.u, dfn.msh, ts.time[1])
    callback!(dfn)
    while ts.time[1] <= ts.Tf[1]
        evolve!(dfn, setBC!, setForcing!, setVisc!; fused = fused)
        callback!(dfn)
        ts.time[1] < 1e-12 && break
    end
    nothing
end

# the four-side bc characters of a structured mask (for the fused operator)
function _bc_of(M::AbstractMatrix, msh::Mesh)
    for bc in Iterators.product("DN", "DN", "DN", "DN")
        c = collect(bc)
        generateMask(c, msh) == M && return c
    end
    error("the mask is not one of generateMask's")
end

# ConvectionDiffusion (convectionDiffusion.jl:5-56) with device fields
struct HIPConvectionDiffusion
    name::String
    fld::HIPFieldSet
    vx::F64
    vy::F64
    ν::F64
    f::F64
    rhs::F64
    exH::Vector{F64}
    tstep::TimeStepper
    mshV::Mesh
    mshD::Mesh
    set0!::Function
    set∂!::Function
    setF!::Function
    setν!::Function
end
function HIPConvectionDiffusion(name::String, fld::HIPFieldSet, vx, vy, tstep::TimeStepper,
                                mshD::Mesh, set0! = fixU!, set∂! = fixU!, setF! = fixU!,
                                setν! = fixU!)
    HIPConvectionDiffusion(name, fld, _leaf(vx), _leaf(vy), zero(fld.u), zero(fld.u), zero(fld.u),
                           [zero(fld.u) for _ in fld.uh], tstep, fld.msh, mshD, set0!, set∂!,
                           setF!, setν!)
end

# opLHS(u, cdn), convectionDiffusion.jl:76-85
function opLHS(u::F64, cdn::HIPConvectionDiffusion)
    Au = hlmz(u, cdn.ν, cdn.tstep.bdfB[1], cdn.mshV)
    Au .= gatherScatter(Au, cdn.mshV)
    Au .= mask(Au, cdn.fld.M)
    Au
end
# opPrecond(u, cdn), convectionDiffusion.jl:87-91: per-copy B, so the
# iterates are discontinuous (the generic pcg forms <u, Au> on assembled Au)
opPrecond(u::F64, cdn::HIPConvectionDiffusion) = u ./ _leaf(cdn.mshV.B) ./ cdn.tstep.bdfB[1]

# makeRHS!(cdn), convectionDiffusion.jl:93-110
function makeRHS!(cdn::HIPConvectionDiffusion)
    fld, rhs, mshV, a, b = cdn.fld, cdn.rhs, cdn.mshV, cdn.tstep.bdfA, cdn.tstep.bdfB
    rhs .= mass(cdn.f, mshV)
    rhs .-= cdn.ν .* lapl(fld.ub, mshV)
    for i in 1:length(fld.uh)
        cdn.exH[i] .= -advect(fld.uh[i], cdn.vx, cdn.vy, mshV, cdn.mshD)
        rhs .-= b[1+i] .* mass(fld.uh[i], mshV)
        rhs .+= a[i] .* cdn.exH[i]
    end
    rhs .= mask(rhs, fld.M)
    rhs .= gatherScatter(rhs, mshV)
    nothing
end

# solve!(cdn), convectionDiffusion.jl:112-122
function solve!(cdn::HIPConvectionDiffusion)
    fld, mshV = cdn.fld, cdn.mshV
    opL(u) = opLHS(u, cdn)
    opM(u) = opPrecond(u, cdn)
    pcg!(fld.u, cdn.rhs, opL; opM = opM, mult = mshV.mult, ifv = false)
    fld.u .+= fld.ub
    nothing
end

updateHist!(cdn::HIPConvectionDiffusion) = updateHist!(cdn.fld)

# evolve!(cdn), convectionDiffusion.jl:131-144
function evolve!(cdn::HIPConvectionDiffusion)
    t, m = cdn.tstep.time[1], cdn.mshV
    _set!(cdn.set∂!, cdn.fld.ub, m, t)
    _set!(cdn.setF!, cdn.f, m, t)
    _set!(cdn.setν!, cdn.ν, m, t)
    makeRHS!(cdn)
    solve!(cdn)
    nothing
end

# step!(cdn), convectionDiffusion.jl:148-155
function step!(cdn::HIPConvectionDiffusion)
    updateHist!(cdn)
    updateHist!(cdn.tstep)
    evolve!(cdn)
    nothing
end

# simulate!(cdn, callback!), convectionDiffusion.jl:159-179
function simulate!(cdn::HIPConvectionDiffusion, callback!::Function)
    ts = cdn.tstep
    _set!(cdn.set0!, cdn.fld.u, cdn.mshV, ts.time[1])
    callback!(cdn)
    while ts.time[1] <= ts.Tf[1]
        step!(cdn)
        callback!(cdn)
        ts.time[1] < 1e-12 && break
    end
    nothing
end

# ------------------------------------------------------------ adjoint
# The rrule of linsolve (adjoint.jl:10-25) for problem(ν) = (SEMOperator with
# ν folded into G, rhs): λ = solver(lhs, ū, true); the pullback of
# resi(u, ν) = rhs - lhs(ν) u at λ is -B .* (∇λ·∇u) per local copy, summed
# over the copies of each node (QQ') when ν is a continuous nodal field.
function adjoint_dnu(u::F64, λ::F64, msh::Mesh)
    g = similar(u)
    check(ccall((:sem_adjoint_dnu, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, u.ptr, λ.ptr, g.ptr))
    g
end

function vardiff_factors(ν::F64, msh::Mesh)
    G = ntuple(_ -> similar(ν), 3)
    check(ccall((:sem_vardiff_factors, libsemhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                ctx(), devmesh(msh).h, ν.ptr, G[1].ptr, G[2].ptr, G[3].ptr))
    G
end

function ChainRulesCore.rrule(::typeof(linsolve), a::F64, problem, solver)
    lhs, rhs = problem(a)
    u = solver(lhs, rhs, false)
    function linsolve_pullback(ū)
        λ = solver(lhs, unthunk(ū), true)
        g = -adjoint_dnu(u, λ, lhs.msh)           # d<λ, resi(u, a)>/da per copy
        lhs.nodal && (g = gatherScatter(g, lhs.msh))
        return NoTangent(), g, NoTangent(), NoTangent()
    end
    return u, linsolve_pullback
end

export HIPField, upload!, SEMOperator, JacobiDiag, MGPrecond, jacobi, pcg_mixed, GSMap,
       HIPFieldSet, HIPDiffusion, HIPConvectionDiffusion, adjoint_dnu, vardiff_factors

end # module
