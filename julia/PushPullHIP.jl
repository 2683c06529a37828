"""
    PushPullHIP

Julia front end of `libpushpull_hip.so` (declared in `include/pushpull_hip.h`) for arrays
that live on an MI355X. It mirrors the reference's Julia surface for the pull/push path —
`pull`, `pull_grad`, `pull_hess`, `push`, `push_grad` (src/pushpull_cpu.jl:17-259,
ext/pushpull_gpu.jl:15-259), `affine_pull` / `affine_push` (src/pushpull_cpu.jl:268-362),
`compose`, `compose!`, `Exp` (src/compositions.jl:9-49) and the three rrules
(src/pushpull_pb.jl:11-51) — on `DArray`, a device buffer in Julia's column-major layout
(image (X,Y,Z,C,B), deformation (X,Y,Z,3,B) holding 1-based voxel coordinates).

It needs no GPU package: device memory comes from `pp_malloc` / `pp_memcpy_*`, every call
is one `ccall` of a Tier B entry point on a caller stream (default: the null stream).
`Settings` is the reference's (src/pushpull_general.jl:15-60); any value with its fields
`deg`, `bnd`, `ext` (PushPull.Settings included) is accepted.

Host arrays keep working through the reference package itself: `drop_in!(PushPull)` points
its `pplib` / `oplib` (src/pushpull_init.jl:4-5) at this library, whose Tier A symbols
have the reference's signatures (artifacts/C/pushpull.h:3-14).

The library path is `ENV["PUSHPULL_HIP_LIB"]`, else `libpushpull_hip.so` on the loader path.
"""
module PushPullHIP

using Libdl

export Settings, DArray, todevice, tohost, free!, id,
       pull, pull_grad, pull_hess, push, push_grad, affine_pull, affine_push, adjust_affine,
       pull_backward, pull_grad_backward, pull_vjp, objective_grad, objective_grad_step!,
       compose, compose!, Exp,
       Volume, BatchPlan, run!, run_static!, destroy!, Graph, launch, capture, drop_in!

const lib = Ref{Ptr{Cvoid}}(C_NULL)
const PP_MAX_CHAN_BND = 64
const PP_BATCH_CHAN_BND = 10

function __init__()
    lib[] = Libdl.dlopen(get(ENV, "PUSHPULL_HIP_LIB", "libpushpull_hip.so"))
    return nothing
end

sym(name::Symbol) = Libdl.dlsym(lib[], name)

function check(st::Integer)
    st == 0 && return nothing
    msg = unsafe_string(ccall(sym(:pp_last_error), Cstring, ()))
    error("pushpull_hip: $msg (status $st)")
end

# ---- Settings (src/pushpull_general.jl:15-60) -----------------------------------------
struct Settings
    deg::NTuple{3,UInt64}
    bnd::Vector{NTuple{3,Int32}}
    ext::Int32
end

function Settings(deg::Union{Integer,NTuple{3,<:Integer}}=(1, 1, 1),
                  bnd::Union{Integer,NTuple{3,<:Integer},Vector{<:NTuple{3,Integer}},Matrix{<:Integer}}=1,
                  ext::Integer=true)
    deg isa Integer && (deg = (deg, deg, deg))
    deg = UInt64.(max.(min.(deg, 4), 0))
    allowed = (0, 1, 2)
    if bnd isa Integer
        @assert bnd ∈ allowed
        b = [Int32.((bnd, bnd, bnd))]
    elseif bnd isa Matrix
        @assert size(bnd, 1) == 3 && all(bnd .∈ (allowed,))
        b = [Int32.((bnd[:, i]...,)) for i in 1:size(bnd, 2)]
    elseif bnd isa Vector
        @assert all(all(t .∈ (allowed,)) for t in bnd)
        b = [Int32.(t) for t in bnd]
    else
        @assert all(bnd .∈ (allowed,))
        b = [Int32.(bnd)]
    end
    return Settings(deg, b, Int32(max(min(ext, 1), 0)))
end

# ---- the call descriptor (include/pushpull_hip.h: pp_desc) ----------------------------
struct Desc
    d0::NTuple{3,Int64}
    d1::NTuple{3,Int64}
    nchan::Int32
    nbatch::Int32
    deg::NTuple{3,Int32}
    ext::Int32
    nbnd::Int32
    bnd::NTuple{3 * PP_MAX_CHAN_BND,Int32}
end

function Desc(sett, d0, d1, C::Integer, B::Integer)
    nb = length(sett.bnd)
    @assert nb == 1 || nb == C "length(sett.bnd) must be 1 or the number of channels"
    @assert nb <= PP_MAX_CHAN_BND
    b = zeros(Int32, 3 * PP_MAX_CHAN_BND)
    for (c, t) in enumerate(sett.bnd), q in 1:3
        b[3(c - 1) + q] = t[q]
    end
    return Desc(Int64.(Tuple(d0)), Int64.(Tuple(d1)), Int32(C), Int32(B), Int32.(Tuple(sett.deg)),
                Int32(sett.ext), Int32(nb), Tuple(b))
end

# ---- device buffers ------------------------------------------------------------------
"""
    DArray(dims...)

`Float32` device buffer of `dims` (column-major, as a Julia `Array`), allocated with
`pp_malloc` and freed by the finalizer or `free!`.
"""
mutable struct DArray{N}
    ptr::Ptr{Float32}
    dims::NTuple{N,Int}
    function DArray(dims::NTuple{N,Integer}) where {N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall(sym(:pp_malloc), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, 4 * max(prod(dims), 1)))
        a = new{N}(Ptr{Float32}(p[]), Int.(dims))
        finalizer(free!, a)
        return a
    end
end
DArray(dims::Integer...) = DArray(dims)

Base.size(a::DArray) = a.dims
Base.size(a::DArray, i::Integer) = i <= length(a.dims) ? a.dims[i] : 1
Base.ndims(a::DArray) = length(a.dims)
Base.length(a::DArray) = prod(a.dims)

function free!(a::DArray)
    if a.ptr != C_NULL
        check(ccall(sym(:pp_free), Cint, (Ptr{Cvoid},), a.ptr))
        a.ptr = C_NULL
    end
    return nothing
end

function todevice(x::Array{Float32}; stream::Ptr{Cvoid}=C_NULL)
    a = DArray(size(x))
    GC.@preserve x check(ccall(sym(:pp_memcpy_h2d), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                               a.ptr, pointer(x), 4 * length(x), stream))
    return a
end

function tohost(a::DArray; stream::Ptr{Cvoid}=C_NULL)
    x = Array{Float32}(undef, a.dims)
    GC.@preserve x check(ccall(sym(:pp_memcpy_d2h), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                               pointer(x), a.ptr, 4 * length(x), stream))
    return x
end

function dzeros(dims; stream::Ptr{Cvoid}=C_NULL)
    a = DArray(Tuple(dims))
    check(ccall(sym(:pp_memset_zero), Cint, (Ptr{Float32}, Csize_t, Ptr{Cvoid}), a.ptr, length(a), stream))
    return a
end

"""
    id(d; stream) -> DArray

The 1-based identity deformation (X,Y,Z,3) of grid `d` (src/pushpull_general.jl `id`),
built on the host and uploaded.
"""
function id(d::NTuple{3,Integer}; stream::Ptr{Cvoid}=C_NULL)
    x = Array{Float32}(undef, d..., 3)
    for k in 1:d[3], j in 1:d[2], i in 1:d[1]
        x[i, j, k, 1] = i
        x[i, j, k, 2] = j
        x[i, j, k, 3] = k
    end
    return todevice(x; stream=stream)
end

nchan(a::DArray) = size(a, 4)
nbatch(a::DArray) = size(a, 5)
spatial(a::DArray) = (size(a, 1), size(a, 2), size(a, 3))

# The reference's argument checks (ext/pushpull_gpu.jl:19-23)
function check_pull_args(f₀::DArray, ϕ::DArray, sett)
    @assert 3 <= ndims(f₀) <= 5 "`f₀` must have between 3 & 5 dimensions"
    @assert 4 <= ndims(ϕ) <= 5 "`ϕ` must have 4 or 5 dimensions"
    @assert size(ϕ, 4) == 3 "`ϕ` must have three channels"
    @assert size(ϕ, 5) == size(f₀, 5) "`ϕ` & `f₀` must have the same batch size"
    @assert length(sett.bnd) == size(f₀, 4) || length(sett.bnd) == 1
end

# ---- the five ops (src/pushpull_cpu.jl:17-259; ext/pushpull_gpu.jl:15-259) -------------
function pull_family(name::Symbol, comp::Tuple, f₀::DArray, ϕ::DArray, sett, stream)
    check_pull_args(f₀, ϕ, sett)
    d1 = spatial(ϕ)
    out = DArray((d1..., comp..., size(f₀)[4:end]...))
    d = Ref(Desc(sett, spatial(f₀), d1, nchan(f₀), nbatch(f₀)))
    check(ccall(sym(name), Cint, (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Cvoid}),
                d, out.ptr, ϕ.ptr, f₀.ptr, stream))
    return out
end

"""pull(f₀, ϕ, sett) (src/pushpull_cpu.jl:17): out (d₁..., C, B)"""
pull(f₀::DArray, ϕ::DArray, sett=Settings(); stream::Ptr{Cvoid}=C_NULL) =
    pull_family(:pp_pull, (), f₀, ϕ, sett, stream)
"""pull_grad(f₀, ϕ, sett) (src/pushpull_cpu.jl:69): out (d₁..., 3, C, B)"""
pull_grad(f₀::DArray, ϕ::DArray, sett=Settings(); stream::Ptr{Cvoid}=C_NULL) =
    pull_family(:pp_pull_grad, (3,), f₀, ϕ, sett, stream)
"""pull_hess(f₀, ϕ, sett) (src/pushpull_cpu.jl:111): out (d₁..., 3, 3, C, B)"""
pull_hess(f₀::DArray, ϕ::DArray, sett=Settings(); stream::Ptr{Cvoid}=C_NULL) =
    pull_family(:pp_pull_hess, (3, 3), f₀, ϕ, sett, stream)

function push_family(name::Symbol, ncomp::Int, f₁::DArray, ϕ::DArray, d₀, sett, stream)
    @assert 4 <= ndims(ϕ) <= 5 "`ϕ` must have 4 or 5 dimensions"
    @assert size(ϕ, 4) == 3 "`ϕ` must have three channels"
    d1 = spatial(ϕ)
    @assert spatial(f₁) == d1 "`f₁` & `ϕ` must have the same volume dimensions"
    lead = ncomp == 1 ? 4 : 5  # push_grad's values are (d₁..., 3, C, B)
    ncomp == 3 && @assert size(f₁, 4) == 3 "push_grad takes (d₁..., 3, C, B) values"
    C, B = size(f₁, lead), size(f₁, lead + 1)
    @assert size(ϕ, 5) == B "`ϕ` & `f₁` must have the same batch size"
    out = DArray((Int.(Tuple(d₀))..., size(f₁)[lead:end]...))
    d = Ref(Desc(sett, d₀, d1, C, B))
    check(ccall(sym(name), Cint, (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Cint, Ptr{Cvoid}),
                d, out.ptr, ϕ.ptr, f₁.ptr, 0, stream))
    return out
end

"""push(f₁, ϕ, d₀, sett) (src/pushpull_cpu.jl:165): out (d₀..., C, B)"""
push(f₁::DArray, ϕ::DArray, d₀::NTuple{3,Integer}, sett=Settings(); stream::Ptr{Cvoid}=C_NULL) =
    push_family(:pp_push, 1, f₁, ϕ, d₀, sett, stream)
"""push_grad(g₁, ϕ, d₀, sett) (src/pushpull_cpu.jl:220): the adjoint of pull_grad"""
push_grad(g₁::DArray, ϕ::DArray, d₀::NTuple{3,Integer}, sett=Settings(); stream::Ptr{Cvoid}=C_NULL) =
    push_family(:pp_push_grad, 3, g₁, ϕ, d₀, sett, stream)

"""adjust_affine (src/pushpull_cpu.jl:357-362): the 0-based 3x4 matrix of the C code."""
function adjust_affine(Aff::Matrix{Float32})
    @assert (size(Aff, 1) == 3 || size(Aff, 1) == 4) && size(Aff, 2) == 4
    A = Float32.(Aff[1:3, :])
    A[:, 4] .= sum(A, dims=2) .- 1.0
    return A
end

"""affine_pull(f₀, Aff, d₁, sett) (src/pushpull_cpu.jl:268)"""
function affine_pull(f₀::DArray, Aff::Matrix{Float32}, d₁::NTuple{3,Integer}, sett=Settings();
                     stream::Ptr{Cvoid}=C_NULL)
    @assert 3 <= ndims(f₀) <= 5 "`f₀` must have between 3 & 5 dimensions"
    A = adjust_affine(Aff)
    out = DArray((Int.(d₁)..., size(f₀)[4:end]...))
    d = Ref(Desc(sett, spatial(f₀), d₁, nchan(f₀), nbatch(f₀)))
    GC.@preserve A check(ccall(sym(:pp_affine_pull), Cint,
                               (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Cvoid}),
                               d, out.ptr, f₀.ptr, pointer(A), stream))
    return out
end

"""affine_push(f₁, Aff, d₀, sett) (src/pushpull_cpu.jl:309)"""
function affine_push(f₁::DArray, Aff::Matrix{Float32}, d₀::NTuple{3,Integer}, sett=Settings();
                     stream::Ptr{Cvoid}=C_NULL)
    @assert 3 <= ndims(f₁) <= 5 "`f₁` must have between 3 & 5 dimensions"
    A = adjust_affine(Aff)
    out = DArray((Int.(d₀)..., size(f₁)[4:end]...))
    d = Ref(Desc(sett, d₀, spatial(f₁), nchan(f₁), nbatch(f₁)))
    GC.@preserve A check(ccall(sym(:pp_affine_push), Cint,
                               (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Cint, Ptr{Cvoid}),
                               d, out.ptr, f₁.ptr, pointer(A), 0, stream))
    return out
end

# ---- pullbacks and fused gradients -------------------------------------------------
"""sum_c δ_c ∇pull(f₀)_c: the ϕ-pullback of pull (src/pushpull_pb.jl:13-17)"""
function pull_backward(f₀::DArray, ϕ::DArray, δ::DArray, sett=Settings(); stream::Ptr{Cvoid}=C_NULL)
    check_pull_args(f₀, ϕ, sett)
    @assert size(δ) == (spatial(ϕ)..., size(f₀)[4:end]...) "δ must have the shape of pull(f₀, ϕ)"
    dϕ = DArray(size(ϕ))
    d = Ref(Desc(sett, spatial(f₀), spatial(ϕ), nchan(f₀), nbatch(f₀)))
    check(ccall(sym(:pp_pull_backward), Cint,
                (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Cvoid}),
                d, dϕ.ptr, ϕ.ptr, f₀.ptr, δ.ptr, stream))
    return dϕ
end

"""sum_{a,c} δ_{a,c} H(f₀)_{a,·,c}: the ϕ-pullback of pull_grad (src/pushpull_pb.jl:26-35)"""
function pull_grad_backward(f₀::DArray, ϕ::DArray, δ::DArray, sett=Settings(); stream::Ptr{Cvoid}=C_NULL)
    check_pull_args(f₀, ϕ, sett)
    @assert size(δ) == (spatial(ϕ)..., 3, size(f₀)[4:end]...) "δ must have the shape of pull_grad(f₀, ϕ)"
    dϕ = DArray(size(ϕ))
    d = Ref(Desc(sett, spatial(f₀), spatial(ϕ), nchan(f₀), nbatch(f₀)))
    check(ccall(sym(:pp_pull_grad_backward), Cint,
                (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Cvoid}),
                d, dϕ.ptr, ϕ.ptr, f₀.ptr, δ.ptr, stream))
    return dϕ
end

"""Both cotangents of pull from one launch: (push(δ, ϕ, d₀), pull_backward(f₀, ϕ, δ))."""
function pull_vjp(f₀::DArray, ϕ::DArray, δ::DArray, sett=Settings(); stream::Ptr{Cvoid}=C_NULL)
    check_pull_args(f₀, ϕ, sett)
    @assert size(δ) == (spatial(ϕ)..., size(f₀)[4:end]...) "δ must have the shape of pull(f₀, ϕ)"
    df₀, dϕ = DArray(size(f₀)), DArray(size(ϕ))
    d = Ref(Desc(sett, spatial(f₀), spatial(ϕ), nchan(f₀), nbatch(f₀)))
    check(ccall(sym(:pp_pull_vjp), Cint,
                (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Cint, Ptr{Cvoid}),
                d, dϕ.ptr, df₀.ptr, ϕ.ptr, f₀.ptr, δ.ptr, 0, stream))
    return df₀, dϕ
end

"""L = ½‖pull(f₀, ϕ) − g‖² with ∂L/∂ϕ and ∂L/∂f₀ from one launch: (L, dϕ, df₀)."""
function objective_grad(f₀::DArray, ϕ::DArray, g::DArray, sett=Settings(); stream::Ptr{Cvoid}=C_NULL)
    check_pull_args(f₀, ϕ, sett)
    @assert size(g) == (spatial(ϕ)..., size(f₀)[4:end]...) "g must have the shape of pull(f₀, ϕ)"
    df₀, dϕ = DArray(size(f₀)), DArray(size(ϕ))
    loss = DArray((2,))  # one Float64
    d = Ref(Desc(sett, spatial(f₀), spatial(ϕ), nchan(f₀), nbatch(f₀)))
    check(ccall(sym(:pp_objective_grad), Cint,
                (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float64}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Cint,
                 Ptr{Cvoid}),
                d, dϕ.ptr, df₀.ptr, Ptr{Float64}(loss.ptr), ϕ.ptr, f₀.ptr, g.ptr, 0, stream))
    L = Ref{Float64}(0.0)
    check(ccall(sym(:pp_memcpy_d2h), Cint, (Ref{Float64}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}), L, loss.ptr, 8, stream))
    free!(loss)
    return L[], dϕ, df₀
end

"""
    objective_grad_step!(f₀, ϕ, g, step, sett; ϕnext=ϕ)

`objective_grad` with the descent step ϕnext = ϕ + step·∂L/∂ϕ fused into the gradient pass
(in place by default). Returns (L, ∂L/∂ϕ, ∂L/∂f₀).
"""
function objective_grad_step!(f₀::DArray, ϕ::DArray, g::DArray, step::Real, sett=Settings();
                              ϕnext::DArray=ϕ, stream::Ptr{Cvoid}=C_NULL)
    check_pull_args(f₀, ϕ, sett)
    @assert size(g) == (spatial(ϕ)..., size(f₀)[4:end]...) "g must have the shape of pull(f₀, ϕ)"
    @assert size(ϕnext) == size(ϕ) "`ϕnext` must have the shape of ϕ"
    df₀, dϕ = DArray(size(f₀)), DArray(size(ϕ))
    loss = DArray((2,))  # one Float64
    d = Ref(Desc(sett, spatial(f₀), spatial(ϕ), nchan(f₀), nbatch(f₀)))
    check(ccall(sym(:pp_objective_grad_step), Cint,
                (Ref{Desc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float64}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                 Ptr{Float32}, Cfloat, Cint, Ptr{Cvoid}),
                d, dϕ.ptr, df₀.ptr, Ptr{Float64}(loss.ptr), ϕ.ptr, f₀.ptr, g.ptr, ϕnext.ptr, Float32(step), 0, stream))
    L = Ref{Float64}(0.0)
    check(ccall(sym(:pp_memcpy_d2h), Cint, (Ref{Float64}, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}), L, loss.ptr, 8, stream))
    free!(loss)
    return L[], dϕ, df₀
end

# ---- compositions (src/compositions.jl:9-49) ------------------------------------------
function compose!(ψ::DArray, ϕ::DArray, θ::DArray; stream::Ptr{Cvoid}=C_NULL)
    @assert size(ϕ, 4) == 3 "`ϕ` must have three channels"
    @assert size(θ, 4) == 3 "`θ` must have three channels"
    @assert size(ψ) == size(θ) "`ψ` and `θ` must have the same dimensions"
    @assert size(ϕ, 5) == size(θ, 5)
    check(ccall(sym(:pp_compose), Cint,
                (Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ref{NTuple{3,Int64}}, Ref{NTuple{3,Int64}}, Cint, Ptr{Cvoid}),
                ψ.ptr, ϕ.ptr, θ.ptr, Int64.(spatial(ϕ)), Int64.(spatial(θ)), size(θ, 5), stream))
    return ψ
end

compose(ϕ::DArray, θ::DArray; stream::Ptr{Cvoid}=C_NULL) = compose!(DArray(size(θ)), ϕ, θ; stream=stream)

"""Exp(v): scaling and squaring as the reference writes it (src/compositions.jl:41-49)."""
function Exp(v::DArray; nsq::Integer=8, stream::Ptr{Cvoid}=C_NULL)
    @assert size(v, 4) == 3 "`v` must have three channels"
    out, work = DArray(size(v)), DArray(size(v))
    check(ccall(sym(:pp_exp), Cint,
                (Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ref{NTuple{3,Int64}}, Cint, Cint, Ptr{Cvoid}),
                out.ptr, v.ptr, work.ptr, Int64.(spatial(v)), size(v, 5), nsq, stream))
    free!(work)
    return out
end

# ---- batches of uneven volumes (include/pushpull_hip.h: pp_volume) --------------------
const PP_OP = Dict(:pull => 0, :push => 3, :push_grad => 4, :pull_backward => 5, :objective => 7, :pull_vjp => 8)

struct Volume
    f::Ptr{Float32}
    phi::Ptr{Float32}
    aux::Ptr{Float32}
    out::Ptr{Float32}
    out2::Ptr{Float32}
    loss::Ptr{Float64}
    d0::NTuple{3,Int64}
    d1::NTuple{3,Int64}
    nchan::Int32
    deg::NTuple{3,Int32}
    ext::Int32
    bnd::NTuple{3,Int32}
    nbnd::Int32
    chan_bnd::NTuple{3 * PP_BATCH_CHAN_BND,Int32}
    phi_next::Ptr{Float32}
    step::Float32
end

"""
    Volume(f, ϕ; deg, bnd=(1,1,1), ext=1, aux, out, out2, loss, d₀)

One volume of a batch (`deg` one degree or a degree per axis; `bnd` one triple, or a vector
of triples, one per channel); the DArrays must outlive every run of the plan. For the
objective, `ϕnext` (may be ϕ itself) receives ϕ + `step`·∂L/∂ϕ, a descent step fused into
the gradient pass.
"""
function Volume(f::DArray, ϕ::DArray; deg::Union{Integer,NTuple{3,<:Integer}}, bnd=(1, 1, 1), ext::Integer=1,
                aux=nothing, out::DArray, out2=nothing, loss::Ptr{Float64}=Ptr{Float64}(C_NULL), d₀=spatial(f),
                ϕnext=nothing, step::Real=0)
    p(a) = a === nothing ? Ptr{Float32}(C_NULL) : a.ptr
    C = size(f, 4)
    per = bnd isa AbstractVector
    per && @assert length(bnd) == C <= PP_BATCH_CHAN_BND
    codes = zeros(Int32, 3 * PP_BATCH_CHAN_BND)
    per && for c in 1:C, q in 1:3
        codes[3 * (c - 1) + q] = bnd[c][q]
    end
    Volume(f.ptr, ϕ.ptr, p(aux), out.ptr, p(out2), loss, Int64.(Tuple(d₀)), Int64.(spatial(ϕ)),
           Int32(C), Int32.(deg isa Integer ? (deg, deg, deg) : deg), Int32(ext),
           Int32.(Tuple(per ? bnd[1] : bnd)), Int32(per ? C : 0), Tuple(codes), p(ϕnext), Float32(step))
end

mutable struct BatchPlan
    handle::Ptr{Cvoid}
    vols::Vector{Volume}
end

function BatchPlan(op::Symbol, vols::Vector{Volume})
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall(sym(:pp_batch_plan_create), Cint, (Cint, Ptr{Volume}, Cint, Ref{Ptr{Cvoid}}),
                PP_OP[op], vols, length(vols), h))
    plan = BatchPlan(h[], vols)
    finalizer(destroy!, plan)
    return plan
end

run!(plan::BatchPlan; stream::Ptr{Cvoid}=C_NULL) =
    check(ccall(sym(:pp_batch_run), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), plan.handle, stream))

function destroy!(plan::BatchPlan)
    if plan.handle != C_NULL
        check(ccall(sym(:pp_batch_plan_destroy), Cint, (Ptr{Cvoid},), plan.handle))
        plan.handle = C_NULL
    end
    return nothing
end

"""One batch without a plan or any device allocation (at most 256 volumes per call)."""
run_static!(op::Symbol, vols::Vector{Volume}; stream::Ptr{Cvoid}=C_NULL) =
    check(ccall(sym(:pp_batch_run_static), Cint, (Cint, Ptr{Volume}, Cint, Ptr{Cvoid}),
                PP_OP[op], vols, length(vols), stream))

# ---- HIP graphs ----------------------------------------------------------------------
mutable struct Graph
    exec::Ptr{Cvoid}
end

"""
    capture(f, stream) -> Graph

Record every library call `f()` makes on `stream` (not the null stream) into a HIP graph;
buffers must be allocated before the capture.
"""
function capture(f, stream::Ptr{Cvoid})
    check(ccall(sym(:pp_graph_begin), Cint, (Ptr{Cvoid},), stream))
    exec = Ref{Ptr{Cvoid}}(C_NULL)
    try
        f()
    finally
        check(ccall(sym(:pp_graph_end), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), stream, exec))
    end
    g = Graph(exec[])
    finalizer(g) do x
        x.exec != C_NULL && ccall(sym(:pp_graph_destroy), Cint, (Ptr{Cvoid},), x.exec)
        x.exec = C_NULL
    end
    return g
end

launch(g::Graph, stream::Ptr{Cvoid}) = check(ccall(sym(:pp_graph_launch), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.exec, stream))

# ---- host arrays through the reference package (Tier A) -------------------------------
"""
    drop_in!(PushPull)

Point the reference package's native handles (src/pushpull_init.jl:4-5) at this library:
its `pull`, `push`, ... on host `Array`s then `ccall` the Tier A symbols here unchanged.
"""
function drop_in!(mod::Module)
    Core.eval(mod, :(global pplib = $(lib[])))
    Core.eval(mod, :(global oplib = $(lib[])))
    return nothing
end

# ---- ChainRulesCore (src/pushpull_pb.jl:11-51) ----------------------------------------
@static if Base.find_package("ChainRulesCore") !== nothing
    import ChainRulesCore
    using ChainRulesCore: NoTangent, unthunk, @thunk

    function ChainRulesCore.rrule(::typeof(pull), f₀::DArray, ϕ::DArray, s=Settings())
        f₁ = pull(f₀, ϕ, s)
        function pull_pb(δ)
            δ = unthunk(δ)
            return (NoTangent(), @thunk(push(δ, ϕ, spatial(f₀), s)), @thunk(pull_backward(f₀, ϕ, δ, s)), NoTangent())
        end
        return f₁, pull_pb
    end

    function ChainRulesCore.rrule(::typeof(pull_grad), f₀::DArray, ϕ::DArray, s=Settings())
        ∇f = pull_grad(f₀, ϕ, s)
        function pull_grad_pb(δ)
            δ = unthunk(δ)
            return (NoTangent(), @thunk(push_grad(δ, ϕ, spatial(f₀), s)), @thunk(pull_grad_backward(f₀, ϕ, δ, s)),
                    NoTangent())
        end
        return ∇f, pull_grad_pb
    end

    function ChainRulesCore.rrule(::typeof(push), f₁::DArray, ϕ::DArray, dim, s=Settings())
        f₀ = push(f₁, ϕ, dim, s)
        function push_pb(δ)
            δ = unthunk(δ)
            return (NoTangent(), @thunk(pull(δ, ϕ, s)), @thunk(pull_backward(δ, ϕ, f₁, s)), NoTangent(), NoTangent())
        end
        return f₀, push_pb
    end
end

end # module
