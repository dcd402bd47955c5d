# KernelMatricesHIP: ccall shim that puts libkmhodlr (HIP, gfx950) under the
# KernelMatrices.jl HODLR API.
#
# Mirrors (KernelMatrices.jl 0.9.0):
#   KernelMatrix(pts, pts2, parms, kernel)              src/structstypes.jl:17-23
#   HODLR.FixedLevel / HODLR.LogLevel                    src/HODLR/src/structstypes.jl:30-37
#   HODLR.KernelHODLR(K, ep, maxrank, lvl; nystrom, plel, sorted)
#                                                        src/HODLR/src/constructor.jl:6-58
#   HODLR.symmetricfactorize!(HK; plel, verbose)         src/HODLR/src/symfact.jl:6-50
#   HK*v, HK\v, HK\M, logdet(HK), full(HK)               src/HODLR/src/baseoverloads.jl:152-326
#   mul!/ldiv!(t, HK.W or adjoint(HK.W), v)              src/HODLR/src/baseoverloads.jl:95-150
#   HODLR.stoch_gradient(K, HK, y, dfuns, vecs)          src/HODLR/src/stochasticgradient.jl:30-45
#
# A Julia kernel closure cannot cross the C ABI, so kernels are callable
# structs (MaternKernel{ν}, RationalKernel, CauchyKernel): valid Julia kernel
# functions k(x, y, p) for the reference CPU path, and descriptors for the
# GPU path.  Errors are rethrown with the reference's exception types and
# messages (ErrorException, PosDefException).
module KernelMatricesHIP

using LinearAlgebra

export KernelMatrix, MaternKernel, RationalKernel, CauchyKernel, HODLR

const libkmhodlr = get(ENV, "KMHODLR_LIB",
                       joinpath(@__DIR__, "..", "kmhodlr", "libkmhodlr.so"))

# ---- C structs (include/kmhodlr.h) ----
struct kmh_kernel
  family::Int32
  deriv::Int32
  params::NTuple{4,Float64}
  nugget::Float64
end

struct kmh_build_opts
  method::Int32
  level_kind::Int32
  level::Int32
  maxrank::Int32
  tol::Float64
  precision::Int32
  reserved::NTuple{3,Int32}
end

mutable struct kmh_info
  n::Int64
  levels::Int32
  nleaves::Int32
  factorized::Int32
  method::Int32
  device_bytes::Int64
end
kmh_info() = kmh_info(0, 0, 0, 0, 0, 0)

const KMH_E_NOT_PD = 2
const KMH_W, KMH_WT, KMH_WINV, KMH_WTINV = 0, 1, 2, 3

function check(rc::Integer)
  rc == 0 && return nothing
  msg = unsafe_string(ccall((:kmh_last_error, libkmhodlr), Cstring, ()))
  rc == KMH_E_NOT_PD && throw(PosDefException(-1))
  throw(ErrorException(msg))
end

# ---- kernels: callable structs ----
abstract type GPUKernel <: Function end

struct MaternKernel{N} <: GPUKernel   # N = 2ν, ν ∈ {1/2, 3/2}
  nugget::Float64
  deriv::Int32
end
MaternKernel(nu::Real; nugget=0.0) = MaternKernel{Int(2nu)}(nugget, 0)
struct RationalKernel <: GPUKernel; nugget::Float64; end   # README.md:42-48
struct CauchyKernel   <: GPUKernel; nugget::Float64; end   # example/basicexample.jl:62-68

family(::MaternKernel{1}) = Int32(3)
family(::MaternKernel{3}) = Int32(4)
family(::RationalKernel) = Int32(1)
family(::CauchyKernel)   = Int32(2)
deriv(k::MaternKernel) = k.deriv
deriv(::GPUKernel) = Int32(0)

# derivative kernels with respect to parms[1] (t0) and parms[2] (rho), and
# the second derivatives (src/covariancefunctions.jl:268-330)
d1(k::MaternKernel{N}) where {N} = MaternKernel{N}(0.0, 1)
d2(k::MaternKernel{N}) where {N} = MaternKernel{N}(0.0, 2)
d12(k::MaternKernel{N}) where {N} = MaternKernel{N}(0.0, 12)
d22(k::MaternKernel{N}) where {N} = MaternKernel{N}(0.0, 22)

_cor(::Val{1}, x) = x <= 1e-8 ? 1.0 : exp(-x)
_cor(::Val{3}, x) = x <= 1e-8 ? 1.0 : (1 + sqrt(3)x) * exp(-sqrt(3)x)
function (k::MaternKernel{N})(x, y, p) where {N}
  r = norm(x - y)
  k.deriv == 0 || error("evaluate derivative kernels on the GPU path")
  out = r > 0 ? p[1] * _cor(Val(N), r / p[2]) : p[1]
  return x == y ? out + k.nugget : out
end
(k::RationalKernel)(x, y, p) = abs2(p[1]) / abs2(1 + abs2(norm(x - y) / p[2])) + (x == y ? k.nugget : 0.0)
(k::CauchyKernel)(x, y, p)   = abs2(p[1]) / (1 + abs2(p[2] * norm(x - y))) + (x == y ? k.nugget : 0.0)

descriptor(k::GPUKernel, parms) =
  kmh_kernel(family(k), deriv(k), ntuple(i -> i <= length(parms) ? Float64(parms[i]) : 0.0, 4),
             k isa MaternKernel ? k.nugget : (k.deriv == 0 ? k.nugget : 0.0))
descriptor(k::Union{RationalKernel,CauchyKernel}, parms) =
  kmh_kernel(family(k), 0, ntuple(i -> i <= length(parms) ? Float64(parms[i]) : 0.0, 4), k.nugget)
# family 0: the reference's ZeroFunction in stoch_hessian's second-derivative list
zero_descriptor() = kmh_kernel(Int32(0), Int32(0), (0.0, 0.0, 0.0, 0.0), 0.0)

# ---- context ----
mutable struct Context
  ptr::Ptr{Cvoid}
end
function Context(device::Integer=0; hbm_budget::Integer=0)
  p = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:kmh_ctx_create, libkmhodlr), Cint, (Cint, UInt64, Ptr{Ptr{Cvoid}}), device, hbm_budget, p))
  c = Context(p[])
  finalizer(c -> ccall((:kmh_ctx_destroy, libkmhodlr), Cint, (Ptr{Cvoid},), c.ptr), c)
  return c
end
const DEFAULT = Ref{Union{Nothing,Context}}(nothing)
default_context() = (DEFAULT[] === nothing && (DEFAULT[] = Context(0)); DEFAULT[])

# ---- KernelMatrix ----
struct KernelMatrix{T,N,A,Fn}
  x1::AbstractVector{A}
  x2::AbstractVector{A}
  parms::NTuple{N,Float64}
  kernel::Fn
end
KernelMatrix(pts, pts2, prms, fn) = KernelMatrix{Float64,length(prms),eltype(pts),typeof(fn)}(pts, pts2, Tuple(Float64.(prms)), fn)
Base.size(K::KernelMatrix) = (length(K.x1), length(K.x2))

module HODLR
  using LinearAlgebra
  import ..KernelMatricesHIP: check, libkmhodlr, kmh_build_opts, kmh_info, descriptor, zero_descriptor, default_context,
                              KernelMatrix, KMH_W, KMH_WT, KMH_WINV, KMH_WTINV
  import LinearAlgebra: logdet, mul!, ldiv!

  abstract type HierLevel end
  struct FixedLevel <: HierLevel; lv::Int64; end
  struct LogLevel   <: HierLevel; lv::Int64; end
  kind(::FixedLevel) = Int32(0)
  kind(::LogLevel)   = Int32(1)

  mutable struct KernelHODLR{T<:Number}
    ptr::Ptr{Cvoid}
    n::Int
    nys::Bool
    mrnk::Int
  end

  struct FactorHODLR
    HK::KernelHODLR
  end

  function points_matrix(x)
    P = Matrix{Float64}(undef, 2, length(x))       # row-major n x 2 == column-major 2 x n
    for (i, p) in enumerate(x); P[1, i] = p[1]; P[2, i] = p[2]; end
    return P
  end

  function KernelHODLR(K::KernelMatrix, ep::Float64, maxrank::Int64, lvl::HierLevel;
                       nystrom::Bool=false, plel::Bool=false, sorted=false, ctx=default_context())
    K.x1 == K.x2 || error("This function builds symmetric matrices. For non-symmetric matrices,
                 please use the recursive RKernelHODLR, which will only allow matvecs.")
    sorted || @warn "No default point sorting is done for you, and if your points are
    not sorted properly the approximation can be very poor." maxlog=1
    kd = Ref(descriptor(K.kernel, K.parms))
    op = Ref(kmh_build_opts(nystrom ? 1 : 0, kind(lvl), lvl.lv, maxrank, ep, 0, (0, 0, 0)))
    P = points_matrix(K.x1)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:kmh_hodlr_build, libkmhodlr), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
                ctx.ptr, P, length(K.x1), 2, kd, op, h))
    HK = KernelHODLR{Float64}(h[], length(K.x1), nystrom, maxrank)
    finalizer(x -> ccall((:kmh_hodlr_destroy, libkmhodlr), Cint, (Ptr{Cvoid},), x.ptr), HK)
    return HK
  end

  function info(HK::KernelHODLR)
    i = Ref(kmh_info())
    check(ccall((:kmh_hodlr_info, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{kmh_info}), HK.ptr, i))
    return i[]
  end

  # Callers test `HK.U == nothing` as "is factorized" (stochasticgradient.jl:35)
  function Base.getproperty(HK::KernelHODLR, s::Symbol)
    s === :U && return info(HK).factorized == 1 ? nothing : :gpu_resident
    s === :V && return info(HK).factorized == 1 ? nothing : :gpu_resident
    s === :W && return info(HK).factorized == 1 ? FactorHODLR(HK) : nothing
    return getfield(HK, s)
  end
  Base.size(HK::KernelHODLR) = (HK.n, HK.n)

  function symmetricfactorize!(HK::KernelHODLR; plel::Bool=false, verbose::Bool=false)
    check(ccall((:kmh_hodlr_factorize, libkmhodlr), Cint, (Ptr{Cvoid},), HK.ptr))
    return nothing
  end

  function mul!(t::StridedVecOrMat{Float64}, HK::KernelHODLR, s::StridedVecOrMat{Float64})
    check(ccall((:kmh_hodlr_matvec, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64),
                HK.ptr, s, t, size(s, 2)))
    return t
  end
  Base.:*(HK::KernelHODLR, s::VecOrMat{Float64}) = mul!(similar(s), HK, s)

  function ldiv!(t::StridedVecOrMat{Float64}, HK::KernelHODLR, s::StridedVecOrMat{Float64})
    check(ccall((:kmh_hodlr_solve, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64),
                HK.ptr, s, t, size(s, 2)))
    return t
  end
  Base.:\(HK::KernelHODLR, s::VecOrMat{Float64}) = ldiv!(similar(s), HK, s)

  function logdet(HK::KernelHODLR)
    out = Ref(0.0)
    check(ccall((:kmh_hodlr_logdet, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}), HK.ptr, out))
    return out[]
  end

  function full(HK::KernelHODLR)
    out = Matrix{Float64}(undef, HK.n, HK.n)
    check(ccall((:kmh_hodlr_full, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}), HK.ptr, out))
    return out
  end

  function factor_apply!(t, W::FactorHODLR, op::Integer, s)
    check(ccall((:kmh_hodlr_factor_apply, libkmhodlr), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Int64), W.HK.ptr, op, s, t, size(s, 2)))
    return t
  end
  mul!(t::StridedVector{Float64}, W::FactorHODLR, s::StridedVector{Float64}) = factor_apply!(t, W, KMH_W, s)
  mul!(t::StridedVector{Float64}, W::Adjoint{<:Any,FactorHODLR}, s::StridedVector{Float64}) =
    factor_apply!(t, W.parent, KMH_WT, s)
  ldiv!(t::StridedVector{Float64}, W::FactorHODLR, s::StridedVector{Float64}) = factor_apply!(t, W, KMH_WINV, s)
  ldiv!(t::StridedVector{Float64}, W::Adjoint{<:Any,FactorHODLR}, s::StridedVector{Float64}) =
    factor_apply!(t, W.parent, KMH_WTINV, s)
  LinearAlgebra.adjoint(W::FactorHODLR) = Adjoint{Float64,FactorHODLR}(W)

  function stoch_gradient(K::KernelMatrix, HK::KernelHODLR, dat::Vector{Float64},
                          dfuns::Vector, vecs::Vector{Vector{Float64}};
                          plel::Bool=false, verbose::Bool=false, shuffle::Bool=false)
    HK.U == nothing                  || error("The matrix needs to be factorized for this to work.")
    length(dfuns) == length(K.parms) || error("You didn't supply the right number of gradient funs.")
    dk = [descriptor(f, K.parms) for f in dfuns]
    saa = reduce(hcat, vecs)
    g = zeros(length(dfuns))
    check(ccall((:kmh_stoch_gradient, libkmhodlr), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}),
                HK.ptr, dk, length(dk), dat, saa, length(vecs), g))
    return g
  end

  # stochasticgradient.jl:48-86
  function stoch_profile_gradient(K::KernelMatrix, HK::KernelHODLR, dat::Vector{Float64},
                                  dfuns::Vector, vecs::Vector{Vector{Float64}};
                                  plel::Bool=false, verbose::Bool=false, shuffle::Bool=false)
    HK.U == nothing                  || error("The matrix needs to be factorized for this to work.")
    length(dfuns) == length(K.parms) || error("You didn't supply the right number of gradient funs.")
    dk = [descriptor(f, K.parms) for f in dfuns]
    saa = reduce(hcat, vecs)
    g = zeros(length(dfuns))
    check(ccall((:kmh_stoch_profile_gradient, libkmhodlr), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}),
                HK.ptr, dk, length(dk), dat, saa, length(vecs), g))
    return g
  end

  # stochastichessian.jl:91-115 (derivative HODLRs given by their kernels)
  function stoch_fisher(K::KernelMatrix, HK::KernelHODLR, dfuns::Vector, vecs::Vector{Vector{Float64}};
                        plel::Bool=false, shuffle::Bool=false)
    HK.U == nothing || error("The matrix needs to be factorized.")
    dk = [descriptor(f, K.parms) for f in dfuns]
    saa = reduce(hcat, vecs)
    F = zeros(length(dk), length(dk))
    check(ccall((:kmh_stoch_fisher, libkmhodlr), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Float64}),
                HK.ptr, dk, length(dk), saa, length(vecs), F))
    return Symmetric(F)
  end

  # stochastichessian.jl:117-147; d2funs[j][k-j] is the (j,k) kernel or nothing (ZeroFunction)
  function stoch_hessian(K::KernelMatrix, HK::KernelHODLR, dat::Vector{Float64}, d1funs::Vector,
                         d2funs::Vector, vecs::Vector{Vector{Float64}}; plel::Bool=false,
                         verbose::Bool=false, shuffle::Bool=false, profile::Bool=false,
                         exact_second::Bool=false)
    profile && error("the profile-likelihood Hessian is not provided")
    HK.U == nothing || error("The matrix needs to be factorized.")
    d1 = [descriptor(f, K.parms) for f in d1funs]
    d2 = [f === nothing ? zero_descriptor() : descriptor(f, K.parms) for row in d2funs for f in row]
    saa = reduce(hcat, vecs)
    Hm = zeros(length(d1), length(d1))
    check(ccall((:kmh_stoch_hessian, libkmhodlr), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Int32, Ptr{Float64}),
                HK.ptr, d1, length(d1), d2, dat, saa, length(vecs), Int32(exact_second), Hm))
    return Symmetric(Hm)
  end

  negloglik(HK::KernelHODLR, dat::Vector{Float64}) = 0.5 * logdet(HK) + 0.5 * dot(dat, HK \ dat)
  # maxlikfunctions.jl:11-19
  profile_negloglik(HK::KernelHODLR, dat::Vector{Float64}) =
    0.5 * (logdet(HK) + length(dat) * log(dot(dat, HK \ dat)))

  # maxlikdata(; pts, data, sortmethod=:KDTree)  constructor.jl:171-181
  function kd_order(pts::Vector; leafsize::Integer=8)
    d = length(first(pts))
    P = Matrix{Float64}(undef, d, length(pts))        # row-major n x d for the C side
    for (i, x) in enumerate(pts); P[:, i] .= x; end
    perm = Vector{Int64}(undef, length(pts))
    check(ccall((:kmh_kd_order, libkmhodlr), Cint, (Ptr{Float64}, Int64, Int32, Int32, Ptr{Int64}),
                P, length(pts), d, leafsize, perm))
    return perm .+ 1
  end

  # HODLR.RKernelHODLR(K, tol, maxrank=0, lvl=LogLevel(7))  constructor.jl:130-161:
  # recursive HODLR of K(x1, x2) for matvecs (x1 and x2 may differ).
  mutable struct RKernelHODLR
    ptr::Ptr{Cvoid}
    n1::Int
    n2::Int
  end

  function RKernelHODLR(K::KernelMatrix, tol::Float64, maxrank::Int64=0, lvl::HierLevel=LogLevel(7);
                        ctx=default_context())
    kd = Ref(descriptor(K.kernel, K.parms))
    op = Ref(kmh_build_opts(0, kind(lvl), lvl.lv, maxrank, tol, 0, (0, 0, 0)))
    P1, P2 = points_matrix(K.x1), points_matrix(K.x2)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:kmh_rhodlr_build, libkmhodlr), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Int32, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Ptr{Cvoid}}),
                ctx.ptr, P1, length(K.x1), P2, length(K.x2), 2, kd, op, h))
    R = RKernelHODLR(h[], length(K.x1), length(K.x2))
    finalizer(x -> ccall((:kmh_rhodlr_destroy, libkmhodlr), Cint, (Ptr{Cvoid},), x.ptr), R)
    return R
  end

  Base.size(R::RKernelHODLR) = (R.n1, R.n2)
  Base.size(R::RKernelHODLR, j::Int64) = size(R)[j]

  # baseoverloads.jl:299-303
  function Base.:*(R::RKernelHODLR, V::VecOrMat{Float64})
    size(V, 1) == R.n2 || throw(DimensionMismatch("matvec with a vector of the wrong length"))
    out = V isa Vector ? Vector{Float64}(undef, R.n1) : Matrix{Float64}(undef, R.n1, size(V, 2))
    check(ccall((:kmh_rhodlr_matvec, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64),
                R.ptr, V, out, size(V, 2)))
    return out
  end
end # module HODLR

end # module KernelMatricesHIP
