# KernelMatricesHIP: puts libkmhodlr (HIP, gfx950) under KernelMatrices.jl's
# own HODLR API.  It adds methods to the reference's functions and types; it
# defines no KernelMatrix or HODLR module of its own, so callers keep
# `using KernelMatrices` and their calls, and pick the GPU path by the type of
# the kernel function they put in the KernelMatrix.
#
# Methods added (KernelMatrices.jl 0.9.0):
#   HODLR.KernelHODLR(K, ep, maxrank, lvl; nystrom, plel, sorted)
#                                                        src/HODLR/src/constructor.jl:6-58
#   HODLR.symmetricfactorize!(HK; plel, verbose)         src/HODLR/src/symfact.jl:6-50
#   HK*v, HK\v, HK\M, logdet(HK), full(HK), size(HK)     src/HODLR/src/baseoverloads.jl:2-11,152-326
#   mul!/ldiv!(t, HK.W or HK.W', v)                      src/HODLR/src/baseoverloads.jl:95-150
#   HODLR.DerivativeHODLR(K, dfun, HK)                   src/HODLR/src/constructor.jl:66-95
#   HODLR.stoch_gradient / stoch_profile_gradient        src/HODLR/src/stochasticgradient.jl:30-86
#   HODLR.stoch_fisher / stoch_hessian                   src/HODLR/src/stochastichessian.jl:91-147
#   HODLR.negloglik / profile_negloglik                  src/HODLR/src/maxlikfunctions.jl:2-19
#   HODLR.RKernelHODLR(K, tol, maxrank, lvl), R*v        src/HODLR/src/constructor.jl:130-161,
#                                                        src/HODLR/src/baseoverloads.jl:299-303
#
# Every method dispatches on K::KernelMatrix{T,N,A,<:GPUKernel} or on a type
# defined here, so the reference's own methods are untouched.  A Julia
# closure cannot cross the C ABI, so the GPU kernels are callable structs
# (MaternKernel{ν}, RationalKernel, CauchyKernel): valid kernel functions
# k(x, y, p) for the reference's KernelMatrix (and its CPU path), and kernel
# descriptors for the library.  Errors are rethrown with the reference's
# exception types and messages (ErrorException, PosDefException).
module KernelMatricesHIP

using LinearAlgebra
import KernelMatrices
import KernelMatrices: KernelMatrix, full
import KernelMatrices.HODLR
import KernelMatrices.HODLR: HierLevel, FixedLevel, LogLevel
import LinearAlgebra: logdet, mul!, ldiv!

export MaternKernel, RationalKernel, CauchyKernel, d1, d2, d12, d22, kd_order, HIPContext

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
const PRECISION = Dict(:fp64 => Int32(0), :mixed => Int32(1), :mixed_q => Int32(2))

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
             k isa MaternKernel ? k.nugget : 0.0)
descriptor(k::Union{RationalKernel,CauchyKernel}, parms) =
  kmh_kernel(family(k), 0, ntuple(i -> i <= length(parms) ? Float64(parms[i]) : 0.0, 4), k.nugget)
# family 0: the reference's ZeroFunction in stoch_hessian's second-derivative list
zero_descriptor() = kmh_kernel(Int32(0), Int32(0), (0.0, 0.0, 0.0, 0.0), 0.0)
descriptor(::HODLR.ZeroFunction, parms) = zero_descriptor()
descriptor(f::Function, parms) =
  error("the GPU path takes kernels of type MaternKernel, RationalKernel or CauchyKernel, not $(typeof(f))")

# ---- context ----
mutable struct HIPContext
  ptr::Ptr{Cvoid}
end
function HIPContext(device::Integer=0; hbm_budget::Integer=0)
  p = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:kmh_ctx_create, libkmhodlr), Cint, (Cint, UInt64, Ptr{Ptr{Cvoid}}), device, hbm_budget, p))
  c = HIPContext(p[])
  finalizer(c -> ccall((:kmh_ctx_destroy, libkmhodlr), Cint, (Ptr{Cvoid},), c.ptr), c)
  return c
end
const DEFAULT = Ref{Union{Nothing,HIPContext}}(nothing)
default_context() = (DEFAULT[] === nothing && (DEFAULT[] = HIPContext(0)); DEFAULT[])

kind(::FixedLevel) = Int32(0)
kind(::LogLevel)   = Int32(1)

# row-major n x d for the C side (column-major d x n)
function points_matrix(x)
  d = length(first(x))
  P = Matrix{Float64}(undef, d, length(x))
  for (i, p) in enumerate(x); P[:, i] .= p; end
  return P
end

# ---- the GPU-resident HODLR matrix ----
mutable struct HIPHODLR{T<:Number}
  ptr::Ptr{Cvoid}
  n::Int
  nys::Bool
  mrnk::Int
  ctx::HIPContext
end

# the symmetric factor W (K = W W'), as the reference's FactorHODLR
struct HIPFactor{T} <: LinearAlgebra.Factorization{T}
  HK::HIPHODLR{T}
end
struct HIPFactorAdjoint{T}
  W::HIPFactor{T}
end
LinearAlgebra.adjoint(W::HIPFactor) = HIPFactorAdjoint(W)
Base.size(W::HIPFactor) = (W.HK.n, W.HK.n)

function info(HK::HIPHODLR)
  i = Ref(kmh_info())
  check(ccall((:kmh_hodlr_info, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{kmh_info}), HK.ptr, i))
  return i[]
end

# Callers test `HK.U == nothing` as "is factorized" (stochasticgradient.jl:35)
function Base.getproperty(HK::HIPHODLR, s::Symbol)
  s === :U && return info(HK).factorized == 1 ? nothing : :gpu_resident
  s === :V && return info(HK).factorized == 1 ? nothing : :gpu_resident
  s === :W && return info(HK).factorized == 1 ? HIPFactor(HK) : nothing
  return getfield(HK, s)
end
Base.size(HK::HIPHODLR) = (HK.n, HK.n)
Base.size(HK::HIPHODLR, j::Int64) = size(HK)[j]

# constructor.jl:6-58, for a KernelMatrix whose kernel is a GPU kernel.
# Extra keywords: ctx (a HIPContext) and precision (:fp64, :mixed, :mixed_q).
function HODLR.KernelHODLR(K::KernelMatrix{T,N,A,Fn}, ep::Float64, maxrank::Int64, lvl::HierLevel;
                           nystrom::Bool=false, plel::Bool=false, sorted=false,
                           ctx::HIPContext=default_context(),
                           precision::Symbol=:fp64) where {T<:Number,N,A,Fn<:GPUKernel}
  K.x1 == K.x2 || error("This function builds symmetric matrices. For non-symmetric matrices,
               please use the recursive RKernelHODLR, which will only allow matvecs.")
  sorted || @warn "No default point sorting is done for you, and if your points are
  not sorted properly the approximation can be very poor." maxlog=1
  kd = Ref(descriptor(K.kernel, K.parms))
  op = Ref(kmh_build_opts(nystrom ? 1 : 0, kind(lvl), lvl.lv, maxrank, ep, PRECISION[precision], (0, 0, 0)))
  P = points_matrix(K.x1)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:kmh_hodlr_build, libkmhodlr), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Int64, Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
              ctx.ptr, P, length(K.x1), size(P, 1), kd, op, h))
  HK = HIPHODLR{T}(h[], length(K.x1), nystrom, maxrank, ctx)
  finalizer(x -> ccall((:kmh_hodlr_destroy, libkmhodlr), Cint, (Ptr{Cvoid},), x.ptr), HK)
  return HK
end

function HODLR.symmetricfactorize!(HK::HIPHODLR; plel::Bool=false, verbose::Bool=false)
  check(ccall((:kmh_hodlr_factorize, libkmhodlr), Cint, (Ptr{Cvoid},), HK.ptr))
  return nothing
end

function mul!(t::StridedVecOrMat{Float64}, HK::HIPHODLR, s::StridedVecOrMat{Float64})
  check(ccall((:kmh_hodlr_matvec, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64),
              HK.ptr, s, t, size(s, 2)))
  return t
end
Base.:*(HK::HIPHODLR, s::VecOrMat{Float64}) = mul!(similar(s), HK, s)

function ldiv!(t::StridedVecOrMat{Float64}, HK::HIPHODLR, s::StridedVecOrMat{Float64})
  check(ccall((:kmh_hodlr_solve, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64),
              HK.ptr, s, t, size(s, 2)))
  return t
end
Base.:\(HK::HIPHODLR, s::VecOrMat{Float64}) = ldiv!(similar(s), HK, s)

function logdet(HK::HIPHODLR)
  out = Ref(0.0)
  check(ccall((:kmh_hodlr_logdet, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}), HK.ptr, out))
  return out[]
end

function full(HK::HIPHODLR)
  out = Matrix{Float64}(undef, HK.n, HK.n)
  check(ccall((:kmh_hodlr_full, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}), HK.ptr, out))
  return out
end

function factor_apply!(t, W::HIPFactor, op::Integer, s)
  check(ccall((:kmh_hodlr_factor_apply, libkmhodlr), Cint,
              (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Int64), W.HK.ptr, op, s, t, size(s, 2)))
  return t
end
mul!(t::StridedVector{Float64}, W::HIPFactor, s::StridedVector{Float64}) = factor_apply!(t, W, KMH_W, s)
mul!(t::StridedVector{Float64}, W::HIPFactorAdjoint, s::StridedVector{Float64}) = factor_apply!(t, W.W, KMH_WT, s)
ldiv!(t::StridedVector{Float64}, W::HIPFactor, s::StridedVector{Float64}) = factor_apply!(t, W, KMH_WINV, s)
ldiv!(t::StridedVector{Float64}, W::HIPFactorAdjoint, s::StridedVector{Float64}) = factor_apply!(t, W.W, KMH_WTINV, s)
Base.:*(W::HIPFactor, s::Vector{Float64}) = mul!(similar(s), W, s)
Base.:*(W::HIPFactorAdjoint, s::Vector{Float64}) = mul!(similar(s), W, s)
Base.:\(W::HIPFactor, s::Vector{Float64}) = ldiv!(similar(s), W, s)
Base.:\(W::HIPFactorAdjoint, s::Vector{Float64}) = ldiv!(similar(s), W, s)

# constructor.jl:66-95: the derivative of the Nystrom HODLR matrix with
# respect to one parameter; on the GPU path it is the derivative kernel,
# evaluated inside the library's gradient, Fisher and Hessian routines.
struct HIPDerivativeHODLR{T}
  kernel::kmh_kernel
  HK::HIPHODLR{T}
end
function HODLR.DerivativeHODLR(K::KernelMatrix{T,N,A,Fn}, dfun::Function, HK::HIPHODLR{T};
                               plel::Bool=false) where {T<:Number,N,A,Fn<:GPUKernel}
  HK.nys || error("This is only valid for Nystrom-block matrices.")
  return HIPDerivativeHODLR{T}(descriptor(dfun, K.parms), HK)
end

saa_matrix(vecs) = reduce(hcat, vecs)

# stochasticgradient.jl:30-45
function HODLR.stoch_gradient(K::KernelMatrix{T,N,A,Fn}, HK::HIPHODLR{T}, dat::Vector{T},
                              dfuns::Vector, vecs::Vector{Vector{T}};
                              plel::Bool=false, verbose::Bool=false,
                              shuffle::Bool=false) where {T<:Number,N,A,Fn<:GPUKernel}
  HK.U == nothing                  || error("The matrix needs to be factorized for this to work.")
  length(dfuns) == length(K.parms) || error("You didn't supply the right number of gradient funs.")
  dk = [descriptor(f, K.parms) for f in dfuns]
  g = zeros(length(dfuns))
  check(ccall((:kmh_stoch_gradient, libkmhodlr), Cint,
              (Ptr{Cvoid}, Ptr{kmh_kernel}, Int32, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}),
              HK.ptr, dk, length(dk), dat, saa_matrix(vecs), length(vecs), g))
  return g
end

# stochasticgradient.jl:68-86
function HODLR.stoch_profile_gradient(K::KernelMatrix{T,N,A,Fn}, HK::HIPHODLR{T}, dat::Vector{T},
                                      dfuns::Vector, vecs::Vector{Vector{T}};
                                      plel::Bool=false, verbose::Bool=false,
                                      shuffle::Bool=false) where {T<:Number,N,A,Fn<:GPUKernel}
  HK.U == nothing                  || error("The matrix needs to be factorized for this to work.")
  length(dfuns) == length(K.parms) || error("You didn't supply the right number of gradient funs.")
  dk = [descriptor(f, K.parms) for f in dfuns]
  g = zeros(length(dfuns))
  check(ccall((:kmh_stoch_profile_gradient, libkmhodlr), Cint,
              (Ptr{Cvoid}, Ptr{kmh_kernel}, Int32, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}),
              HK.ptr, dk, length(dk), dat, saa_matrix(vecs), length(vecs), g))
  return g
end

# stochastichessian.jl:91-115
function HODLR.stoch_fisher(K::KernelMatrix{T,N,A,Fn}, HK::HIPHODLR{T},
                            DKs::Vector{HIPDerivativeHODLR{T}}, vecs::Vector{Vector{T}};
                            plel::Bool=false, shuffle::Bool=false) where {T<:Number,N,A,Fn<:GPUKernel}
  HK.U == nothing || error("The matrix needs to be factorized.")
  dk = [D.kernel for D in DKs]
  F = zeros(length(dk), length(dk))
  check(ccall((:kmh_stoch_fisher, libkmhodlr), Cint,
              (Ptr{Cvoid}, Ptr{kmh_kernel}, Int32, Ptr{Float64}, Int64, Ptr{Float64}),
              HK.ptr, dk, length(dk), saa_matrix(vecs), length(vecs), F))
  return Symmetric(F)
end

# stochastichessian.jl:117-147; d2funs[j][k-j+1] is the (j,k) second-derivative
# kernel or the reference's ZeroFunction.  exact_second=true uses the exact
# second derivative of the Nystrom blocks instead of the reference's terms.
function HODLR.stoch_hessian(K::KernelMatrix{T,N,A,Fn}, HK::HIPHODLR{T}, dat::Vector{T},
                             d1funs::Vector, d2funs::Vector, vecs::Vector{Vector{T}};
                             plel::Bool=false, verbose::Bool=false, shuffle::Bool=false,
                             profile::Bool=false, exact_second::Bool=false) where {T<:Number,N,A,Fn<:GPUKernel}
  profile && error("the profile-likelihood Hessian is not provided")
  HK.U == nothing || error("The matrix needs to be factorized.")
  d1 = [descriptor(f, K.parms) for f in d1funs]
  d2 = [descriptor(f, K.parms) for row in d2funs for f in row]
  Hm = zeros(length(d1), length(d1))
  check(ccall((:kmh_stoch_hessian, libkmhodlr), Cint,
              (Ptr{Cvoid}, Ptr{kmh_kernel}, Int32, Ptr{kmh_kernel}, Ptr{Float64}, Ptr{Float64}, Int64, Int32,
               Ptr{Float64}),
              HK.ptr, d1, length(d1), d2, dat, saa_matrix(vecs), length(vecs), Int32(exact_second), Hm))
  return Symmetric(Hm)
end

# maxlikfunctions.jl:2-19
function HODLR.negloglik(HK::HIPHODLR{T}, dat::Vector{T}) where {T<:Number}
  HK.U == nothing || error("The HODLR matrix needs to be factorized for this.")
  return 0.5 * logdet(HK) + 0.5 * dot(dat, HK \ dat)
end
function HODLR.profile_negloglik(HK::HIPHODLR{T}, dat::Vector{T}) where {T<:Number}
  HK.U == nothing || error("The HODLR matrix needs to be factorized for this.")
  return 0.5 * (logdet(HK) + length(dat) * log(dot(dat, HK \ dat)))
end

# The KD ordering the reference takes from NearestNeighbors.KDTree(pts).data
# (maxlikdata, constructor.jl:171-181), computed by the library; returns the
# 1-based permutation.
function kd_order(pts::AbstractVector; leafsize::Integer=8)
  P = points_matrix(pts)
  perm = Vector{Int64}(undef, length(pts))
  check(ccall((:kmh_kd_order, libkmhodlr), Cint, (Ptr{Float64}, Int64, Int32, Int32, Ptr{Int64}),
              P, length(pts), size(P, 1), leafsize, perm))
  return perm .+ 1
end

# constructor.jl:130-161: recursive HODLR of K(x1, x2) for matvecs (x1 and
# x2 may differ), for a KernelMatrix whose kernel is a GPU kernel.
mutable struct HIPRKernelHODLR
  ptr::Ptr{Cvoid}
  n1::Int
  n2::Int
  ctx::HIPContext
end

function HODLR.RKernelHODLR(K::KernelMatrix{T,N,A,Fn}, tol::Float64, maxrank::Int64=0,
                            lvl::HierLevel=LogLevel(7); ctx::HIPContext=default_context()) where {T,N,A,Fn<:GPUKernel}
  kd = Ref(descriptor(K.kernel, K.parms))
  op = Ref(kmh_build_opts(0, kind(lvl), lvl.lv, maxrank, tol, 0, (0, 0, 0)))
  P1, P2 = points_matrix(K.x1), points_matrix(K.x2)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:kmh_rhodlr_build, libkmhodlr), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Int32, Ptr{Cvoid}, Ptr{Cvoid},
               Ptr{Ptr{Cvoid}}),
              ctx.ptr, P1, length(K.x1), P2, length(K.x2), size(P1, 1), kd, op, h))
  R = HIPRKernelHODLR(h[], length(K.x1), length(K.x2), ctx)
  finalizer(x -> ccall((:kmh_rhodlr_destroy, libkmhodlr), Cint, (Ptr{Cvoid},), x.ptr), R)
  return R
end

Base.size(R::HIPRKernelHODLR) = (R.n1, R.n2)
Base.size(R::HIPRKernelHODLR, j::Int64) = size(R)[j]

# baseoverloads.jl:299-303
function Base.:*(R::HIPRKernelHODLR, V::VecOrMat{Float64})
  size(V, 1) == R.n2 || throw(DimensionMismatch("matvec with a vector of the wrong length"))
  out = V isa Vector ? Vector{Float64}(undef, R.n1) : Matrix{Float64}(undef, R.n1, size(V, 2))
  check(ccall((:kmh_rhodlr_matvec, libkmhodlr), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64),
              R.ptr, V, out, size(V, 2)))
  return out
end

end # module KernelMatricesHIP
