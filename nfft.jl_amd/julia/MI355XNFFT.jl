"""
    MI355XNFFT

Julia binding of libnfftm.so: an `AbstractNFFTPlan` whose transforms run on
one MI355X.  It plugs into AbstractNFFTs exactly where NFFT.jl's own GPU
extension does (ext/NFFTGPUArraysExt/implementation.jl:21-59) and exposes the
same operator surface: plan_nfft, mul!, adjoint, nodes!, size_in/size_out,
convolve!/convolve_transpose!/deconvolve!/deconvolve_transpose!, copy and
TimingStats, and NFFTTools' sdc! and Toeplitz kernel on the device.  Data buffers are AMDGPU `ROCArray`s passed as device pointers;
node matrices stay on the host, as `k::Matrix` does in the reference.

Usage:

    using AMDGPU, AbstractNFFTs, MI355XNFFT
    MI355XNFFT.activate!()
    p = plan_nfft(ROCArray, k, N; m=4, σ=2.0, precompute=AbstractNFFTs.LINEAR)
    fHat = p * ROCArray(f)

This file is written against the C ABI in include/nfftm.h.  No Julia
toolchain exists where the library was built, so it has not been executed;
the same calls are exercised from Python ctypes by the test suite.
"""
module MI355XNFFT

using AbstractNFFTs
using AbstractNFFTs: AbstractNFFTBackend, TimingStats
using LinearAlgebra
using AMDGPU
import ChainRulesCore
import NFFTTools

const libnfftm = get(ENV, "NFFTM_LIB", joinpath(@__DIR__, "..", "nfftm", "libnfftm.so"))

const NFFTM_OK = Cint(0)
const NFFTM_ERR_ARG = Cint(1)
const NFFTM_ERR_DIM = Cint(2)
const NFFTM_ERR_UNSUPPORTED = Cint(3)
const NFFTM_F32 = Cint(0)
const NFFTM_F64 = Cint(1)
const NFFTM_FLAG_BLOCKING = Cint(1)
const NFFTM_FLAG_TIMING = Cint(2)
const NFFTM_FLAG_GRAPHS = Cint(4)
const NFFTM_FLAG_ROCFFT = Cint(8)
const NFFTM_FLAG_SEPARATE_FOLD = Cint(16)
const NFFTM_FLAG_SPARSE = Cint(32)

# Error classes as the reference raises them (src/utils.jl:46-55, :98-105).
function check(rc::Cint)
  rc == NFFTM_OK && return nothing
  msg = unsafe_string(ccall((:nfftm_last_error, libnfftm), Cstring, ()))
  rc == NFFTM_ERR_ARG && throw(ArgumentError(msg))
  rc == NFFTM_ERR_DIM && throw(DimensionMismatch(msg))
  error(msg)
end

struct MI355XBackend <: AbstractNFFTBackend end
activate!() = AbstractNFFTs.set_active_backend!(MI355XNFFT)
backend() = MI355XBackend()

# R = 1, or D - length(dims) + 1 for a directional plan (dims ≠ 1:D), whose
# output puts J in place of the transformed axes (src/precomputation.jl:38-48).
mutable struct MI355XNFFTPlan{T,D,R} <: AbstractNFFTs.AbstractNFFTPlan{T,D,R}
  handle::Ptr{Cvoid}
  N::NTuple{D,Int64}
  NOut::NTuple{R,Int64}
  J::Int64
  k::Matrix{T}
  Ñ::NTuple{D,Int64}                # oversampled along dims only
  dims::UnitRange{Int64}
  tmpVec::ROCArray{Complex{T}}      # wraps the plan-owned grid (no copy); one slice for a directional plan
end

nout(N::NTuple{D,Int}, dims::UnitRange{Int64}, J) where {D} =
  (N[1:first(dims)-1]..., J, N[last(dims)+1:D]...)

precision_of(::Type{Float32}) = NFFTM_F32
precision_of(::Type{Float64}) = NFFTM_F64

# getWindow's names (src/windowFunctions.jl:4-19) and the C ABI's ids
const WINDOW_IDS = Dict(:kaiser_bessel => Cint(0), :gauss => Cint(1), :spline => Cint(2),
                        :kaiser_bessel_rev => Cint(3), :cosh_type => Cint(4))
window_id(w::Symbol) = haskey(WINDOW_IDS, w) ? WINDOW_IDS[w] : error("Window $(w) not yet implemented!")

function wrap_grid(h::Ptr{Cvoid}, ::Type{T}, Ñ::NTuple{D,Int}) where {T,D}
  ptr = Ref{Ptr{Cvoid}}()
  check(ccall((:nfftm_grid, libnfftm), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), h, ptr))
  return unsafe_wrap(ROCArray, Ptr{Complex{T}}(ptr[]), Ñ; own=false)
end

function sizes(h::Ptr{Cvoid}, D)
  Ñ = zeros(Int64, D); M = Ref{Int64}(); σ = Ref{Float64}()
  check(ccall((:nfftm_sizes, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ref{Int64}, Ref{Float64}),
              h, Ñ, M, σ))
  return NTuple{D,Int64}(Ñ), M[]
end

# plan_nfft(::Backend, ::Type{<:ROCArray}, k, N): ext/NFFTGPUArraysExt/implementation.jl:21-59,
# parameters resolved as AbstractNFFTs/src/misc.jl:66-81.
function AbstractNFFTs.plan_nfft(::MI355XBackend, ::Type{<:ROCArray}, k::Matrix{T}, N::NTuple{D,Int};
    dims::Union{Integer,UnitRange{Int64}}=1:D, fftflags=nothing, sortNodes=false,
    blocking=true, window=:kaiser_bessel, precompute=AbstractNFFTs.POLYNOMIAL,
    timing::Union{Nothing,TimingStats}=nothing, graphs=false, rocfft=false, fused_fold=true,
    sparse_matrix=false, kargs...) where {T<:Union{Float32,Float64},D}
  dims_ = dims isa Integer ? (dims:dims) : dims
  (1 <= first(dims_) <= last(dims_) <= D) || throw(ArgumentError("dims = $(dims) is not a range of 1:$(D)"))
  size(k, 1) == length(dims_) ||
    throw(ArgumentError("Nodes x have dimension $(size(k, 1)) != $(length(dims_))"))
  reltol, m, σ = AbstractNFFTs.accuracyParams(; kargs...)
  sortNodes && (k = sortslices(k, dims=2))
  kd = Matrix{Float64}(k)
  h = Ref{Ptr{Cvoid}}(C_NULL)
  flags = (blocking ? NFFTM_FLAG_BLOCKING : Cint(0)) | (timing === nothing ? Cint(0) : NFFTM_FLAG_TIMING) |
          (graphs ? NFFTM_FLAG_GRAPHS : Cint(0)) | (rocfft ? NFFTM_FLAG_ROCFFT : Cint(0)) |
          (fused_fold ? Cint(0) : NFFTM_FLAG_SEPARATE_FOLD) | (sparse_matrix ? NFFTM_FLAG_SPARSE : Cint(0))
  t = @elapsed if dims_ == 1:D
    check(ccall((:nfftm_plan_create, libnfftm), Cint,
      (Ref{Ptr{Cvoid}}, Cint, Cint, Ptr{Int64}, Int64, Ptr{Float64}, Cint, Cdouble, Cint, Cint, Cint, Ptr{Cvoid}),
      h, precision_of(T), D, collect(Int64, N), size(k, 2), kd, m, σ, window_id(window),
      Cint(Int(precompute)), flags, AMDGPU.stream().stream))
  else  # directional (src/directional.jl): slices transformed in place, strided
    check(ccall((:nfftm_plan_create_dims, libnfftm), Cint,
      (Ref{Ptr{Cvoid}}, Cint, Cint, Ptr{Int64}, Cint, Cint, Int64, Ptr{Float64}, Cint, Cdouble, Cint, Cint, Cint,
       Ptr{Cvoid}),
      h, precision_of(T), D, collect(Int64, N), first(dims_) - 1, last(dims_) - 1, size(k, 2), kd, m, σ,
      window_id(window), Cint(Int(precompute)), flags, AMDGPU.stream().stream))
  end
  timing === nothing || (timing.pre = t)
  Ñ, M = sizes(h[], D)
  NOut = nout(N, dims_, Int64(M))
  p = MI355XNFFTPlan{T,D,length(NOut)}(h[], N, NOut, Int64(M), k, Ñ, dims_, wrap_grid(h[], T, Ñ[dims_]))
  finalizer(p) do q   # Wrappers/FINUFFT.jl:52-56
    ccall((:nfftm_plan_destroy, libnfftm), Cvoid, (Ptr{Cvoid},), q.handle)
  end
  return p
end

AbstractNFFTs.size_in(p::MI355XNFFTPlan) = p.N
AbstractNFFTs.size_out(p::MI355XNFFTPlan) = p.NOut

# nodes!(p, k): src/implementation.jl:108-141
function AbstractNFFTs.nodes!(p::MI355XNFFTPlan{T,D}, k::Matrix{T}) where {T,D}
  check(ccall((:nfftm_nodes, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64),
              p.handle, Matrix{Float64}(k), size(k, 2)))
  p.k = k; p.J = size(k, 2); p.NOut = nout(p.N, p.dims, p.J)
  return p
end

function Base.copy(p::MI355XNFFTPlan{T,D,R}) where {T,D,R}   # src/implementation.jl:45-66
  h = Ref{Ptr{Cvoid}}(C_NULL)
  check(ccall((:nfftm_plan_copy, libnfftm), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), p.handle, h))
  q = MI355XNFFTPlan{T,D,R}(h[], p.N, p.NOut, p.J, copy(p.k), p.Ñ, p.dims, wrap_grid(h[], T, p.Ñ[p.dims]))
  finalizer(x -> ccall((:nfftm_plan_destroy, libnfftm), Cvoid, (Ptr{Cvoid},), x.handle), q)
  return q
end

devptr(a::ROCArray) = Ptr{Cvoid}(pointer(a))

# src/utils.jl:98-105
function consistency(p::MI355XNFFTPlan{T,D}, f::ROCArray, fHat::ROCArray) where {T,D}
  size(f) == p.N || throw(DimensionMismatch("Data f has size $(size(f)) but the plan expects $(p.N)"))
  size(fHat) == p.NOut || throw(DimensionMismatch("fHat has size $(size(fHat)) but the plan expects $(p.NOut)"))
  (eltype(f) == Complex{T} && eltype(fHat) == Complex{T}) ||
    throw(ArgumentError("data must be Complex{$T} for this plan"))
end

function fill_timing!(p, timing, fwd::Bool)
  timing === nothing && return
  t = zeros(Float64, 7)
  check(ccall((:nfftm_timing, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Float64}), p.handle, t))
  if fwd
    timing.deconv, timing.fft, timing.conv = t[4], t[3], t[2]
  else
    timing.conv_adjoint, timing.fft_adjoint, timing.deconv_adjoint = t[5], t[6], t[7]
  end
end

# mul!(fHat, p, f): src/implementation.jl:155-172
function LinearAlgebra.mul!(fHat::ROCArray, p::MI355XNFFTPlan{T}, f::ROCArray;
                            verbose=false, timing::Union{Nothing,TimingStats}=nothing) where {T}
  consistency(p, f, fHat)
  check(ccall((:nfftm_set_timing, libnfftm), Cint, (Ptr{Cvoid}, Cint), p.handle, timing !== nothing))
  check(ccall((:nfftm_forward, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
              p.handle, devptr(f), devptr(fHat)))
  fill_timing!(p, timing, true)
  verbose && timing !== nothing && @info "Timing: deconv=$(timing.deconv) fft=$(timing.fft) conv=$(timing.conv)"
  return fHat
end

# mul!(f, adjoint(p), fHat): src/implementation.jl:176-193
function LinearAlgebra.mul!(f::ROCArray, pl::Adjoint{Complex{T},<:MI355XNFFTPlan{T}}, fHat::ROCArray;
                            verbose=false, timing::Union{Nothing,TimingStats}=nothing) where {T}
  p = pl.parent
  consistency(p, f, fHat)
  check(ccall((:nfftm_set_timing, libnfftm), Cint, (Ptr{Cvoid}, Cint), p.handle, timing !== nothing))
  check(ccall((:nfftm_adjoint, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
              p.handle, devptr(fHat), devptr(f)))
  fill_timing!(p, timing, false)
  verbose && timing !== nothing && @info "Timing: conv=$(timing.conv_adjoint) fft=$(timing.fft_adjoint) deconv=$(timing.deconv_adjoint)"
  return f
end

# Element types of the typed stage operators (include/nfftm.h NFFTM_R32 ...).
const ELT = Dict{DataType,Cint}(Float32 => Cint(0), Float64 => Cint(1), ComplexF32 => Cint(2),
                                ComplexF64 => Cint(3), Float16 => Cint(4), ComplexF16 => Cint(5))
elt(::Type{E}) where {E} = haskey(ELT, E) ? ELT[E] :
  throw(ArgumentError("element type $(E) is not supported by the MI355X stage operators"))

# convolve!(p, g, fHat): src/convolution.jl:20-53.  g and fHat may be real or
# complex of any precision (RealOrComplex, as NFFTTools.sdc! passes real
# weights and grids, samplingDensity.jl:106,112); the library converts
# through the plan's complex type and checks the lengths again.
function AbstractNFFTs.convolve!(p::MI355XNFFTPlan{T,D}, g::ROCArray{<:Any,D}, fHat::ROCArray) where {T,D}
  (eltype(g) <: Complex && !(eltype(fHat) <: Complex)) &&
    throw(ArgumentError("Complex input g requires Complex output fHat"))
  size(g) == p.Ñ || throw(DimensionMismatch("size(g)=$(size(g)) ≠ Ñ = $(p.Ñ)"))
  size(fHat) == p.NOut || throw(DimensionMismatch("size(fHat)=$(size(fHat)) ≠ $(p.NOut)"))
  if p.dims != 1:D  # directional (src/directional.jl:57-101): the plan's complex type
    (eltype(g) == Complex{T} && eltype(fHat) == Complex{T}) ||
      throw(ArgumentError("a directional plan convolves Complex{$T} arrays"))
    check(ccall((:nfftm_convolve, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                p.handle, devptr(g), devptr(fHat)))
    return fHat
  end
  check(ccall((:nfftm_convolve_typed, libnfftm), Cint,
              (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}, Cint, Int64),
              p.handle, devptr(g), elt(eltype(g)), length(g), devptr(fHat), elt(eltype(fHat)), length(fHat)))
  return fHat
end

# convolve_transpose!(p, fHat, g): src/convolution.jl:115-149.
function AbstractNFFTs.convolve_transpose!(p::MI355XNFFTPlan{T,D}, fHat::ROCArray, g::ROCArray{<:Any,D}) where {T,D}
  (eltype(fHat) <: Complex && !(eltype(g) <: Complex)) &&
    throw(ArgumentError("Complex input fHat requires Complex output g"))
  size(g) == p.Ñ || throw(DimensionMismatch("size(g)=$(size(g)) ≠ Ñ = $(p.Ñ)"))
  size(fHat) == p.NOut || throw(DimensionMismatch("size(fHat)=$(size(fHat)) ≠ $(p.NOut)"))
  if p.dims != 1:D  # directional (src/directional.jl:103-156)
    (eltype(g) == Complex{T} && eltype(fHat) == Complex{T}) ||
      throw(ArgumentError("a directional plan convolves Complex{$T} arrays"))
    check(ccall((:nfftm_convolve_transpose, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                p.handle, devptr(fHat), devptr(g)))
    return g
  end
  check(ccall((:nfftm_convolve_transpose_typed, libnfftm), Cint,
              (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}, Cint, Int64),
              p.handle, devptr(fHat), elt(eltype(fHat)), length(fHat), devptr(g), elt(eltype(g)), length(g)))
  return g
end

# deconvolve!(p, f, g) and deconvolve_transpose!(p, g, f):
# src/deconvolution.jl:2-92, on Complex{T} buffers of the plan's sizes.
function AbstractNFFTs.deconvolve!(p::MI355XNFFTPlan{T,D}, f::ROCArray{Complex{T},D},
                                   g::ROCArray{Complex{T},D}) where {T,D}
  size(f) == p.N || throw(DimensionMismatch("size(f)=$(size(f)) ≠ N = $(p.N)"))
  size(g) == p.Ñ || throw(DimensionMismatch("size(g)=$(size(g)) ≠ Ñ = $(p.Ñ)"))
  check(ccall((:nfftm_deconvolve, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
              p.handle, devptr(f), devptr(g)))
  return g
end

function AbstractNFFTs.deconvolve_transpose!(p::MI355XNFFTPlan{T,D}, g::ROCArray{Complex{T},D},
                                             f::ROCArray{Complex{T},D}) where {T,D}
  size(f) == p.N || throw(DimensionMismatch("size(f)=$(size(f)) ≠ N = $(p.N)"))
  size(g) == p.Ñ || throw(DimensionMismatch("size(g)=$(size(g)) ≠ Ñ = $(p.Ñ)"))
  check(ccall((:nfftm_deconvolve_transpose, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
              p.handle, devptr(g), devptr(f)))
  return f
end

# NFFTTools on the plan.  sdc!(p, iters, weights, weights_tmp, workg[, workf,
# workv]) (NFFTTools/src/samplingDensity.jl:93-155): the iterations run on
# the device (nfftm_sdc); the work buffers of the generic method are not
# needed.  Throws "non-positive weights" as the reference does.
function NFFTTools.sdc!(p::MI355XNFFTPlan{T}, iters::Int, weights::ROCVector{T}, weights_tmp::AbstractVector,
                        workg::AbstractArray) where {T<:Real}
  length(weights) == p.J || throw(DimensionMismatch("length(weights)=$(length(weights)) ≠ J = $(p.J)"))
  rc = ccall((:nfftm_sdc, libnfftm), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Cint), p.handle, iters, devptr(weights), 0)
  rc == NFFTM_ERR_UNSUPPORTED && throw("non-positive weights")
  check(rc)
  return weights
end

function NFFTTools.sdc!(p::MI355XNFFTPlan{T}, iters::Int, weights::ROCVector{T}, weights_tmp::AbstractVector,
                        workg::AbstractArray, workf::AbstractVector, workv::AbstractArray) where {T<:Real}
  length(weights) == p.J || throw(DimensionMismatch("length(weights)=$(length(weights)) ≠ J = $(p.J)"))
  rc = ccall((:nfftm_sdc, libnfftm), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Cint), p.handle, iters, devptr(weights), 1)
  rc == NFFTM_ERR_UNSUPPORTED && throw("non-positive weights")
  check(rc)
  return weights
end

# Toeplitz embedding (NFFTTools/src/Toeplitz.jl): one library object per
# (precision, image size), created on first use and freed with the process.
mutable struct ToeplitzObj
  handle::Ptr{Cvoid}
end
const TOEPLITZ = Dict{Tuple{DataType,Tuple},ToeplitzObj}()
function toeplitz_obj(::Type{T}, N::NTuple{D,Int}) where {T,D}
  get!(TOEPLITZ, (T, N)) do
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:nfftm_toeplitz_create, libnfftm), Cint, (Ref{Ptr{Cvoid}}, Cint, Cint, Ptr{Int64}, Ptr{Cvoid}),
                h, precision_of(T), D, collect(Int64, N), AMDGPU.stream().stream))
    o = ToeplitzObj(h[])
    finalizer(x -> ccall((:nfftm_toeplitz_destroy, libnfftm), Cvoid, (Ptr{Cvoid},), x.handle), o)
    o
  end
end

# calculateToeplitzKernel!(f, p, tr, fftplan) (Toeplitz.jl:119-125): nodes!(p,
# tr), then FFT(fftshift(adjoint(p) * ones)) on the device into f (size 2N =
# the plan's N); fftplan is not used.
function NFFTTools.calculateToeplitzKernel!(f::ROCArray{Complex{T},D}, p::MI355XNFFTPlan{T,D}, tr::Matrix{T},
                                            fftplan=nothing) where {T,D}
  size(f) == p.N || throw(DimensionMismatch("size(f)=$(size(f)) ≠ the plan's N = $(p.N)"))
  all(iseven, p.N) || throw(ArgumentError("the Toeplitz kernel needs a plan of size 2N"))
  AbstractNFFTs.nodes!(p, tr)
  t = toeplitz_obj(T, p.N .÷ 2)
  check(ccall((:nfftm_toeplitz_kernel, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
              t.handle, p.handle, devptr(f)))
  return f
end

# convolveToeplitzKernel!(y, λ, ...) (Toeplitz.jl:236-250): y overwritten by
# the N-corner of IFFT(FFT(pad(y)) .* λ); the optional plans and buffers of
# the generic method are not needed.
function NFFTTools.convolveToeplitzKernel!(y::ROCArray{Complex{T},D}, λ::ROCArray{Complex{T},D}, args...) where {T,D}
  size(λ) == 2 .* size(y) || throw(DimensionMismatch("size(λ)=$(size(λ)) ≠ 2 .* size(y) = $(2 .* size(y))"))
  t = toeplitz_obj(T, size(y))
  check(ccall((:nfftm_toeplitz_apply, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
              t.handle, devptr(λ), devptr(y)))
  return y
end

"""
    grad_nodes(p, f, ȳ) -> k̄::Matrix{Float64}

Node gradient of y = p * f contracted with the cotangent ȳ:
k̄[d, j] = Re(conj(ȳ[j]) * ∂y[j]/∂k[d, j]), evaluated on the device with the
analytic derivative of the Kaiser-Bessel window.
"""
function grad_nodes(p::MI355XNFFTPlan{T,D}, f::ROCArray, ȳ::ROCArray) where {T,D}
  consistency(p, f, ȳ)
  kbar = ROCArray{Float64}(undef, D, p.J)
  check(ccall((:nfftm_grad_nodes, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
              p.handle, devptr(f), devptr(ȳ), devptr(kbar)))
  return Array(kbar)
end

"""
    grad_nodes_grid(p, g, ȳ) -> k̄::Matrix{Float64}

The node gradient of `grad_nodes` from the transformed grid g of the forward
transform (`p.tmpVec` right after `p * f`, or a copy of it): only the
window-derivative gather runs.
"""
function grad_nodes_grid(p::MI355XNFFTPlan{T,D}, g::ROCArray{Complex{T}}, ȳ::ROCArray) where {T,D}
  size(g) == p.Ñ || throw(DimensionMismatch("size(g) = $(size(g)) ≠ Ñ = $(p.Ñ)"))
  (eltype(ȳ) == Complex{T} && length(ȳ) == p.J) ||
    throw(DimensionMismatch("ȳ must be a Complex{$T} vector of length J = $(p.J)"))
  kbar = ROCArray{Float64}(undef, D, p.J)
  check(ccall((:nfftm_grad_nodes_grid, libnfftm), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
              p.handle, devptr(g), devptr(ȳ), devptr(kbar)))
  return Array(kbar)
end

# AbstractNFFTs/ext/AbstractNFFTsChainRulesCoreExt.jl:26-45 returns NoTangent()
# for the nodes; this rule returns their gradient as well, from the forward
# pass's transformed grid.
function ChainRulesCore.rrule(::typeof(AbstractNFFTs.nfft), k::Matrix{T}, x::ROCArray) where {T}
  p = AbstractNFFTs.plan_nfft(MI355XBackend(), ROCArray, k, size(x))
  y = p * x
  g = copy(p.tmpVec)
  function nfft_pullback(ȳ)
    ȳ = ROCArray(ChainRulesCore.unthunk(ȳ))
    k̄ = T.(grad_nodes_grid(p, g, ȳ))
    x̄ = adjoint(p) * ȳ
    return ChainRulesCore.NoTangent(), k̄, x̄
  end
  return y, nfft_pullback
end

end # module
