# FunicularHIPExt: Funicular's device contract (src/device_api.jl:12-138) on
# libfunicular_hip.so, for a Funicular.jl checkout that declares the stub
#
#     function hip_backend(args...; kwargs...)
#         throw(ArgumentError("the HIP backend arrives with FunicularHIP, which has to be loaded first"))
#     end
#
# next to cuda_backend (src/device_api.jl:173-175) and lists this file as a
# package extension. It imports only the contract names, as the CUDA
# extension does (ext/FunicularCUDAExt.jl:8-12), and reaches the GPU through
# the C ABI declared in include/funicular_hip.h. Nothing here is Julia-side
# GPU code: arrays are device pointers with a shape and a leading dimension,
# the kernels are the library's.
#
# This image has no Julia, so this file is not exercised here; the Python host
# package funicular_hip is the same mapping and is what the tests run.
module FunicularHIPExt

import Funicular: alloc_device, alloc_host_slab, d2h!, devicetype, h2d!, hip_backend, ka_backend,
                  make_queue, make_sweep_queues, record_event, submit!, sync_all, sync_event, sync_queue,
                  wait_event, supports_eltype, gram_accumulate!, rightmul_gemm!, rdiv_upper!, convert_copy!

import Funicular
using Funicular: DeviceBackend, SLAB_ALIGN, aligned_slab
using LinearAlgebra

const LIB = Ref{String}(get(ENV, "FUNICULAR_HIP_LIB", joinpath(@__DIR__, "..", "funicular_hip", "libfunicular_hip.so")))

const FHIP_F32, FHIP_F64, FHIP_C32, FHIP_C64 = 0, 1, 2, 3
const FHIP_H2D, FHIP_D2H, FHIP_D2D = 1, 2, 3

eltcode(::Type{Float32}) = FHIP_F32
eltcode(::Type{Float64}) = FHIP_F64
eltcode(::Type{ComplexF32}) = FHIP_C32
eltcode(::Type{ComplexF64}) = FHIP_C64

function check(status::Cint)
    status == 0 && return nothing
    msg = unsafe_string(ccall((:fhip_last_error, LIB[]), Cstring, ()))
    status == 2 && throw(DimensionMismatch(msg))
    status == 1 && throw(ArgumentError(msg))
    error(msg)
end

struct HIPBackend <: DeviceBackend end
hip_backend() = HIPBackend()

# A device matrix: pointer, shape, leading dimension. Views keep the parent's
# leading dimension (the CUDA extension's region_pointer/copy_shape,
# ext/FunicularCUDAExt.jl:157-182).
mutable struct HIPMatrix{T} <: AbstractMatrix{T}
    ptr::Ptr{T}
    dims::Dims{2}
    ld::Int
    owner::Any
end
Base.size(A::HIPMatrix) = A.dims
Base.getindex(::HIPMatrix, ::Int...) = throw(ArgumentError("HIPMatrix has no scalar indexing; copy it to the host"))
Base.view(A::HIPMatrix{T}, rows::AbstractUnitRange, cols::AbstractUnitRange) where {T} =
    HIPMatrix{T}(A.ptr + ((first(rows) - 1) + (first(cols) - 1) * A.ld) * sizeof(T), (length(rows), length(cols)), A.ld, A.owner)
Base.view(A::HIPMatrix, ::Colon, cols::AbstractUnitRange) = view(A, 1:size(A, 1), cols)
Base.view(A::HIPMatrix, rows::AbstractUnitRange, ::Colon) = view(A, rows, 1:size(A, 2))

devicetype(::HIPBackend, ::Type{T}) where {T} = HIPMatrix{T}
supports_eltype(::HIPBackend, ::Type{T}) where {T} = T in (Float32, Float64, ComplexF32, ComplexF64)

# Funicular reaches ka_backend for two things: to launch its KernelAbstractions
# kernels (src/kernels.jl), which this backend replaces with the library's
# (convert_copy!, gram_accumulate!, rightmul_gemm!, rdiv_upper! below), and to
# wait after a bare convert_staged! (src/panels.jl:365-371). KernelAbstractions
# has no AMD backend without AMDGPU.jl, so HIPKernels is a handle whose
# synchronize drains the device; nothing is ever launched through it.
const KA = Funicular.KernelAbstractions
struct HIPKernels <: KA.GPU end
KA.synchronize(::HIPKernels) = (check(ccall((:fhip_device_sync, LIB[]), Cint, ())); nothing)
ka_backend(::HIPBackend) = HIPKernels()

function alloc_device(::HIPBackend, ::Type{T}, dims::Dims) where {T}
    d2 = length(dims) == 1 ? (dims[1], 1) : dims
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fhip_malloc, LIB[]), Cint, (Ptr{Ptr{Cvoid}}, Csize_t), p, max(prod(d2) * sizeof(T), 1)))
    owner = p[]
    A = HIPMatrix{T}(Ptr{T}(owner), d2, max(d2[1], 1), owner)
    finalizer(a -> ccall((:fhip_free, LIB[]), Cint, (Ptr{Cvoid},), a.owner), A)
    A
end

function alloc_host_slab(::HIPBackend, nbytes::Integer)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fhip_host_alloc, LIB[]), Cint, (Ptr{Ptr{Cvoid}}, Csize_t), p, nbytes + SLAB_ALIGN))
    owner = Ref(p[])
    finalizer(o -> ccall((:fhip_host_free, LIB[]), Cint, (Ptr{Cvoid},), o[]), owner)
    aligned_slab(owner, Ptr{UInt8}(p[]), nbytes)
end

# make_sweep_queues runs once per sweep (src/executor.jl:193), so each queue
# owns its stream and releases it when collected, as CUDA.jl's streams do.
# Work still queued on the stream completes before its resources go.
mutable struct HIPQueue
    stream::Ptr{Cvoid}
end

function make_queue(::HIPBackend)
    s = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fhip_stream_create, LIB[]), Cint, (Ptr{Ptr{Cvoid}},), s))
    q = HIPQueue(s[])
    finalizer(x -> ccall((:fhip_stream_destroy, LIB[]), Cint, (Ptr{Cvoid},), x.stream), q)
    q
end

mutable struct HIPEvent
    handle::Ptr{Cvoid}
end

# The stream that kernels launched from a panel function land on: the task
# that runs submit! stores it here (the task-local stream CUDA.jl gave the
# CUDA extension, ext/FunicularCUDAExt.jl:83-89).
current_stream() = get(task_local_storage(), :fhip_stream, C_NULL)

function record_event(::HIPBackend, q::HIPQueue)
    e = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fhip_event_create, LIB[]), Cint, (Ptr{Ptr{Cvoid}}, Cint), e, 1))
    ev = HIPEvent(e[])
    finalizer(x -> ccall((:fhip_event_destroy, LIB[]), Cint, (Ptr{Cvoid},), x.handle), ev)
    check(ccall((:fhip_event_record, LIB[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ev.handle, q.stream))
    ev
end

wait_event(::HIPBackend, q::HIPQueue, e::HIPEvent) =
    (check(ccall((:fhip_stream_wait_event, LIB[]), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), q.stream, e.handle)); nothing)
sync_queue(::HIPBackend, q::HIPQueue) = (check(ccall((:fhip_stream_sync, LIB[]), Cint, (Ptr{Cvoid},), q.stream)); nothing)
sync_event(::HIPBackend, e::HIPEvent) = (check(ccall((:fhip_event_sync, LIB[]), Cint, (Ptr{Cvoid},), e.handle)); nothing)
sync_all(::HIPBackend) = (check(ccall((:fhip_device_sync, LIB[]), Cint, ())); nothing)

function submit!(b::HIPBackend, q::HIPQueue, work)
    task_local_storage(work, :fhip_stream, q.stream)
    record_event(b, q)
end

# Every copy is one 2-D copy with both pitches (ext/FunicularCUDAExt.jl:134-152).
region(A::HIPMatrix{T}) where {T} = (Ptr{Cvoid}(A.ptr), A.ld * sizeof(T))
function region(A::StridedMatrix{T}) where {T}
    stride(A, 1) == 1 || throw(ArgumentError("a copy needs its columns contiguous"))
    (Ptr{Cvoid}(pointer(A)), stride(A, 2) * sizeof(T))
end

# A byte copy between two arrays of one eltype, queued on `stream`.
function raw_copy!(dst, src, kind, stream)
    size(dst) == size(src) || throw(DimensionMismatch("sizes differ: destination $(size(dst)), source $(size(src))"))
    eltype(dst) === eltype(src) ||
        throw(ArgumentError("a byte copy needs one eltype on both sides, got $(eltype(src)) into $(eltype(dst)); " *
                            "h2d! and d2h! convert through a device staging buffer"))
    isempty(dst) && return nothing
    T = eltype(dst)
    dp, dpitch = region(dst)
    sp, spitch = region(src)
    GC.@preserve dst src check(ccall((:fhip_memcpy2d_async, LIB[]), Cint,
        (Ptr{Cvoid}, Csize_t, Ptr{Cvoid}, Csize_t, Csize_t, Csize_t, Cint, Ptr{Cvoid}),
        dp, dpitch, sp, spitch, size(dst, 1) * sizeof(T), size(dst, 2), kind, stream))
    nothing
end

function copy2d!(dst, src, kind, queue)
    raw_copy!(dst, src, kind, queue === nothing ? C_NULL : queue.stream)
    queue === nothing && check(ccall((:fhip_stream_sync, LIB[]), Cint, (Ptr{Cvoid},), C_NULL))
    nothing
end

# The contract lets h2d! and d2h! change the eltype (src/device_api.jl:31-44).
# Outside a sweep there is no staging buffer of the source eltype, so, as the
# CUDA extension does (ext/FunicularCUDAExt.jl:195-212), a converting copy
# allocates one on the device, moves the bytes, converts with the library's
# kernel on the same stream, and waits before the buffer is let go.
function converting_copy!(dst, src, b::HIPBackend, queue)
    size(dst) == size(src) || throw(DimensionMismatch("sizes differ: destination $(size(dst)), source $(size(src))"))
    stream = queue === nothing ? C_NULL : queue.stream
    staging = nothing
    task_local_storage(:fhip_stream, stream) do
        if dst isa HIPMatrix && src isa HIPMatrix
            convert_copy!(b, dst, src)
        elseif dst isa HIPMatrix
            staging = alloc_device(b, eltype(src), size(src))
            raw_copy!(staging, src, FHIP_H2D, stream)
            convert_copy!(b, dst, staging)
        else
            staging = alloc_device(b, eltype(dst), size(dst))
            convert_copy!(b, staging, src)
            raw_copy!(dst, staging, FHIP_D2H, stream)
        end
    end
    GC.@preserve staging check(ccall((:fhip_stream_sync, LIB[]), Cint, (Ptr{Cvoid},), stream))
    queue === nothing ? nothing : record_event(b, queue)
end

function h2d!(dst::HIPMatrix, src, b::HIPBackend; queue=nothing)
    eltype(dst) === eltype(src) || return converting_copy!(dst, src, b, queue)
    copy2d!(dst, src, FHIP_H2D, queue)
    queue === nothing ? nothing : record_event(b, queue)
end

function d2h!(dst, src::HIPMatrix, b::HIPBackend; queue=nothing)
    eltype(dst) === eltype(src) || return converting_copy!(dst, src, b, queue)
    copy2d!(dst, src, FHIP_D2H, queue)
    queue === nothing ? nothing : record_event(b, queue)
end

# The kernel overrides (the Metal extension's pattern, ext/FunicularMetalExt.jl:187-189).
gram_accumulate!(::HIPBackend, C::HIPMatrix{T}, A::HIPMatrix{T}, B::HIPMatrix{T}, α, β) where {T} =
    gram_raw!(C, A, B, α, β, eltcode(T))

# C <- α A^H B + β C with the products in eltype `compute` (FHIP_C32 on
# ComplexF64 storage is the mixed first pass: FP32 products, FP64 sums).
function gram_raw!(C::HIPMatrix{T}, A::HIPMatrix{T}, B::HIPMatrix{T}, α, β, compute::Integer) where {T}
    herm = Cint(A.ptr == B.ptr && A.ld == B.ld && size(A, 2) == size(B, 2))
    n = Ref{Csize_t}(0)
    check(ccall((:fhip_gram_workspace_bytes, LIB[]), Cint, (Ptr{Csize_t}, Int64, Int64, Int64, Cint, Cint, Cint),
                n, size(A, 1), size(A, 2), size(B, 2), eltcode(T), compute, herm))
    work = workspace(n[])
    a, bb = ComplexF64(α), ComplexF64(β)
    check(ccall((:fhip_gram, LIB[]), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64, Int64, Float64, Float64, Float64, Float64,
         Cint, Cint, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
        C.ptr, C.ld, A.ptr, A.ld, B.ptr, B.ld, size(A, 1), size(A, 2), size(B, 2), real(a), imag(a), real(bb), imag(bb),
        eltcode(T), compute, work.ptr, n[], current_stream()))
    C
end

rdiv_upper!(::HIPBackend, Y::HIPMatrix{T}, R::HIPMatrix{T}) where {T} = trsm_raw!(Y, R, 1, eltcode(T))

# With the block index passed to rdiv_upper! by rdiv_rows! (src/panelblas.jl:157-169;
# the default method for other backends drops it), R is folded into the solve's
# workspace once, on the sweep's first block, and the later blocks reuse it: the
# solve's workspace is its own, so a Gram between the blocks (the fused next
# pass) does not overwrite the folded factor. Every block of one rdiv_rows!
# runs on the one compute stream, so the fold is complete before a later block
# reads it.
rdiv_upper!(::HIPBackend, Y::HIPMatrix{T}, R::HIPMatrix{T}, block::Integer) where {T} =
    trsm_raw!(Y, R, block, eltcode(T))

# Y <- Y R^-1 with the products in eltype `compute`; block 1 folds R into the
# solve's workspace, later blocks of the same sweep reuse it.
function trsm_raw!(Y::HIPMatrix{T}, R::HIPMatrix{T}, block::Integer, compute::Integer) where {T}
    n = Ref{Csize_t}(0)
    check(ccall((:fhip_trsm_workspace_bytes, LIB[]), Cint, (Ptr{Csize_t}, Int64, Cint), n, size(R, 1), compute))
    work = workspace(n[]; which=TRSM_WORK)
    if block == 1
        check(ccall((:fhip_trsm_runn, LIB[]), Cint,
            (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Cvoid}, Int64, Cint, Cint, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
            Y.ptr, Y.ld, size(Y, 1), size(Y, 2), R.ptr, R.ld, eltcode(T), compute, work.ptr, n[], current_stream()))
    else
        check(ccall((:fhip_trsm_apply, LIB[]), Cint,
            (Ptr{Cvoid}, Int64, Int64, Int64, Cint, Cint, Cint, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
            Y.ptr, Y.ld, size(Y, 1), size(Y, 2), eltcode(T), compute, Cint(0), work.ptr, n[], current_stream()))
    end
    Y
end

# Y <- Y R^-H (R upper triangular), the solve of the QR adjoint, folded once
# per sweep in the same way. No counterpart in the reference.
function rdiv_upper_adjoint!(::HIPBackend, Y::HIPMatrix{T}, R::HIPMatrix{T}, block::Integer=1) where {T}
    n = Ref{Csize_t}(0)
    check(ccall((:fhip_trsm_workspace_bytes, LIB[]), Cint, (Ptr{Csize_t}, Int64, Cint), n, size(R, 1), eltcode(T)))
    work = workspace(n[]; which=TRSM_WORK)
    if block == 1
        check(ccall((:fhip_trsm_runc, LIB[]), Cint,
            (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Cvoid}, Int64, Cint, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
            Y.ptr, Y.ld, size(Y, 1), size(Y, 2), R.ptr, R.ld, eltcode(T), work.ptr, n[], current_stream()))
    else
        check(ccall((:fhip_trsm_apply, LIB[]), Cint,
            (Ptr{Cvoid}, Int64, Int64, Int64, Cint, Cint, Cint, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
            Y.ptr, Y.ld, size(Y, 1), size(Y, 2), eltcode(T), eltcode(T), Cint(1), work.ptr, n[], current_stream()))
    end
    Y
end

# out[i] (+)= Re sum_j conj(A[i, j]) B[i, j]: the sensitivity of S = Q^H G Q
# to a real diagonal of G, one panel at a time. No counterpart in the reference.
function rowdot!(::HIPBackend, out::HIPMatrix{Float64}, A::HIPMatrix{T}, B::HIPMatrix{T}, accumulate::Bool) where {T}
    size(A) == size(B) && size(out, 1) >= size(A, 1) ||
        throw(DimensionMismatch("rowdot of $(size(A)) and $(size(B)) into $(size(out))"))
    check(ccall((:fhip_rowdot, LIB[]), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64, Cint, Cint, Ptr{Cvoid}),
        out.ptr, A.ptr, A.ld, B.ptr, B.ld, size(A, 1), size(A, 2), eltcode(T), Cint(accumulate), current_stream()))
    out
end

function rightmul_gemm!(::HIPBackend, dst::HIPMatrix{T}, A::HIPMatrix{T}, C::HIPMatrix{T}) where {T}
    check(ccall((:fhip_gemm_nn, LIB[]), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64, Int64, Cint, Ptr{Cvoid}),
        dst.ptr, dst.ld, A.ptr, A.ld, C.ptr, C.ld, size(A, 1), size(A, 2), size(C, 2), eltcode(T), current_stream()))
    dst
end

function convert_copy!(::HIPBackend, dst::HIPMatrix{S}, src::HIPMatrix{T}) where {S,T}
    size(dst) == size(src) || throw(DimensionMismatch("sizes differ: destination $(size(dst)), source $(size(src))"))
    check(ccall((:fhip_convert_copy, LIB[]), Cint,
        (Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}, Cint, Int64, Int64, Int64, Ptr{Cvoid}),
        dst.ptr, eltcode(S), dst.ld, src.ptr, eltcode(T), src.ld, size(dst, 1), size(dst, 2), current_stream()))
    dst
end

# scale! and axpy! run unchanged: their panel steps are rmul! and axpy! on the
# staged device buffers (src/panelblas.jl:311-331).
function LinearAlgebra.rmul!(Y::HIPMatrix{T}, α::Number) where {T}
    a = ComplexF64(α)
    check(ccall((:fhip_scale, LIB[]), Cint,
        (Ptr{Cvoid}, Int64, Int64, Int64, Float64, Float64, Cint, Cint, Ptr{Cvoid}),
        Y.ptr, Y.ld, size(Y, 1), size(Y, 2), real(a), imag(a), Cint(α isa Real), eltcode(T), current_stream()))
    Y
end

# The broadcasts of a foreachpanel closure (`panel .*= a`, `panel .= a .* panel
# .+ b`, `panel .= v`) have no device array package to compile them here; a
# closure spells them as rmul!(panel, a), affine!(panel, a, b) and
# fill!(panel, v), which give the same values as the broadcasts.
function affine!(Y::HIPMatrix{T}, α::Number, β::Number) where {T}
    a, b = ComplexF64(α), ComplexF64(β)
    check(ccall((:fhip_affine, LIB[]), Cint,
        (Ptr{Cvoid}, Int64, Int64, Int64, Float64, Float64, Cint, Float64, Float64, Cint, Ptr{Cvoid}),
        Y.ptr, Y.ld, size(Y, 1), size(Y, 2), real(a), imag(a), Cint(α isa Real), real(b), imag(b), eltcode(T),
        current_stream()))
    Y
end

function Base.fill!(Y::HIPMatrix{T}, v) where {T}
    x = ComplexF64(v)
    check(ccall((:fhip_fill, LIB[]), Cint, (Ptr{Cvoid}, Cint, Int64, Int64, Int64, Float64, Float64, Ptr{Cvoid}),
        Y.ptr, eltcode(T), Y.ld, size(Y, 1), size(Y, 2), real(x), imag(x), current_stream()))
    Y
end

function LinearAlgebra.axpy!(α::Number, X::HIPMatrix{T}, Y::HIPMatrix{T}) where {T}
    size(X) == size(Y) || throw(DimensionMismatch("axpy! between a $(size(X)) and a $(size(Y)) matrix"))
    a = ComplexF64(α)
    check(ccall((:fhip_axpby, LIB[]), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64, Float64, Float64, Float64, Float64, Cint, Ptr{Cvoid}),
        Y.ptr, Y.ld, X.ptr, X.ld, size(Y, 1), size(Y, 2), real(a), imag(a), 1.0, 0.0, eltcode(T), current_stream()))
    Y
end

# dot and norm (src/panelblas.jl:333-370) broadcast into an N x w device
# accumulator, which a HIPMatrix cannot do. With a per-panel hook in Funicular,
#
#     panel_reduce!(b, acc, Xb, Yb, first) = <the broadcast step of dot/norm>
#
# (Yb === nothing for norm), this method sums the panel straight into a 1 x 1
# ComplexF64 device scalar in FP64, in an order fixed by the panel's shape.
function panel_reduce!(::HIPBackend, acc::HIPMatrix{ComplexF64}, Xb::HIPMatrix{T}, Yb, first::Bool) where {T}
    n = Ref{Csize_t}(0)
    check(ccall((:fhip_dot_workspace_bytes, LIB[]), Cint, (Ptr{Csize_t}, Int64, Int64), n, size(Xb, 1), size(Xb, 2)))
    work = workspace(n[])
    y = Yb === nothing ? Xb : Yb
    check(ccall((:fhip_dot, LIB[]), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64, Cint, Cint, Cint, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
        acc.ptr, Xb.ptr, Xb.ld, y.ptr, y.ld, size(Xb, 1), size(Xb, 2), eltcode(T), Cint(Yb === nothing),
        Cint(!first), work.ptr, n[], current_stream()))
    acc
end

# With a `factor_gram!(b, R, G, N) -> shifted` hook in cholqr_pass!
# (src/panelblas.jl:141-146; its default copies G to the host and runs
# cholesky_upper), the plain factorization runs on the device; a bad pivot
# hands G to cholesky_upper on the host, which retries and shifts as the
# reference does.
const INFO = Ref{Any}(nothing)

function factor_gram!(b::HIPBackend, R::HIPMatrix{T}, G::HIPMatrix{T}, N::Integer) where {T}
    INFO[] === nothing && (INFO[] = alloc_device(b, Cint, (1, 1)))  # one status word, kept for every pass
    info = INFO[]
    check(ccall((:fhip_potrf_upper, LIB[]), Cint,
        (Ptr{Cvoid}, Int64, Cint, Ptr{Cvoid}, Int64, Cint, Int64, Ptr{Cvoid}, Ptr{Cvoid}),
        R.ptr, R.ld, eltcode(T), G.ptr, G.ld, eltcode(T), size(G, 1), info.ptr, current_stream()))
    bad = Ref{Cint}(0)
    check(ccall((:fhip_memcpy2d_async, LIB[]), Cint,
        (Ptr{Cvoid}, Csize_t, Ptr{Cvoid}, Csize_t, Csize_t, Csize_t, Cint, Ptr{Cvoid}),
        bad, sizeof(Cint), info.ptr, sizeof(Cint), sizeof(Cint), 1, FHIP_D2H, current_stream()))
    check(ccall((:fhip_stream_sync, LIB[]), Cint, (Ptr{Cvoid},), current_stream()))
    bad[] == 0 && return false
    G_host = Array{T}(undef, size(G))
    copy2d!(G_host, G, FHIP_D2H, nothing)
    R_host, shifted = Funicular.cholesky_upper(G_host, N, size(G, 1))  # the reference's retry and shift
    copy2d!(R, R_host, FHIP_H2D, nothing)
    shifted
end

# cholqr2! with a mixed-precision first pass, on the reference's step order
# (src/panelblas.jl:130-146): the first pass forms its Gram on the FP32 matrix
# cores with FP64 sums across row chunks, factors it in FP32 on the host
# (unshifted first with shift=:adaptive, or always with the shift
# 11 (N k + k (k + 1)) u32 ||Y||_2^2 with shift=:fukaya) and divides by that
# factor with FP32 products (solve=:fp32) or FP64 ones (solve=:fp64); the
# second pass is the reference's FP64 cholqr_pass!, and a third runs when
# either factorization was shifted. Returns R with Q R = Y; Y holds Q. No
# counterpart in the reference.
function cholqr2_mixed!(Y::Funicular.PanelMatrix{ComplexF64}; nbuffers=nothing, shift::Symbol=:adaptive,
                        solve::Symbol=:fp32)
    shift in (:adaptive, :fukaya) || throw(ArgumentError("shift must be :adaptive or :fukaya, got $shift"))
    solve in (:fp32, :fp64) || throw(ArgumentError("solve must be :fp32 or :fp64, got $solve"))
    Y.plan.backend isa HIPBackend || throw(ArgumentError("cholqr2_mixed! runs on the HIP backend"))
    G = gram_fp32(Y; nbuffers=nbuffers)
    R32, s = cholesky_upper_fp32(G, Y.N, Y.k, shift)
    R1 = Matrix{ComplexF64}(R32)
    rdiv_rows_compute!(Y, R1, solve === :fp32 ? FHIP_C32 : FHIP_C64; nbuffers=nbuffers)
    R2, shifted2 = Funicular.cholqr_pass!(Y, nbuffers)
    R = R2 * R1
    if s > 0 || shifted2
        R3, _ = Funicular.cholqr_pass!(Y, nbuffers)
        R = R3 * R
    end
    R
end

# The first pass's Gram: gram's row sweep (src/panelblas.jl:33-53) with FP32
# products.
function gram_fp32(Y::Funicular.PanelMatrix{ComplexF64}; nbuffers=nothing)
    plan = Y.plan
    acc = only(Funicular.checkout_device_buffers!(plan, ComplexF64, (Y.k, Y.k), 1))
    try
        Funicular.sweep!(Funicular.readpanels(Y); nbuffers=nbuffers, traversal=Funicular.rowtraversal(Y)) do block, Yb
            gram_raw!(acc, Yb, Yb, 1.0, block == 1 ? 0.0 : 1.0, FHIP_C32)
        end
        Funicular.collect_small(acc, plan)
    finally
        Funicular.checkin_device_buffers!(plan, (acc,))
    end
end

function cholesky_upper_fp32(G::Matrix{ComplexF64}, N::Integer, k::Integer, shift::Symbol)
    if shift === :adaptive
        plain = cholesky(Hermitian(Matrix{ComplexF32}(G)); check=false)
        issuccess(plain) && return Matrix(plain.U), 0.0
    end
    s = 11 * (N * k + k * (k + 1)) * (eps(Float32) / 2) * eigmax(Hermitian(G))
    shifted = cholesky(Hermitian(Matrix{ComplexF32}(G + s * I)); check=false)
    issuccess(shifted) || throw(ArgumentError("the $(k)×$(k) FP32 Gram matrix is not positive definite after " *
                                              "the shift $s; the columns are numerically dependent"))
    Matrix(shifted.U), s
end

# rdiv_rows! (src/panelblas.jl:157-169) with the solve's products in `compute`.
function rdiv_rows_compute!(Y::Funicular.PanelMatrix{ComplexF64}, R::Matrix{ComplexF64}, compute::Integer;
                            nbuffers=nothing)
    plan = Y.plan
    factor = only(Funicular.checkout_device_buffers!(plan, ComplexF64, size(R), 1))
    try
        h2d!(factor, R, plan.backend)
        Funicular.sweep!(Funicular.updatepanels(Y); nbuffers=nbuffers, traversal=Funicular.rowtraversal(Y)) do block, Yb
            trsm_raw!(Yb, factor, block, compute)
        end
    finally
        Funicular.checkin_device_buffers!(plan, (factor,))
    end
    Y
end

# Workspaces that grow to the largest request and are reused: the first
# sweep of a shape allocates, later ones do not. The solve keeps its own (it
# holds the folded factor between row blocks). Before a smaller buffer is
# dropped the device is drained, since a kernel already queued may still be
# using it and dropping it lets the finalizer free it.
const WORK = Ref{Any}(nothing)
const TRSM_WORK = Ref{Any}(nothing)
function workspace(n; which=WORK)
    w = which[]
    if w === nothing || w.nbytes < n
        w === nothing || check(ccall((:fhip_device_sync, LIB[]), Cint, ()))
        buf = alloc_device(HIPBackend(), UInt8, (max(Int(n), 1), 1))
        which[] = (ptr=buf.ptr, nbytes=Int(n), owner=buf)
        w = which[]
    end
    w
end

# The two stencils of the benchmark configurations as Funicular operators
# (src/operator.jl:1-25: size, eltype, adjoint, mul!, and the panel_capable and
# ishermitian_op traits), applied to whole device panels by the library:
# the 3-D 7-point periodic Helmholtz stencil on n1 x n2 x n3 (linear index
# i + n1 (j + n2 l)), and the periodic tridiagonal with an optional real
# diagonal kept on the device.
struct HelmholtzStencil{T}
    n::NTuple{3,Int}
    d::ComplexF64
    o::ComplexF64
    adjoint::Bool
end
HelmholtzStencil{T}(n1, n2, n3, d, o=-1.0) where {T} = HelmholtzStencil{T}((Int(n1), Int(n2), Int(n3)), d, o, false)
Base.size(G::HelmholtzStencil) = (prod(G.n), prod(G.n))
Base.eltype(::HelmholtzStencil{T}) where {T} = T
LinearAlgebra.adjoint(G::HelmholtzStencil{T}) where {T} = HelmholtzStencil{T}(G.n, G.d, G.o, !G.adjoint)
Funicular.panel_capable(::HelmholtzStencil) = true
Funicular.ishermitian_op(G::HelmholtzStencil) = imag(G.d) == 0 && imag(G.o) == 0
function LinearAlgebra.mul!(Y::HIPMatrix{T}, G::HelmholtzStencil{T}, X::HIPMatrix{T}) where {T}
    size(X) == size(Y) && size(X, 1) == size(G, 1) ||
        throw(DimensionMismatch("a $(size(G)) stencil cannot map $(size(X)) onto $(size(Y))"))
    check(ccall((:fhip_stencil_helmholtz3d, LIB[]), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64, Int64, Int64, Float64, Float64, Float64, Float64, Cint,
         Cint, Ptr{Cvoid}),
        Y.ptr, Y.ld, X.ptr, X.ld, G.n[1], G.n[2], G.n[3], size(X, 2), real(G.d), imag(G.d), real(G.o), imag(G.o),
        Cint(G.adjoint), eltcode(T), current_stream()))
    Y
end

struct TridiagonalStencil{T}
    N::Int
    d::ComplexF64
    o::ComplexF64
    diag::Union{Nothing,HIPMatrix{Float64}}
    adjoint::Bool
end
TridiagonalStencil{T}(N, d, o, diag=nothing) where {T} = TridiagonalStencil{T}(Int(N), d, o, diag, false)
Base.size(G::TridiagonalStencil) = (G.N, G.N)
Base.eltype(::TridiagonalStencil{T}) where {T} = T
LinearAlgebra.adjoint(G::TridiagonalStencil{T}) where {T} = TridiagonalStencil{T}(G.N, G.d, G.o, G.diag, !G.adjoint)
Funicular.panel_capable(::TridiagonalStencil) = true
Funicular.ishermitian_op(G::TridiagonalStencil) = imag(G.d) == 0 && imag(G.o) == 0
function LinearAlgebra.mul!(Y::HIPMatrix{T}, G::TridiagonalStencil{T}, X::HIPMatrix{T}) where {T}
    size(X) == size(Y) && size(X, 1) == G.N ||
        throw(DimensionMismatch("a $(size(G)) stencil cannot map $(size(X)) onto $(size(Y))"))
    diag = G.diag === nothing ? C_NULL : G.diag.ptr
    check(ccall((:fhip_stencil_tridiag, LIB[]), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64, Float64, Float64, Float64, Float64, Ptr{Cvoid}, Cint,
         Cint, Ptr{Cvoid}),
        Y.ptr, Y.ld, X.ptr, X.ld, G.N, size(X, 2), real(G.d), imag(G.d), real(G.o), imag(G.o), diag,
        Cint(G.adjoint), eltcode(T), current_stream()))
    Y
end

# Ghost panels generated straight into a device buffer (rows row0+1 ..
# row0+size(dst, 1) of columns col0 .. col0+size(dst, 2)-1), for a hook in the
# ghost tier's materialize! whose default is the reference's host generation
# (src/ghosts.jl:97-122) and an upload. The library's stream is its own
# counter-based one, not Julia's Xoshiro stream, so a GhostPanels matrix has to
# opt in to it: the values are standard complex normals with the same
# contract (they depend on (seed, col) alone), but not the same numbers.
function fill_ghost!(::HIPBackend, dst::HIPMatrix{T}, row0::Integer, col0::Integer, seed::Integer) where {T}
    check(ccall((:fhip_ghost_randn, LIB[]), Cint,
        (Ptr{Cvoid}, Cint, Int64, Int64, Int64, Int64, Int64, UInt64, Ptr{Cvoid}),
        dst.ptr, eltcode(T), dst.ld, row0, size(dst, 1), col0, size(dst, 2), UInt64(seed), current_stream()))
    dst
end

end # module FunicularHIPExt
