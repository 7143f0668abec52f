# HIPBackend.jl -- the MI355X backend of OpenADMIXTURE.jl's GPU branch.
#
# The reference plugs a GPU in through device types and a handful of methods
# that its generic host code calls when `d_cu !== nothing`
# (src/algorithms_inner.jl:185-195, 260-271; src/algorithms_outer.jl:12-45,
# 77-156); src/cuda/*.jl defines them for CUDA.jl.  This file defines the same
# methods for libadmixhip (include/admixhip.h) through ccall, so the
# reference's own files run unchanged:
#
#     using OpenADMIXTURE
#     OpenADMIXTURE.include("/path/to/openadmixture.jl_amd/julia/HIPBackend.jl")
#     d = OpenADMIXTURE.run_admixture("cohort.bed", 8; use_gpu=true)
#
# `Module.include` evaluates the file inside OpenADMIXTURE, as the reference's
# `@require CUDA` block includes src/cuda/*.jl (src/OpenADMIXTURE.jl:11-20).
# run_admixture(...; use_gpu=true) then calls _cu_admixture_base
# (src/driver.jl:150-151), defined below for (AdmixData, SnpLinAlg), and the
# generic update_q!/update_p!/init_em!/admixture_qn! take their GPU branches
# with these types:
#
#   reference call (GPU branch)                    libadmixhip entry point
#   _cu_admixture_base(d, g_la, I, J)              admx_create + admx_upload_bed
#   copyto_sync!([d_cu.q, d_cu.p], [q, p])         admx_set_block (x, side Q / P)
#   copyto_sync!([d.p, d.q], [d_cu.p, d_cu.q])     admx_get_block
#   update_q_cuda!(d_cu, g_cu) / update_p_cuda!    admx_grad_hessian (Q / P)
#   copyto_sync!([XtX, Xtz], [d_cu.XtX_q, ...])    admx_get_hessians
#   em_q!(d_cu, g_cu) / em_p!(d_cu, g_cu)          admx_em_q / admx_em_p
#   d_.p .= d_.p_next ; d_.q .= d_.q_next          admx_copy_block (x <- x_next)
#   loglikelihood(d_cu, g_cu)                      admx_loglikelihood (x)
#
# The QPs, quasi-Newton steps and projections then stay in the reference's
# Julia code on the CPU, as with CUDA.  hip_init_em! and hip_admixture_qn! at
# the end are the device-resident path (QPs, ZAL/LBQN and projections on the
# device, only log-likelihoods cross the bus per iteration), an extra entry
# beside the drop-in one.
#
# Load one GPU backend per session: CUDA.jl's src/cuda/*.jl defines
# _cu_admixture_base and copyto_sync! for the same argument types.

const libadmixhip = get(ENV, "ADMIXHIP_LIB", normpath(joinpath(@__DIR__, "..", "libadmixhip.so")))

const ADMX_OK = Cint(0)
const ADMX_F64 = Cint(0)
const ADMX_F32 = Cint(1)
const ADMX_X = Cint(0)
const ADMX_X_NEXT = Cint(1)
const ADMX_X_NEXT2 = Cint(2)
const ADMX_X_TMP = Cint(3)
const ADMX_SIDE_Q = Cint(0)
const ADMX_SIDE_P = Cint(1)
const ADMX_TUNE_QN_MODE = Cint(4)
const ADMX_TUNE_SKIPMISSING = Cint(5)
const ADMX_QN_ZAL = Int64(0)
const ADMX_QN_LBQN = Int64(1)

struct AdmixHIPError <: Exception
    code::Cint
    msg::String
end
Base.showerror(io::IO, e::AdmixHIPError) = print(io, "libadmixhip error ", e.code, ": ", e.msg)

function admx_check(rc::Cint)
    if rc != ADMX_OK
        msg = unsafe_string(ccall((:admx_last_error, libadmixhip), Cstring, ()))
        throw(AdmixHIPError(rc, msg))
    end
    return nothing
end

# ---------------------------------------------------------------- context

"""
    HipContext(I, J, K, Q, T)

One libadmixhip context (`admx_ctx`): the device mirror of `AdmixData{T}` --
the BED payload in HBM, the four state vectors x, x_next, x_next2, x_tmp
(K x (I+J) each), the Hessian and gradient arrays of both sides, the
quasi-Newton history and the log-likelihood slots.  Freed by its finalizer
or `close!`.
"""
mutable struct HipContext
    ptr::Ptr{Cvoid}
    I::Int
    J::Int
    K::Int
    Q::Int
    function HipContext(I::Integer, J::Integer, K::Integer, Q::Integer, ::Type{T}) where {T}
        dtype = T === Float64 ? ADMX_F64 : T === Float32 ? ADMX_F32 :
                throw(ArgumentError("libadmixhip runs Float64 or Float32 contexts"))
        out = Ref{Ptr{Cvoid}}(C_NULL)
        admx_check(ccall((:admx_create, libadmixhip), Cint, (Ref{Ptr{Cvoid}}, Int64, Int64, Int64, Int64, Cint),
                         out, I, J, K, Q, dtype))
        ctx = new(out[], Int(I), Int(J), Int(K), Int(Q))
        finalizer(close!, ctx)
        return ctx
    end
end

function close!(ctx::HipContext)
    if ctx.ptr != C_NULL
        ccall((:admx_destroy, libadmixhip), Cint, (Ptr{Cvoid},), ctx.ptr)
        ctx.ptr = C_NULL
    end
    return nothing
end

function _live(ctx::HipContext)
    ctx.ptr == C_NULL && throw(ArgumentError("libadmixhip context already closed"))
    return ctx.ptr
end

# ---------------------------------------------------------------- device arrays

"""
    HipArray{T,N}

A device array of a `HipContext`: block `side` (Q: K x I, P: K x J) of state
vector `which`, or (`hess = true`) the context's XtX (N = 3) or Xtz (N = 2) of
`side`.  It supports exactly what the reference's GPU branch does with CuArrays:
`copyto!` to and from host `Array`s, and `dst .= src` between two blocks.
"""
struct HipArray{T,N} <: AbstractArray{T,N}
    ctx::HipContext
    which::Cint
    side::Cint
    hess::Bool
    dims::NTuple{N,Int}
end

Base.size(A::HipArray) = A.dims
Base.getindex(::HipArray, ::Int...) =
    throw(ArgumentError("HipArray has no scalar access; copyto! it into a host Array"))
Base.show(io::IO, A::HipArray{T,N}) where {T,N} =
    print(io, "HipArray{", T, ",", N, "}(", join(A.dims, "x"), A.hess ? ", Hessian" : "", ")")
Base.show(io::IO, ::MIME"text/plain", A::HipArray) = show(io, A)

# contiguous column-major memory, the layout the C ABI reads and writes
function _dense_ptr(A::StridedArray)
    expect = 1
    for (n, s) in zip(size(A), strides(A))
        (n == 1 || s == expect) || throw(ArgumentError("host array must be contiguous column-major memory"))
        expect *= n
    end
    return pointer(A)
end

function Base.copyto!(dst::HipArray{T,N}, src::StridedArray{T,N}) where {T,N}
    size(dst) == size(src) || throw(DimensionMismatch("device $(size(dst)) vs host $(size(src))"))
    dst.hess && throw(ArgumentError("XtX and Xtz are written by the device passes only"))
    GC.@preserve src admx_check(ccall((:admx_set_block, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}),
                                      _live(dst.ctx), dst.which, dst.side, _dense_ptr(src)))
    return dst
end

function Base.copyto!(dst::StridedArray{T,N}, src::HipArray{T,N}) where {T,N}
    size(dst) == size(src) || throw(DimensionMismatch("host $(size(dst)) vs device $(size(src))"))
    if src.hess
        XtX = N == 3 ? _dense_ptr(dst) : C_NULL
        Xtz = N == 2 ? _dense_ptr(dst) : C_NULL
        GC.@preserve dst admx_check(ccall((:admx_get_hessians, libadmixhip), Cint,
                                          (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                                          _live(src.ctx), src.side, XtX, Xtz))
    else
        GC.@preserve dst admx_check(ccall((:admx_get_block, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}),
                                          _live(src.ctx), src.which, src.side, _dense_ptr(dst)))
    end
    return dst
end

function Base.copyto!(dst::HipArray{T,N}, src::HipArray{T,N}) where {T,N}
    (dst.ctx === src.ctx && dst.side == src.side && !dst.hess && !src.hess) ||
        throw(ArgumentError("device copies run between same-side state blocks of one context"))
    admx_check(ccall((:admx_copy_block, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Cint, Cint),
                     _live(dst.ctx), dst.which, src.which, dst.side))
    return dst
end

# `d_.p .= d_.p_next` (src/algorithms_outer.jl:28-29) lowers to
# materialize!(d_.p, broadcasted(identity, d_.p_next)): a device block copy.
function Base.Broadcast.materialize!(dst::HipArray, bc::Base.Broadcast.Broadcasted)
    if bc.f === identity && length(bc.args) == 1 && bc.args[1] isa HipArray
        return copyto!(dst, bc.args[1])
    end
    throw(ArgumentError("HipArray broadcasts only `dst .= src` between device arrays"))
end

# copyto_sync! of src/cuda/transfer.jl:1-7: every libadmixhip copy returns
# after its stream has synchronised.
function copyto_sync!(As_d1::Vector{<:AbstractArray{T}}, As_d2::Vector{<:AbstractArray{T}}) where {T}
    for (A_d1, A_d2) in zip(As_d1, As_d2)
        copyto!(A_d1, A_d2)
    end
end

# ---------------------------------------------------------------- the CuAdmixData counterpart

"""
    HipAdmixData{T}

The fields of `CuAdmixData{T}` (src/cuda/structs.jl:1-31) as blocks of one
context: q, p = blocks of x; q_next, p_next = blocks of x_next; p_tmp = the
p block of x_tmp; XtX_q, Xtz_q, XtX_p, Xtz_p = the context's Hessian arrays.
"""
struct HipAdmixData{T}
    ctx::HipContext
    q::HipArray{T,2}
    q_next::HipArray{T,2}
    p::HipArray{T,2}
    p_next::HipArray{T,2}
    p_tmp::HipArray{T,2}
    XtX_q::HipArray{T,3}
    Xtz_q::HipArray{T,2}
    XtX_p::HipArray{T,3}
    Xtz_p::HipArray{T,2}
end

function HipAdmixData{T}(ctx::HipContext) where {T}
    I, J, K = ctx.I, ctx.J, ctx.K
    blk(which, side) = HipArray{T,2}(ctx, which, side, false, (K, side == ADMX_SIDE_Q ? I : J))
    HipAdmixData{T}(ctx,
                    blk(ADMX_X, ADMX_SIDE_Q), blk(ADMX_X_NEXT, ADMX_SIDE_Q),
                    blk(ADMX_X, ADMX_SIDE_P), blk(ADMX_X_NEXT, ADMX_SIDE_P), blk(ADMX_X_TMP, ADMX_SIDE_P),
                    HipArray{T,3}(ctx, Cint(-1), ADMX_SIDE_Q, true, (K, K, I)),
                    HipArray{T,2}(ctx, Cint(-1), ADMX_SIDE_Q, true, (K, I)),
                    HipArray{T,3}(ctx, Cint(-1), ADMX_SIDE_P, true, (K, K, J)),
                    HipArray{T,2}(ctx, Cint(-1), ADMX_SIDE_P, true, (K, J)))
end

"""
    HipBED

`g_cu` of the GPU branch: the genotypes that `_cu_admixture_base` uploaded
into the context (the payload `g_la.s.data`, ceil(I/4) x J bytes, kept 2-bit
in HBM).
"""
struct HipBED
    ctx::HipContext
end

# _cu_admixture_base (stub src/driver.jl:204-206; CUDA form src/cuda/structs.jl:33-38)
function _cu_admixture_base(d::AdmixData{T}, g_la::SnpLinAlg{T}, I::Int, J::Int) where {T}
    ctx = HipContext(I, J, d.K, d.Q, T)
    admx_check(ccall((:admx_set_tuning, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Int64),
                     ctx.ptr, ADMX_TUNE_SKIPMISSING, Int64(d.skipmissing)))
    data = g_la.s.data
    GC.@preserve data admx_check(ccall((:admx_upload_bed, libadmixhip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64),
                                       ctx.ptr, pointer(data), Int64(size(data, 1))))
    return HipAdmixData{T}(ctx), HipBED(ctx)
end

function _same(d::HipAdmixData, g::HipBED)
    d.ctx === g.ctx || throw(ArgumentError("d_cu and g_cu belong to different contexts"))
    return _live(d.ctx)
end

# src/cuda/runners.jl:2-20
function loglikelihood(d::HipAdmixData{T}, g_cu::HipBED) where {T}
    out = Ref{Float64}(NaN)
    admx_check(ccall((:admx_loglikelihood, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Ref{Float64}),
                     _same(d, g_cu), ADMX_X, out))
    return out[]
end

# src/cuda/runners.jl:22-43: q_next from (q, p), divided by 2J
function em_q!(d::HipAdmixData{T}, g_cu::HipBED; verbose=false) where {T}
    admx_check(ccall((:admx_em_q, libadmixhip), Cint, (Ptr{Cvoid},), _same(d, g_cu)))
    return d.q_next
end

# src/cuda/runners.jl:45-67: p_next = p_tmp / (p_tmp + p_next)
function em_p!(d::HipAdmixData{T}, g_cu::HipBED; verbose=false) where {T}
    admx_check(ccall((:admx_em_p, libadmixhip), Cint, (Ptr{Cvoid},), _same(d, g_cu)))
    return d.p_next
end

# src/cuda/runners.jl:69-111: XtX and Xtz of (d_cu.q, d_cu.p), zeroed then summed
function update_q_cuda!(d::HipAdmixData{T}, g_cu::HipBED) where {T}
    admx_check(ccall((:admx_grad_hessian, libadmixhip), Cint, (Ptr{Cvoid}, Cint), _same(d, g_cu), ADMX_SIDE_Q))
    return nothing
end

function update_p_cuda!(d::HipAdmixData{T}, g_cu::HipBED) where {T}
    admx_check(ccall((:admx_grad_hessian, libadmixhip), Cint, (Ptr{Cvoid}, Cint), _same(d, g_cu), ADMX_SIDE_P))
    return nothing
end

# ---------------------------------------------------------------- device-resident path

"""
    hip_init_em!(d, d_cu, iter)

init_em! (src/algorithms_outer.jl:12-45) with the whole loop on the device:
q and p go up once, `iter` FRAPPE EM steps run there, the state and
`d.ll_new` come back.  Returns the per-step log-likelihoods.
"""
function hip_init_em!(d::AdmixData{T}, d_cu::HipAdmixData{T}, iter::Integer) where {T}
    ctx = _live(d_cu.ctx)
    GC.@preserve d admx_check(ccall((:admx_set_x, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Ptr{T}),
                                    ctx, ADMX_X, pointer(d.x)))
    lls = zeros(Float64, max(iter, 1))
    admx_check(ccall((:admx_init_em, libadmixhip), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}), ctx, iter, lls))
    GC.@preserve d admx_check(ccall((:admx_get_x, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Ptr{T}),
                                    ctx, ADMX_X, pointer(d.x)))
    prev, new = Ref{Float64}(NaN), Ref{Float64}(NaN)
    admx_check(ccall((:admx_get_ll, libadmixhip), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}), ctx, prev, new))
    d.ll_new = new[]
    return lls[1:iter]
end

"""
    hip_admixture_qn!(d, d_cu, iter=1000, rtol=1e-7; mode=:ZAL, fix_p=false, fix_q=false)

admixture_qn! (src/algorithms_outer.jl:63-193) with every iteration on the
device: update_q!/update_p! twice with their QPs and projections, the two
log-likelihoods, the ZAL or LBQN step and the accept rule.  The state goes up
once and comes back at the end; each iteration returns only
(prev, basic, qn, new, reldiff).  Returns the log-likelihood history.
"""
function hip_admixture_qn!(d::AdmixData{T}, d_cu::HipAdmixData{T}, iter::Int=1000, rtol=1e-7;
                           mode=:ZAL, fix_p=false, fix_q=false, iter_count_offset=0) where {T}
    mode in (:ZAL, :LBQN) || throw(ArgumentError("Invalid mode"))
    ctx = _live(d_cu.ctx)
    admx_check(ccall((:admx_set_tuning, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Int64),
                     ctx, ADMX_TUNE_QN_MODE, mode == :ZAL ? ADMX_QN_ZAL : ADMX_QN_LBQN))
    GC.@preserve d begin
        admx_check(ccall((:admx_set_x, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Ptr{T}), ctx, ADMX_X, pointer(d.x)))
        admx_check(ccall((:admx_set_uv, libadmixhip), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}), ctx,
                         pointer(d.U), pointer(d.V)))
    end
    if isnan(d.ll_new)
        out = Ref{Float64}(NaN)
        admx_check(ccall((:admx_loglikelihood, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Ref{Float64}), ctx, ADMX_X, out))
        d.ll_new = out[]
    end
    admx_check(ccall((:admx_set_ll, libadmixhip), Cint, (Ptr{Cvoid}, Float64, Float64), ctx, d.ll_prev, d.ll_new))
    println("initial ll: ", d.ll_new)
    llhist = [d.ll_new]
    lls = zeros(Float64, 5)
    converged = false
    i = iter_count_offset
    for outer i in (iter_count_offset + 1):iter
        t = @elapsed admx_check(ccall((:admx_qn_iteration, libadmixhip), Cint,
                                      (Ptr{Cvoid}, Int64, Cint, Cint, Ptr{Float64}), ctx, i, fix_q, fix_p, lls))
        println("Iteration $i ($t sec): LogLikelihood = $(lls[4]), reldiff = $(lls[5])")
        println("    LogLikelihoods: Previous = $(lls[1]) QN = $(lls[3]), Basic = $(lls[2])")
        push!(llhist, lls[4])
        if lls[5] < rtol
            converged = true
            break
        end
    end
    println(converged ? "Main algorithm converged in $i iterations." :
            "Main algorithm failed to converge after $i iterations.")
    GC.@preserve d begin
        for (which, arr) in ((ADMX_X, d.x), (ADMX_X_NEXT, d.x_next), (ADMX_X_NEXT2, d.x_next2), (ADMX_X_TMP, d.x_tmp))
            admx_check(ccall((:admx_get_x, libadmixhip), Cint, (Ptr{Cvoid}, Cint, Ptr{T}), ctx, which, pointer(arr)))
        end
        admx_check(ccall((:admx_get_uv, libadmixhip), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}), ctx,
                         pointer(d.U), pointer(d.V)))
    end
    prev, new = Ref{Float64}(NaN), Ref{Float64}(NaN)
    admx_check(ccall((:admx_get_ll, libadmixhip), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}), ctx, prev, new))
    d.ll_prev, d.ll_new = prev[], new[]
    return llhist
end
