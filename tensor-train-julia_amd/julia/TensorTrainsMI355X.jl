"""
    TensorTrainsMI355X

Julia binding of libttround (include/ttround.h): the MI355X rounding path of
TensorTrains behind `ccall`.  Loading this module after `TensorTrains` adds
GPU methods with the reference's names and keyword defaults
(src/tt_rounding.jl:146-167, src/tt_randtools.jl:112-283,
src/sketches.jl:175-315,317-472,645-817,930-1001,1047-1070,1126-1156,
src/tt_randtools.jl:552-638,
src/tt_operations.jl:81-112,170-194,252-258,273-290,
302-421, the NTuple methods src/tt_randtools.jl:364-473, src/sketches.jl:474-615,
819-909,1003-1040, and expand_basis! src/tt_solvers.jl:354-455); the
solver callers (tt_gmres, arnoldi, davidson, sketched_rr, sketched_rayleigh_ritz,
als/mals/dmrg via orthogonalize) need no change.  Every call uploads the cores (column-major (n, r, r') arrays,
exactly TTvector.ttv_vec), runs on the GPU and copies the result back into new
Julia arrays.

Sketch draws: by default (`reference_draws=true`) the binding seeds
`Random.default_rng()` with `seed` and calls the reference's own
`generate_sketch_blocks` core by core in the reference's order
(src/sketches.jl:177-178,222-225,279-281), then hands those blocks to the library
(`ttr_rand_opts.blocks`), so a given `seed` gives the reference's sketch and
result; `reference_draws=false` uses the library's counter-based draws instead
(no host work).  Every method accepts the reference's `timer` keyword (TimerOutputs
timing of the host code only; the GPU work is one call).

Status codes map to the reference's exceptions:
TTR_INVALID_ARG -> ArgumentError, TTR_DIM_MISMATCH -> DimensionMismatch,
TTR_NUMERICAL_FAILURE -> LinearAlgebra.LAPACKException-like ErrorException,
others -> ErrorException with ttr_last_error.
"""
module TensorTrainsMI355X

using TensorTrains
using Random
import TensorTrains: tt_rounding, tt_rounding!, orthogonalize, ttrand_rounding, tt_recursive_sketch, expand_basis!,
                     tt_sketch, tt_combined_sketch, stta, stta_sketch, dot_par, dot_operator, partial_contraction
import LinearAlgebra
import LinearAlgebra: dot
import Base: *

const lib = get(ENV, "TTROUND_LIB", joinpath(@__DIR__, "..", "ttround", "libttround.so"))

const TTR_OK = 0
const _handle = Ref{Ptr{Cvoid}}(C_NULL)

function handle()
    if _handle[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:ttr_create, lib), Cint, (Ref{Ptr{Cvoid}}, Cint), h, 0)
        st == TTR_OK || error("ttr_create failed with status $st")
        _handle[] = h[]
        atexit(() -> ccall((:ttr_destroy, lib), Cint, (Ptr{Cvoid},), _handle[]))
    end
    return _handle[]
end

"""
    trim()

Return the handle's idle device workspace and its cached TT storage to the runtime
(`ttr_trim`); live results stay valid.  For callers that move between workloads of very
different sizes.  No counterpart in `TensorTrains` (Julia's GC owns its buffers).
"""
trim() = check(ccall((:ttr_trim, lib), Cint, (Ptr{Cvoid},), handle()))

function check(st)
    st == TTR_OK && return
    msg = unsafe_string(ccall((:ttr_last_error, lib), Cstring, (Ptr{Cvoid},), handle()))
    st == 1 && throw(ArgumentError(msg))
    st == 2 && throw(DimensionMismatch(msg))
    st == 4 && error("LAPACK-like failure (libttround): $msg")
    error("libttround status $st: $msg")
end

function upload(x::TTvector{Float64,N}) where {N}
    id = Ref{Int64}(0)
    ptrs = [pointer(c) for c in x.ttv_vec]
    GC.@preserve x begin
        check(ccall((:ttr_tt_upload, lib), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Float64}}, Ref{Int64}),
                    handle(), N, collect(Int64, x.ttv_dims), x.ttv_rks, x.ttv_ot, ptrs, id))
    end
    return id[]
end

function download(id::Int64, ::Val{N}) where {N}
    dims = zeros(Int64, N); rks = zeros(Int64, N + 1); ot = zeros(Int64, N)
    check(ccall((:ttr_tt_info, lib), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}),
                handle(), id, C_NULL, dims, rks, ot))
    vec = Vector{Array{Float64,3}}(undef, N)
    for k in 1:N
        vec[k] = Array{Float64,3}(undef, dims[k], rks[k], rks[k+1])
        check(ccall((:ttr_tt_download_core, lib), Cint, (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}),
                    handle(), id, k - 1, vec[k]))
    end
    check(ccall((:ttr_tt_free, lib), Cint, (Ptr{Cvoid}, Int64), handle(), id))
    return TTvector{Float64,N}(N, vec, Tuple(dims), rks, ot)
end

free(id) = check(ccall((:ttr_tt_free, lib), Cint, (Ptr{Cvoid}, Int64), handle(), id))

# entrywise product of TTs on the device (ttr_hadamard); first_fastest: the first
# factor's rank index runs fastest (the NTuple methods), else the last (x * y)
function _hadamard(y, first_fastest::Bool)
    ids = Int64[upload(t) for t in y]
    out = Ref{Int64}(0)
    check(ccall((:ttr_hadamard, lib), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}, Int32, Ref{Int64}),
                handle(), length(ids), ids, Int32(first_fastest ? 1 : 0), out))
    foreach(free, ids)
    return out[]
end

# x * y (src/tt_operations.jl:273-290): slices kron(x_k[i,:,:], y_k[i,:,:])
*(x::TTvector{Float64,N}, y::TTvector{Float64,N}) where {N} = download(_hadamard((x, y), false), Val(N))

# an NTuple argument of the sketches: the product, built on the device at upload
struct _Product{N}
    factors::Vector{TTvector{Float64,N}}
    ttv_dims::NTuple{N,Int}
    ttv_rks::Vector{Int}
end
_Product(y::NTuple{M,TTvector{Float64,N}}) where {M,N} =
    _Product{N}(collect(y), y[1].ttv_dims, [prod(t.ttv_rks[k] for t in y) for k in 1:N+1])
upload(p::_Product) = _hadamard(p.factors, true)
_bond_shape(A::TTvector, k) = (A.ttv_rks[k],)
_bond_shape(A::_Product, k) = Tuple(t.ttv_rks[k] for t in A.factors)

# ---------------------------------------------------------------- deterministic

function tt_rounding(x::TTvector{Float64,N}; tol=1e-12, rmax=2^14, direction=:left) where {N}
    direction in (:left, :right) || error("Invalid direction. Use direction=:right or direction=:left")
    id = upload(x); out = Ref{Int64}(0)
    check(ccall((:ttr_round_svd, lib), Cint, (Ptr{Cvoid}, Int64, Ref{Int64}, Cdouble, Int64, Cint, Ref{Int64}),
                handle(), 1, id, tol, rmax, direction == :left ? 0 : 1, out))
    free(id)
    return download(out[], Val(N))
end

# The reference rounds x in place only when it is already orthogonal in the sweep's
# sense; otherwise it rounds an orthogonalised copy and x stays as it was
# (src/tt_rounding.jl:158-167)
function tt_rounding!(x::TTvector{Float64,N}; tol=1e-12, rmax=2^14, direction=:left) where {N}
    orth = direction == :left ? TensorTrains.is_leftorthogonal(x) : TensorTrains.is_rightorthogonal(x)
    y = tt_rounding(x; tol=tol, rmax=rmax, direction=direction)
    orth || return
    copyto!(x.ttv_vec, y.ttv_vec); copyto!(x.ttv_rks, y.ttv_rks); copyto!(x.ttv_ot, y.ttv_ot)
    return
end

# tt_rounding(A::TToperator; tol, rmax) (src/tt_rounding.jl:275-277): the rounding of the
# (n^2, R, R') reshape of the cores
tt_rounding(A::TToperator{Float64,N}; tol=1e-12, rmax=2^14) where {N} =
    TensorTrains.ttv_to_tto(tt_rounding(TensorTrains.tto_to_ttv(A); tol=tol, rmax=rmax))

function orthogonalize(x::TTvector{Float64,N}; i=1::Int) where {N}
    id = upload(x); out = Ref{Int64}(0)
    check(ccall((:ttr_orthogonalize, lib), Cint, (Ptr{Cvoid}, Int64, Ref{Int64}, Int64, Ref{Int64}),
                handle(), 1, id, i, out))
    free(id)
    return download(out[], Val(N))
end

# ---------------------------------------------------------------- inner products

# dot(A, B) (src/tt_operations.jl:302-321); dot_par (:339-370) is the same number
function dot(A::TTvector{Float64,N}, B::TTvector{Float64,N}) where {N}
    A.ttv_dims == B.ttv_dims || throw(DimensionMismatch("TT dimensions are not compatible"))
    a, b = upload(A), upload(B); out = Ref{Float64}(0.0)
    check(ccall((:ttr_dot, lib), Cint, (Ptr{Cvoid}, Int64, Ref{Int64}, Ref{Int64}, Ref{Float64}),
                handle(), 1, a, b, out))
    free(a); free(b)
    return out[]
end
dot_par(A::TTvector{Float64,N}, B::TTvector{Float64,N}) where {N} = dot(A, B)

# partial_contraction(A, B; reverse) (src/sketches.jl:1047-1070)
function partial_contraction(A::TTvector{Float64,N}, B::TTvector{Float64,N}; reverse=true) where {N}
    A.ttv_dims == B.ttv_dims || throw(DimensionMismatch("TT dimensions are not compatible"))
    W = [zeros(Float64, A.ttv_rks[i], B.ttv_rks[i]) for i in 1:N+1]
    a, b = upload(A), upload(B)
    GC.@preserve W begin
        check(ccall((:ttr_partial_contraction, lib), Cint, (Ptr{Cvoid}, Int64, Int64, Cint, Ptr{Ptr{Float64}}),
                    handle(), a, b, reverse ? 1 : 0, [pointer(w) for w in W]))
    end
    free(a); free(b)
    return W
end

# dot_operator(psi, H, phi) = <psi|H|phi> (src/tt_operations.jl:385-421)
function dot_operator(ψ::TTvector{Float64,N}, H::TToperator{Float64,N}, φ::TTvector{Float64,N}) where {N}
    a, op, b = upload(ψ), upload(H), upload(φ); out = Ref{Float64}(0.0)
    check(ccall((:ttr_dot_operator, lib), Cint,
                (Ptr{Cvoid}, Int64, Ref{Int64}, Ref{Int64}, Ref{Int64}, Ref{Float64}), handle(), 1, a, op, b, out))
    free(a); free(op); free(b)
    return out[]
end

# ---------------------------------------------------------------- randomized

struct RandItem
    nterms::Int64
    alpha::Ptr{Float64}
    term_ids::Ptr{Int64}
    op_id::Int64
    rks::Ptr{Int64}
end

struct RandOpts
    orthogonal::Int32
    sketch_prec::Int32
    block_rks::Int64
    seed::UInt64
    p::Ptr{Int64}
    oversampling::Float64
    forward::Int32                    # 0: reverse = true (right-to-left), 1: reverse = false
    blocks::Ptr{Ptr{Float64}}         # caller's blocks per core, or C_NULL (library draws)
    blocks_right::Ptr{Ptr{Float64}}   # stta: the right sketch's blocks
    sketch_accum::Int32               # 1: only the blocks in sketch precision, contractions FP64
end

# sketch precision: Float64, Float32, or BF16 (BFloat16s.BFloat16, or the symbol :bf16)
const PREC = Dict{Any,Int32}(Float64 => Int32(0), Float32 => Int32(1), :bf16 => Int32(2))
_prec(T) = haskey(PREC, T) ? PREC[T] :
           (T isa DataType && nameof(T) == :BFloat16) ? Int32(2) :
           throw(ArgumentError("sketch precision must be Float64, Float32 or BFloat16"))

_opts(orthogonal, T, block_rks, seed; p=Ptr{Int64}(C_NULL), forward=false, blocks=Ptr{Ptr{Float64}}(C_NULL),
      blocks_right=Ptr{Ptr{Float64}}(C_NULL), accum64::Bool=false) =
    Ref(RandOpts(orthogonal ? 1 : 0, _prec(T), block_rks, seed % UInt64, p, 1.0, Int32(forward ? 1 : 0), blocks,
                 blocks_right, Int32(accum64 ? 1 : 0)))

# The reference's sketch blocks for `seed`: default_rng seeded, then
# generate_sketch_blocks core by core in the reference's order (right to left for the
# reverse sketch, left to right for the forward one), each array as it returns it
# (src/sketches.jl:177-178,222-225,279-281,700-715).  Drawn in the element type TD of
# the reference's call (Float64 for ttrand_rounding of Float64 TTs) and passed as FP64.
function _reference_blocks(::Type{TD}, dims, bv, pp, orthogonal::Bool, seed, reverse::Bool) where {TD}
    rng = Random.default_rng()
    Random.seed!(rng, seed)
    N = length(dims)
    B = Vector{Array{Float64,4}}(undef, N)
    for k in (reverse ? (N:-1:1) : (1:N))
        S = reverse ? TensorTrains.generate_sketch_blocks(rng, TD, bv[k+1], dims[k], bv[k], pp[k], orthogonal) :
                      TensorTrains.generate_sketch_blocks(rng, TD, bv[k], dims[k], bv[k+1], pp[k+1], orthogonal)
        B[k] = Array{Float64,4}(S)
    end
    return B
end

# the reverse sketch's structure of the default block counts (src/sketches.jl:186-198)
function _reverse_structure(dims, rks, block_rks, N; p=0, oversampling=1)
    bv = _block_ranks(dims, block_rks, N)
    pp = p == 0 ? TensorTrains.compute_sketch_blocks_heuristic(collect(Int, rks), bv, N; reverse=true) : collect(Int, p)
    oversampling != 1 && (pp[1:N] .= ceil.(Int, oversampling .* pp[1:N]))
    return bv, pp
end

function upload(A::TToperator{Float64,N}) where {N}
    id = Ref{Int64}(0)
    ptrs = [pointer(c) for c in A.tto_vec]
    GC.@preserve A begin
        check(ccall((:ttr_tto_upload, lib), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Float64}}, Ref{Int64}),
                    handle(), N, collect(Int64, A.tto_dims), collect(Int64, A.tto_rks), ptrs, id))
    end
    return id[]
end

# alpha_1 (A) y_1 + sum_{j>1} alpha_j y_j; with A given the product is never formed
function rand_round(α::Vector{Float64}, A, y::Vector{TTvector{Float64,N}}, rks;
                    orthogonal=true, seed=1234, block_rks::Int=N, sketch_type=Float64, reference_draws::Bool=true,
                    sketch_accum64::Bool=false, timer=nothing) where {N}
    ids = [upload(t) for t in y]
    op = A === nothing ? -1 : upload(A)
    r = collect(Int64, rks)
    out = Ref{Int64}(0)
    B = Array{Float64,4}[]
    if reference_draws  # the draws of tt_recursive_sketch(Float64, y, rks; seed) (src/tt_randtools.jl:117)
        bv, pp = _reverse_structure(y[1].ttv_dims, rks, block_rks, N)
        B = _reference_blocks(Float64, y[1].ttv_dims, bv, pp, orthogonal, seed, true)
    end
    bptr = [pointer(b) for b in B]
    GC.@preserve α ids r B bptr begin
        item = Ref(RandItem(length(y), pointer(α), pointer(ids), op, pointer(r)))
        opts = _opts(orthogonal, sketch_type, block_rks, seed;
                     blocks=reference_draws ? pointer(bptr) : Ptr{Ptr{Float64}}(C_NULL), accum64=sketch_accum64)
        check(ccall((:ttr_round_rand, lib), Cint, (Ptr{Cvoid}, Int64, Ref{RandItem}, Ref{RandOpts}, Ref{Int64}),
                    handle(), 1, item, opts, out))
    end
    foreach(free, ids)
    op >= 0 && free(op)
    return download(out[], Val(N))
end

ttrand_rounding(α::Vector{Float64}, y::Vector{TTvector{Float64,N}}, rks=TensorTrains.default_rank_heuristic(y);
                kwargs...) where {N} = rand_round(α, nothing, y, rks; kwargs...)

ttrand_rounding(α::Vector{Float64}, A::TToperator{Float64,N}, y::Vector{TTvector{Float64,N}},
                rks=TensorTrains.default_rank_heuristic(y); kwargs...) where {N} = rand_round(α, A, y, rks; kwargs...)
ttrand_rounding(α::Vector{Float64}, A::TToperator{Float64,N}, y::Vector{TTvector{Float64,N}}, rmax::Int;
                kwargs...) where {N} = rand_round(α, A, y, vcat(1, fill(rmax, N - 1), 1); kwargs...)

# A y - b (src/tt_randtools.jl:169-214)
ttrand_rounding(A::TToperator{Float64,N}, y::TTvector{Float64,N}, b::TTvector{Float64,N},
                rks=TensorTrains.default_rank_heuristic(y); kwargs...) where {N} =
    rand_round([1.0, -1.0], A, [y, b], rks; kwargs...)

ttrand_rounding(y::TTvector{Float64,N}, rks=TensorTrains.default_rank_heuristic(y); kwargs...) where {N} =
    ttrand_rounding([1.0], [y], rks; kwargs...)
ttrand_rounding(y::TTvector{Float64,N}, rmax::Int; kwargs...) where {N} =
    ttrand_rounding(y, vcat(1, fill(rmax, N - 1), 1); kwargs...)
ttrand_rounding(α::Vector{Float64}, y::Vector{TTvector{Float64,N}}, rmax::Int; kwargs...) where {N} =
    ttrand_rounding(α, y, vcat(1, fill(rmax, N - 1), 1); kwargs...)

# ------------------------------------------------------------------ sketches (reverse = true only)

# block ranks of the reverse sketch (src/sketches.jl:186-191)
function _block_ranks(dims, block_rks::Int, N::Int)
    bv = ones(Int, N + 1)
    bv[1:N] .= block_rks
    for k = N:-1:1
        bv[k] = min(bv[k], dims[k] * bv[k + 1])
    end
    return bv
end

# tt_recursive_sketch(T, [H,] A, rks; reverse=true, ...) (src/sketches.jl:175-240,300-303,317-459)
function _recursive_sketch(::Type{T}, H, A::Union{TTvector{Float64,N},_Product{N}}, rks::Vector{Int};
                           orthogonal=true, reverse=true,
                           seed=1234, block_rks::Int=N, p=0, oversampling=1, reference_draws::Bool=true,
                           timer=nothing) where {T,N}
    bv = reverse ? _block_ranks(A.ttv_dims, block_rks, N) : _block_ranks_fwd(A.ttv_dims, block_rks, N)
    pp = p == 0 ? TensorTrains.compute_sketch_blocks_heuristic(rks, bv, N; reverse=reverse) : collect(Int, p)
    # the forward operator sketch takes no oversampling (src/sketches.jl:399-402)
    if oversampling != 1 && (reverse || H === nothing)
        rng = reverse ? (1:N) : (2:N+1)
        pp[rng] .= ceil.(Int, oversampling .* pp[rng])
    end
    s = bv .* pp
    Hr = H === nothing ? ones(Int, N + 1) : collect(Int, H.tto_rks)
    W = [zeros(Float64, A.ttv_rks[k] * Hr[k], s[k]) for k in 1:N+1]
    # the sketch of a product runs factor by factor in either direction
    # (ttr_sketch_tuple), the product never formed
    factored = A isa _Product && length(A.factors) > 1
    id = factored ? -1 : upload(A)
    fids = factored ? Int64[upload(t) for t in A.factors] : Int64[]
    op = H === nothing ? -1 : upload(H)
    r, pv, srk = collect(Int64, rks), collect(Int64, pp), zeros(Int64, N + 1)
    B = reference_draws ? _reference_blocks(T, A.ttv_dims, bv, pp, orthogonal, seed, reverse) : Array{Float64,4}[]
    bptr = [pointer(b) for b in B]
    GC.@preserve W r pv B bptr fids begin
        ptrs = [pointer(w) for w in W]
        opts = _opts(orthogonal, T, block_rks, seed; p=pointer(pv), forward=!reverse,
                     blocks=reference_draws ? pointer(bptr) : Ptr{Ptr{Float64}}(C_NULL))
        if factored
            check(ccall((:ttr_sketch_tuple, lib), Cint,
                        (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ref{RandOpts}, Ptr{Ptr{Float64}}, Ptr{Int64}),
                        handle(), length(fids), fids, r, opts, ptrs, srk))
        elseif H === nothing
            check(ccall((:ttr_sketch, lib), Cint,
                        (Ptr{Cvoid}, Int64, Ptr{Int64}, Ref{RandOpts}, Ptr{Ptr{Float64}}, Ptr{Int64}),
                        handle(), id, r, opts, ptrs, srk))
        else
            check(ccall((:ttr_sketch_op, lib), Cint,
                        (Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ref{RandOpts}, Ptr{Ptr{Float64}}, Ptr{Int64}),
                        handle(), op, id, r, opts, ptrs, srk))
        end
    end
    id >= 0 && free(id)
    foreach(free, fids)
    op >= 0 && free(op)
    H === nothing && return [reshape(W[k], _bond_shape(A, k)..., s[k]) for k in 1:N+1], collect(Int, srk)
    return [reshape(W[k], A.ttv_rks[k], Hr[k], s[k]) for k in 1:N+1], collect(Int, srk)
end

# block ranks of the forward sketch (src/sketches.jl:244-248)
function _block_ranks_fwd(dims, block_rks::Int, N::Int)
    bv = ones(Int, N + 1)
    bv[2:N+1] .= block_rks
    for k = 1:N
        bv[k + 1] = min(bv[k + 1], dims[k] * bv[k])
    end
    return bv
end

_rks_of(rmax::Int, N) = (r = fill(rmax, N + 1); r[N + 1] = 1; r)
_rks_of(rks::Vector{Int}, N) = rks

_rks_of_fwd(rmax::Int, N) = (r = fill(rmax, N + 1); r[1] = 1; r)
_rks_of_fwd(rks::Vector{Int}, N) = rks
# the rank vector of an rmax in the direction the keywords ask for (src/sketches.jl:955-958)
_rks_dir(rks_or_rmax, N, kw) = get(kw, :reverse, true) ? _rks_of(rks_or_rmax, N) : _rks_of_fwd(rks_or_rmax, N)

tt_recursive_sketch(::Type{T}, A::TTvector{Float64,N}, rks_or_rmax; reverse=true,
                    kwargs...) where {T<:Union{Float64,Float32},N} =
    _recursive_sketch(T, nothing, A, reverse ? _rks_of(rks_or_rmax, N) : _rks_of_fwd(rks_or_rmax, N);
                      reverse=reverse, kwargs...)
tt_recursive_sketch(::Type{T}, H::TToperator{Float64,N}, A::TTvector{Float64,N}, rks_or_rmax; reverse=true,
                    kwargs...) where {T<:Union{Float64,Float32},N} =
    _recursive_sketch(T, H, A, reverse ? _rks_of(rks_or_rmax, N) : _rks_of_fwd(rks_or_rmax, N);
                      reverse=reverse, kwargs...)
tt_recursive_sketch(A::TTvector{Float64,N}, rks_or_rmax; kwargs...) where {N} =
    tt_recursive_sketch(Float64, A, rks_or_rmax; kwargs...)
tt_recursive_sketch(H::TToperator{Float64,N}, A::TTvector{Float64,N}, rks_or_rmax; kwargs...) where {N} =
    tt_recursive_sketch(Float64, H, A, rks_or_rmax; kwargs...)

# tt_sketch(T, [H,] A, s; reverse=true, ...) (src/sketches.jl:645-817): boundary sketch only
function _tt_sketch(::Type{T}, H, A::Union{TTvector{Float64,N},_Product{N}}, s::Int; reverse=true, orthogonal=true,
                    seed=1234, block_rks::Int=N, reference_draws::Bool=true, timer=nothing) where {T,N}
    reverse || return _tt_sketch_fwd(T, H, A, s; orthogonal=orthogonal, seed=seed, block_rks=block_rks,
                                     reference_draws=reference_draws)
    # the same block count p = cld(s, bv_1) at every core (src/sketches.jl:665-668)
    bv = _block_ranks(A.ttv_dims, block_rks, N)
    pc = cld(s, bv[1])
    B = reference_draws ? _reference_blocks(T, A.ttv_dims, bv, fill(pc, N + 1), orthogonal, seed, true) :
                          Array{Float64,4}[]
    bptr = [pointer(b) for b in B]
    id = upload(A)
    op = H === nothing ? -1 : upload(H)
    sa = Ref{Int64}(0)
    ids, ops = Int64[id], Int64[op]
    out = Float64[]
    GC.@preserve B bptr begin
        opts = _opts(orthogonal, T, block_rks, seed; blocks=reference_draws ? pointer(bptr) : Ptr{Ptr{Float64}}(C_NULL))
        check(ccall((:ttr_tt_sketch, lib), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Ref{RandOpts}, Ptr{Float64}, Int64, Ref{Int64}),
                    handle(), 1, ids, ops, s, opts, C_NULL, 0, sa))
        out = zeros(Float64, sa[])
        check(ccall((:ttr_tt_sketch, lib), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Ref{RandOpts}, Ptr{Float64}, Int64, Ref{Int64}),
                    handle(), 1, ids, ops, s, opts, out, length(out), sa))
    end
    free(id)
    op >= 0 && free(op)
    return out, Int(sa[])
end

# the boundary sketch of a product: its factored recursive sketch with the same
# block count p = cld(s, bv_1) at every core (src/sketches.jl:819-900)
function _tt_sketch(::Type{T}, H::Nothing, A::_Product{N}, s::Int; reverse=true, orthogonal=true, seed=1234,
                    block_rks::Int=N, reference_draws::Bool=true, timer=nothing) where {T,N}
    reverse || return _tt_sketch_fwd(T, nothing, A, s; orthogonal=orthogonal, seed=seed, block_rks=block_rks,
                                     reference_draws=reference_draws)
    bv = _block_ranks(A.ttv_dims, block_rks, N)
    pc = cld(s, bv[1])
    W, _ = _recursive_sketch(T, nothing, A, ones(Int, N + 1); orthogonal=orthogonal, reverse=true, seed=seed,
                             block_rks=block_rks, p=vcat(fill(pc, N), 1), reference_draws=reference_draws)
    return vec(W[1]), bv[1] * pc
end

# tt_sketch(...; reverse=false) (src/sketches.jl:679-694,796-811,894-900): the forward
# recursive sketch with p = cld(s, bv_{N+1}) blocks at every core (the same draws), W_{N+1} kept
function _tt_sketch_fwd(::Type{T}, H, A, s::Int; orthogonal=true, seed=1234, block_rks::Int=length(A.ttv_dims),
                        reference_draws::Bool=true) where {T}
    N = length(A.ttv_dims)
    bv = _block_ranks_fwd(A.ttv_dims, block_rks, N)
    pc = cld(s, bv[N + 1])
    W, _ = _recursive_sketch(T, H, A, ones(Int, N + 1); orthogonal=orthogonal, reverse=false, seed=seed,
                             block_rks=block_rks, p=vcat(1, fill(pc, N)), reference_draws=reference_draws)
    return vec(W[N + 1]), bv[N + 1] * pc
end

tt_sketch(::Type{T}, A::TTvector{Float64,N}, s::Int; kwargs...) where {T<:Union{Float64,Float32},N} =
    _tt_sketch(T, nothing, A, s; kwargs...)
tt_sketch(::Type{T}, H::TToperator{Float64,N}, A::TTvector{Float64,N}, s::Int;
          kwargs...) where {T<:Union{Float64,Float32},N} = _tt_sketch(T, H, A, s; kwargs...)
tt_sketch(A::TTvector{Float64,N}, s::Int; kwargs...) where {N} = tt_sketch(Float64, A, s; kwargs...)
tt_sketch(H::TToperator{Float64,N}, A::TTvector{Float64,N}, s::Int; kwargs...) where {N} =
    tt_sketch(Float64, H, A, s; kwargs...)

# tt_combined_sketch(T, [H,] A, rks_or_rmax, s; reverse=true, ...) (src/sketches.jl:930-1001)
function _combined_sketch(::Type{T}, H, A::Union{TTvector{Float64,N},_Product{N}}, rks::Vector{Int}, s::Int;
                          reverse=true,
                          orthogonal=true, seed=1234, block_rks::Int=N, reference_draws::Bool=true,
                          timer=nothing) where {T,N}
    b, inner = reverse ? (1, 2) : (N + 1, N)
    rks = copy(rks)
    rks[b] = max(rks[b], min(s, rks[inner]))
    W, sketch_rks = _recursive_sketch(T, H, A, rks; orthogonal=orthogonal, reverse=reverse, seed=seed,
                                      block_rks=block_rks, reference_draws=reference_draws)
    sketch = vec(W[b])
    s_rec = sketch_rks[b]
    if s > s_rec
        extra, _ = _tt_sketch(T, H, A, s - s_rec; reverse=reverse, orthogonal=orthogonal, seed=seed + 1,
                              block_rks=block_rks, reference_draws=reference_draws)
        s = s_rec + length(extra)
        sketch = [sqrt(length(sketch) / s) .* sketch; sqrt(length(extra) / s) .* extra]
    end
    W[b] = H === nothing ? ones(Float64, ntuple(_ -> 1, length(_bond_shape(A, b)) + 1)...) : ones(Float64, 1, 1, 1)
    sketch_rks[b] = 1
    # a TT's sketches as Vector{Matrix{Float64}}, the type the reference's solvers keep
    # them in (W_B_window::Vector{Vector{Matrix{T}}}, src/tt_solvers.jl:354-357)
    (H === nothing && A isa TTvector) && return sketch, Matrix{Float64}[w for w in W], sketch_rks
    return sketch, W, sketch_rks
end

# NTuple forms (src/tt_randtools.jl:364-473, src/sketches.jl:474-615,819-909,1003-1040): the
# entrywise product of the tuple with the first factor's rank index fastest, never formed
# (ttr_round_rand_tuple, ttr_sketch_tuple; the forward sketch forms it on the device);
# target ranks are required, as in the reference
function ttrand_rounding(y::NTuple{M,TTvector{Float64,N}}, rks::Vector{Int}; orthogonal=true, seed=1234,
                         block_rks::Int=N, sketch_type=Float64, reference_draws::Bool=true, timer=nothing) where {M,N}
    M == 1 && return ttrand_rounding(y[1], rks; orthogonal=orthogonal, seed=seed, block_rks=block_rks,
                                     sketch_type=sketch_type, reference_draws=reference_draws)
    # the product's sketch uses the draws of a single TT with the same dims (src/sketches.jl:474-540)
    B = Array{Float64,4}[]
    if reference_draws
        bv, pp = _reverse_structure(y[1].ttv_dims, rks, block_rks, N)
        B = _reference_blocks(Float64, y[1].ttv_dims, bv, pp, orthogonal, seed, true)
    end
    bptr = [pointer(b) for b in B]
    ids = Int64[upload(t) for t in y]
    nf, r = Int64[M], collect(Int64, rks)
    out = Ref{Int64}(0)
    GC.@preserve ids nf r B bptr begin
        opts = _opts(orthogonal, sketch_type, block_rks, seed;
                     blocks=reference_draws ? pointer(bptr) : Ptr{Ptr{Float64}}(C_NULL))
        check(ccall((:ttr_round_rand_tuple, lib), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ref{RandOpts}, Ref{Int64}),
                    handle(), 1, nf, ids, r, opts, out))
    end
    foreach(free, ids)
    return download(out[], Val(N))
end
ttrand_rounding(y::NTuple{M,TTvector{Float64,N}}, rmax::Int; kwargs...) where {M,N} =
    ttrand_rounding(y, vcat(1, fill(rmax, N - 1), 1); kwargs...)

tt_recursive_sketch(::Type{T}, A::NTuple{M,TTvector{Float64,N}}, rks_or_rmax; reverse=true,
                    kwargs...) where {T<:Union{Float64,Float32},M,N} =
    _recursive_sketch(T, nothing, _Product(A), reverse ? _rks_of(rks_or_rmax, N) : _rks_of_fwd(rks_or_rmax, N);
                      reverse=reverse, kwargs...)
tt_recursive_sketch(A::NTuple{M,TTvector{Float64,N}}, rks_or_rmax; kwargs...) where {M,N} =
    tt_recursive_sketch(Float64, A, rks_or_rmax; kwargs...)
tt_sketch(::Type{T}, A::NTuple{M,TTvector{Float64,N}}, s::Int; kwargs...) where {T<:Union{Float64,Float32},M,N} =
    _tt_sketch(T, nothing, _Product(A), s; kwargs...)
tt_sketch(A::NTuple{M,TTvector{Float64,N}}, s::Int; kwargs...) where {M,N} = tt_sketch(Float64, A, s; kwargs...)
tt_combined_sketch(::Type{T}, A::NTuple{M,TTvector{Float64,N}}, rks_or_rmax, s::Int;
                   kwargs...) where {T<:Union{Float64,Float32},M,N} =
    _combined_sketch(T, nothing, _Product(A), _rks_dir(rks_or_rmax, N, kwargs), s; kwargs...)
tt_combined_sketch(A::NTuple{M,TTvector{Float64,N}}, rks_or_rmax, s::Int; kwargs...) where {M,N} =
    tt_combined_sketch(Float64, A, rks_or_rmax, s; kwargs...)

tt_combined_sketch(::Type{T}, A::TTvector{Float64,N}, rks_or_rmax, s::Int;
                   kwargs...) where {T<:Union{Float64,Float32},N} =
    _combined_sketch(T, nothing, A, _rks_dir(rks_or_rmax, N, kwargs), s; kwargs...)
tt_combined_sketch(::Type{T}, H::TToperator{Float64,N}, A::TTvector{Float64,N}, rks_or_rmax, s::Int;
                   kwargs...) where {T<:Union{Float64,Float32},N} =
    _combined_sketch(T, H, A, _rks_dir(rks_or_rmax, N, kwargs), s; kwargs...)
tt_combined_sketch(A::TTvector{Float64,N}, rks_or_rmax, s::Int; kwargs...) where {N} =
    tt_combined_sketch(Float64, A, rks_or_rmax, s; kwargs...)
tt_combined_sketch(H::TToperator{Float64,N}, A::TTvector{Float64,N}, rks_or_rmax, s::Int; kwargs...) where {N} =
    tt_combined_sketch(Float64, H, A, rks_or_rmax, s; kwargs...)

# ------------------------------------------------------------------ STTA (src/tt_randtools.jl:552-638)

# the reference's left (forward, oversampling 1.5, seed_left) and right (reverse,
# seed_right) blocks of stta_sketch (src/sketches.jl:1126-1140)
function _stta_blocks(dims, rks, block_rks, N, orthogonal, seed_left, seed_right)
    bvl = _block_ranks_fwd(dims, block_rks, N)
    pl = TensorTrains.compute_sketch_blocks_heuristic(collect(Int, rks), bvl, N; reverse=false)
    pl[2:N+1] .= ceil.(Int, 1.5 .* pl[2:N+1])
    bvr, pr = _reverse_structure(dims, rks, block_rks, N)
    return _reference_blocks(Float64, dims, bvl, pl, orthogonal, seed_left, false),
           _reference_blocks(Float64, dims, bvr, pr, orthogonal, seed_right, true)
end

function _stta(α::Vector{Float64}, y::Vector{TTvector{Float64,N}}, rks::Vector{Int}; seed_left=1234, seed_right=5678,
               orthogonal=true, block_rks::Int=N, reference_draws::Bool=true, timer=nothing) where {N}
    BL, BR = reference_draws ? _stta_blocks(y[1].ttv_dims, rks, block_rks, N, orthogonal, seed_left, seed_right) :
                               (Array{Float64,4}[], Array{Float64,4}[])
    pl, pr = [pointer(b) for b in BL], [pointer(b) for b in BR]
    ids = [upload(t) for t in y]
    r = collect(Int64, rks)
    out = Ref{Int64}(0)
    GC.@preserve α ids r BL BR pl pr begin
        item = Ref(RandItem(length(y), pointer(α), pointer(ids), -1, pointer(r)))
        nb = Ptr{Ptr{Float64}}(C_NULL)
        opts = _opts(orthogonal, Float64, block_rks, seed_left; blocks=reference_draws ? pointer(pl) : nb,
                     blocks_right=reference_draws ? pointer(pr) : nb)
        check(ccall((:ttr_stta, lib), Cint, (Ptr{Cvoid}, Int64, Ref{RandItem}, Ref{RandOpts}, UInt64, Ref{Int64}),
                    handle(), 1, item, opts, seed_right % UInt64, out))
    end
    foreach(free, ids)
    return download(out[], Val(N))
end

_stta_rks(rmax::Int, N) = (r = ones(Int, N + 1); r[2:N] .= rmax; r)
_stta_rks(rks::Vector{Int}, N) = rks

stta(y::TTvector{Float64,N}, rks=TensorTrains.default_rank_heuristic(y); kwargs...) where {N} =
    _stta([1.0], [y], _stta_rks(rks, N); kwargs...)
stta(α::Vector{Float64}, y::Vector{TTvector{Float64,N}}, rks=TensorTrains.default_rank_heuristic(y);
     kwargs...) where {N} = _stta(α, y, _stta_rks(rks, N); kwargs...)

# stta_sketch(x, rks; seed_left, seed_right, orthogonal, block_rks) (src/sketches.jl:1126-1156)
function stta_sketch(x::TTvector{Float64,N}, rks::Vector{Int}; seed_left=1234, seed_right=5678, orthogonal=true,
                     block_rks::Int=N, reference_draws::Bool=true, timer=nothing) where {N}
    sl, sr = zeros(Int64, N + 1), zeros(Int64, N + 1)
    BL, BR = reference_draws ? _stta_blocks(x.ttv_dims, rks, block_rks, N, orthogonal, seed_left, seed_right) :
                               (Array{Float64,4}[], Array{Float64,4}[])
    pl, pr = [pointer(b) for b in BL], [pointer(b) for b in BR]
    id = upload(x)
    r = collect(Int64, rks)
    nb = Ptr{Ptr{Float64}}(C_NULL)
    opts = _opts(orthogonal, Float64, block_rks, seed_left; blocks=reference_draws ? pointer(pl) : nb,
                 blocks_right=reference_draws ? pointer(pr) : nb)
    nullp = fill(Ptr{Float64}(C_NULL), N)
    # the sketch ranks first (no outputs), then the arrays
    check(ccall((:ttr_stta_sketch, lib), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}, Ref{RandOpts}, UInt64,
                Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Int64}, Ptr{Int64}),
                handle(), id, r, opts, seed_right % UInt64, nullp, nullp, sl, sr))
    Ω = [zeros(Float64, sl[k + 1], sr[k + 1]) for k in 1:N-1]
    Ψ = [zeros(Float64, x.ttv_dims[k], sl[k], sr[k + 1]) for k in 1:N]
    GC.@preserve Ω Ψ BL BR pl pr begin
        check(ccall((:ttr_stta_sketch, lib), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}, Ref{RandOpts}, UInt64,
                    Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Int64}, Ptr{Int64}),
                    handle(), id, r, opts, seed_right % UInt64, [pointer(o) for o in Ω], [pointer(q) for q in Ψ],
                    sl, sr))
    end
    free(id)
    return Ω, Ψ
end

# stta_sketch(x, L, R) with explicit sketch TTs (src/sketches.jl:1112-1124)
function stta_sketch(x::TTvector{Float64,N}, L::TTvector{Float64,N}, R::TTvector{Float64,N}) where {N}
    (x.ttv_dims == L.ttv_dims == R.ttv_dims) || throw(DimensionMismatch("TT dimensions are not compatible"))
    Ω = [zeros(Float64, L.ttv_rks[k + 1], R.ttv_rks[k + 1]) for k in 1:N-1]
    Ψ = [zeros(Float64, x.ttv_dims[k], L.ttv_rks[k], R.ttv_rks[k + 1]) for k in 1:N]
    a, l, r = upload(x), upload(L), upload(R)
    GC.@preserve Ω Ψ begin
        check(ccall((:ttr_stta_sketch_tt, lib), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}),
                    handle(), a, l, r, [pointer(o) for o in Ω], [pointer(q) for q in Ψ]))
    end
    free(a); free(l); free(r)
    return Ω, Ψ
end

# ------------------------------------------------------------------ sketched Krylov basis

# expand_basis! (src/tt_solvers.jl:354-455).  Its left-to-right sweep over H b_prev and
# -h_i b_i with the window's sketches (drawn from the same seed at ranks [2rmax, ..., 2rmax, 1])
# is the operator form of the randomized rounding, so it runs as one device call (rand_round
# sketches the terms itself from the reference's draws: the same sketches as the window's
# stored ones up to rounding); h, the normalisation and the window bookkeeping stay on the
# host.  The other solver-side callers, sketched_rr and sketched_rayleigh_ritz
# (src/tt_solvers.jl:466-689), need no method here: they reach the device through
# ttrand_rounding, tt_rounding, tt_sketch, tt_combined_sketch, dot_operator and this method.
function expand_basis!(H::TToperator{Float64,N}, B_window::Vector{TTvector{Float64,N}},
                       W_B_window::Vector{Vector{Matrix{Float64}}}, B_sketch_window::Vector{Vector{Float64}},
                       rmax::Int, s::Int; orthogonal::Bool=true, block_rks=8, seed::Int=1234,
                       k_trunc::Int=length(B_window)) where {N}
    b_prev = B_window[end]
    sketch_Hb_prev, _, _ = tt_combined_sketch(Float64, H, b_prev, 2rmax, s; orthogonal=orthogonal, reverse=true,
                                              seed=seed, block_rks=block_rks)
    h = hcat(B_sketch_window...) \ sketch_Hb_prev
    y = rand_round(vcat(1.0, -h), H, vcat([b_prev], B_window), vcat(fill(2rmax, N), 1);
                   orthogonal=orthogonal, seed=seed, block_rks=block_rks)
    b_new = tt_rounding(y; rmax=rmax)
    sketch_b_new, W_B_new, _ = tt_combined_sketch(Float64, b_new, 2rmax, s; orthogonal=orthogonal, reverse=true,
                                                  seed=seed, block_rks=block_rks)
    β = LinearAlgebra.norm(sketch_b_new)
    sketch_b_new ./= β
    for i = 2:findfirst(isequal(0), b_new.ttv_ot)
        W_B_new[i] ./= β
    end
    b_new = b_new / β
    push!(B_window, b_new);               length(B_window) > k_trunc && popfirst!(B_window)
    push!(B_sketch_window, sketch_b_new); length(B_sketch_window) > k_trunc && popfirst!(B_sketch_window)
    push!(W_B_window, W_B_new);           length(W_B_window) > k_trunc && popfirst!(W_B_window)
    return b_new, sketch_b_new, sketch_Hb_prev
end

end # module
