"""
    CoolPDLPHIP

Julia binding of `libcoolpdlp_hip.so` (the MI355X PDLP/PDHG engine) for CoolPDLP.jl.

It plugs into the extension points CoolPDLP already has, so `solve`, `MOI.optimize!`
and every option name stay as they are:

1. the backend type parameter `B <: Backend` of `Algorithm{A,T,Ti,M,B}`
   (`src/algorithms/common.jl:8-14`, `src/components/conversion.jl:14-15`):
   `HIPEngine <: KernelAbstractions.GPU`.  `solve(milp, sol, algo)` with
   `backend = HIPEngine()` runs the whole PDLP/PDHG loop (scaling, step size, restarts,
   termination) in the library through `cpdlp_solve` and returns the reference's
   `(PrimalDualSolution, ConvergenceStats)`.  The float type `T` selects the engine's
   precision (`Float64`, `Float32`; `HIPEngine(mixed = true)` with `Float64` for FP32 bulk
   with FP64 checks and polish), the matrix type `M` its sparse format
   (`GPUSparseMatrixCSR` or `HIPSparseMatrixCSR` -> CSR, `GPUSparseMatrixELL` -> ELL,
   anything else -> autotuned), the index width of the input's `colptr`/`rowval` the
   C ABI's `index_bits`.  Through MOI: `JuMP.set_attribute(model, "backend", HIPEngine())`.
2. the matrix plug-in (`src/utils/mat_csr.jl:8-92`): `HIPSparseMatrixCSR(A)` builds the
   matrix in HBM and implements `get_backend`, `Adapt.adapt_structure` and the
   5-argument `mul!` on AMDGPU device vectors, so the reference's own loop can run on
   AMDGPU's `ROCBackend()` with the library's SpMV.

This file is not loaded by anything in this repository (there is no Julia in the
image); it is the binding a maintainer would add, written against
`include/coolpdlp_hip.h` (tests/test_julia_binding.py checks its structs and ccall
signatures against the header).
"""
module CoolPDLPHIP

using CoolPDLP
using CoolPDLP: Algorithm, ConvergenceStats, KKTErrors, MILP, PrimalDualSolution,
    GPUSparseMatrixCSR, GPUSparseMatrixELL, OPTIMAL, TIME_LIMIT, ITERATION_LIMIT, STILL_RUNNING
using Adapt: Adapt
using AMDGPU: AMDGPU, ROCBackend, ROCVector
using KernelAbstractions: KernelAbstractions
using LinearAlgebra
using SparseArrays

const libcpdlp = get(ENV, "COOLPDLP_HIP_LIB", "libcoolpdlp_hip.so")

# ---- C structs (include/coolpdlp_hip.h) ----------------------------------------
struct CLP
    m::Int64
    n::Int64
    index_base::Int32
    index_bits::Int32
    colptr::Ptr{Cvoid}
    rowval::Ptr{Cvoid}
    nzval::Ptr{Float64}
    c::Ptr{Float64}
    lv::Ptr{Float64}
    uv::Ptr{Float64}
    lc::Ptr{Float64}
    uc::Ptr{Float64}
end

mutable struct CParams
    algorithm::Int32
    ruiz_iter::Int32
    chambolle_pock_alpha::Float64
    invnorm_scaling::Float64
    primal_weight_damping::Float64
    zero_tol::Float64
    sufficient_decay::Float64
    necessary_decay::Float64
    artificial_decay::Float64
    show_progress::Int32
    check_every::Int32
    record_error_history::Int32
    precision::Int32
    termination_reltol::Float64
    max_kkt_passes::Int64
    time_limit::Float64
    format::Int32
    device::Int32
    powm_tol::Float64
    powm_maxiter::Int64
    powm_seed::UInt64
    polish_reltol::Float64
    use_graph::Int32
    long_row::Int32
    batch_streams::Int32
    detect_infeasibility::Int32   # left at 0: the reference's TerminationStatus has no infeasible values
    infeasibility_tol::Float64
    CParams() = new()
end

struct CKKT
    primal::Float64
    primal_scale::Float64
    dual::Float64
    dual_scale::Float64
    gap::Float64
    gap_scale::Float64
end

struct CStats
    err::CKKT
    time_elapsed::Float64
    kkt_passes::Int64
    termination_status::Int32
    format_A::Int32
    format_At::Int32
    line_tiles::Int32
    eta::Float64
    omega::Float64
    powm_iters::Int64
    restarts::Int64
    n_history::Int64
    fp32_passes::Int64
    setup_seconds::Float64
    loop_seconds::Float64
    step_kernel_ms::Float64
    step_kernel_launches::Int64
end

function check(code::Cint)
    code == 0 && return nothing
    msg = unsafe_string(ccall((:cpdlp_last_error, libcpdlp), Cstring, ()))
    code == 1 && throw(DimensionMismatch(msg))
    code == 2 && throw(ArgumentError(msg))
    code == 3 && throw(OutOfMemoryError())
    error("libcoolpdlp_hip error $code: $msg")
end

# ---- 1. matrix plug-in ----------------------------------------------------------
# the C ABI's index width of a CSC matrix's colptr / rowval (int32 or int64)
function index_bits(A::SparseMatrixCSC)
    Ti = eltype(A.colptr)
    eltype(A.rowval) === Ti || throw(ArgumentError("colptr and rowval must have the same integer type"))
    Ti in (Int32, Int64) || throw(ArgumentError("index type $Ti: the library reads Int32 or Int64 indices"))
    return Int32(8 * sizeof(Ti))
end

"""
    HIPSparseMatrixCSR(A::SparseMatrixCSC{Float64, <:Union{Int32, Int64}}; format = 1, device = 0)

`GPUSparseMatrixCSR` equivalent stored in HBM by libcoolpdlp_hip
(`cpdlp_matrix_create`, replaces `GPUSparseMatrixCSR(A)`, `src/utils/mat_csr.jl:58-61`).
`set_matrix_type` (`src/problems/modify.jl:80-105`) builds one for `A` and one for `At`.
Values are Float64 (the library's matrix plug-in computes in FP64).
"""
mutable struct HIPSparseMatrixCSR <: AbstractSparseMatrix{Float64, Int32}
    handle::Ptr{Cvoid}
    m::Int
    n::Int
    nnz::Int
    device::Int
    function HIPSparseMatrixCSR(A::SparseMatrixCSC{Float64}; format::Integer = 1, device::Integer = 0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:cpdlp_matrix_create, libcpdlp), Cint,
            (Int64, Int64, Int32, Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Int32, Int32, Int32, Ptr{Ptr{Cvoid}}),
            size(A, 1), size(A, 2), 1, index_bits(A), A.colptr, A.rowval, A.nzval, format, 0, device, h))
        M = new(h[], size(A, 1), size(A, 2), nnz(A), device)
        finalizer(x -> ccall((:cpdlp_matrix_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), x.handle), M)
        return M
    end
end

HIPSparseMatrixCSR(A::SparseMatrixCSC{T}; kwargs...) where {T} =
    throw(ArgumentError("HIPSparseMatrixCSR stores Float64 values (got $T); for Float32 use backend = HIPEngine()"))

Base.size(A::HIPSparseMatrixCSR) = (A.m, A.n)
SparseArrays.nnz(A::HIPSparseMatrixCSR) = A.nnz

# src/utils/mat_csr.jl:44-46: the backend of the vectors it multiplies (AMDGPU's arrays)
KernelAbstractions.get_backend(::HIPSparseMatrixCSR) = ROCBackend()

# src/utils/mat_csr.jl:48-56: the matrix already lives in HBM, owned by the library
Adapt.adapt_structure(to, A::HIPSparseMatrixCSR) = A

# mul!(c, A, b, α, β), src/utils/mat_csr.jl:81-92 -> cpdlp_matrix_mul on the device
# pointers of AMDGPU vectors, on AMDGPU's current stream (ordered with its broadcasts)
function LinearAlgebra.mul!(c::ROCVector{Float64}, A::HIPSparseMatrixCSR,
        b::ROCVector{Float64}, α::Number, β::Number)
    length(b) == A.n && length(c) == A.m || throw(DimensionMismatch("mul!: sizes of A, b and c"))
    GC.@preserve b c begin
        check(ccall((:cpdlp_matrix_mul, libcpdlp), Cint,
            (Ptr{Cvoid}, Float64, Ptr{Float64}, Float64, Ptr{Float64}, Ptr{Cvoid}),
            A.handle, Float64(α), reinterpret(Ptr{Float64}, pointer(b)), Float64(β),
            reinterpret(Ptr{Float64}, pointer(c)), AMDGPU.stream().stream))
    end
    return c
end

# ---- 2. the whole loop ------------------------------------------------------------
"""
    HIPEngine(; device = 0, mixed = false)

Backend of `Algorithm{A,T,Ti,M,B}` (`B <: Backend`, `src/algorithms/common.jl:8-14`):
`PDLP(T, Ti, M; backend = HIPEngine())` makes `solve` run on libcoolpdlp_hip.  The
precision follows `T`; `mixed = true` (with `T = Float64`) runs FP32 matrix, iterates and
products with FP64 checks and an FP64 polish.  The whole solve runs in the library, so
nothing is allocated on this backend through KernelAbstractions.
"""
Base.@kwdef struct HIPEngine <: KernelAbstractions.GPU
    device::Int = 0
    mixed::Bool = false
end

KernelAbstractions.allocate(::HIPEngine, ::Type, dims...) = throw(ArgumentError(
    "HIPEngine runs the whole solve in libcoolpdlp_hip; for device arrays use AMDGPU's ROCBackend()"))

# T of Algorithm{A,T,...} -> cpdlp_params.precision (common.jl:55-57)
precision_code(::Type{Float64}, eng::HIPEngine) = Int32(eng.mixed ? 2 : 0)
precision_code(::Type{Float32}, eng::HIPEngine) =
    eng.mixed ? throw(ArgumentError("mixed precision runs with T = Float64")) : Int32(1)
precision_code(::Type{T}, ::HIPEngine) where {T} =
    throw(ArgumentError("float type $T: the library computes in Float64 or Float32"))

# M of Algorithm{A,T,Ti,M,...} -> cpdlp_params.format
format_code(::Type{<:GPUSparseMatrixELL}) = Int32(2)
format_code(::Type{<:GPUSparseMatrixCSR}) = Int32(1)
format_code(::Type{HIPSparseMatrixCSR}) = Int32(1)
format_code(::Type) = Int32(0)   # SparseMatrixCSC and others: autotuned

function cparams(algo::Algorithm{A, T, Ti, M, HIPEngine}) where {A, T, Ti, M}
    eng = algo.conversion.backend
    p = CParams()
    check(ccall((:cpdlp_default_params, libcpdlp), Cint, (Ref{CParams}, Int32), p, A === :PDLP ? 0 : 1))
    (; preconditioning, step_size, restart, generic, termination) = algo
    p.ruiz_iter = preconditioning.ruiz_iter
    p.chambolle_pock_alpha = preconditioning.chambolle_pock_alpha
    p.invnorm_scaling = step_size.invnorm_scaling
    p.primal_weight_damping = step_size.primal_weight_damping
    p.zero_tol = step_size.zero_tol
    p.sufficient_decay = restart.sufficient_decay
    p.necessary_decay = restart.necessary_decay
    p.artificial_decay = restart.artificial_decay
    p.check_every = generic.check_every
    p.record_error_history = generic.record_error_history
    p.show_progress = generic.show_progress
    p.termination_reltol = termination.termination_reltol
    p.max_kkt_passes = termination.max_kkt_passes
    p.time_limit = termination.time_limit
    p.precision = precision_code(T, eng)
    p.format = format_code(M)
    p.device = eng.device
    return p
end

const STATUS = (OPTIMAL, TIME_LIMIT, ITERATION_LIMIT, STILL_RUNNING)

"""
    HostLP(milp)

The C ABI's view of an unscaled CPU `MILP` (as `solve` receives it, `common.jl:186-190`):
CSC arrays with their own index width and Float64 copies of the values where needed.
Keep it alive (`GC.@preserve`) while the library reads it.
"""
struct HostLP
    A::SparseMatrixCSC
    nz::Vector{Float64}
    vecs::NTuple{5, Vector{Float64}}
    lp::CLP
end

function HostLP(milp::MILP)
    A_ = milp.A
    A_ isa SparseMatrixCSC || throw(ArgumentError("solve with HIPEngine takes a CPU MILP (SparseMatrixCSC)"))
    (all(isone, milp.D1.diag) && all(isone, milp.D2.diag)) ||
        throw(ArgumentError("solve with HIPEngine takes an unpreconditioned MILP (D1 = D2 = I)"))
    nz = Vector{Float64}(A_.nzval)
    vecs = map(v -> Vector{Float64}(v), (milp.c, milp.lv, milp.uv, milp.lc, milp.uc))
    lp = CLP(size(A_, 1), size(A_, 2), 1, index_bits(A_), pointer(A_.colptr), pointer(A_.rowval),
        pointer(nz), pointer(vecs[1]), pointer(vecs[2]), pointer(vecs[3]), pointer(vecs[4]), pointer(vecs[5]))
    return HostLP(A_, nz, vecs, lp)
end

kkt_errors(::Type{T}, e::CKKT) where {T} = KKTErrors{T}(; primal = T(e.primal), primal_scale = T(e.primal_scale),
    dual = T(e.dual), dual_scale = T(e.dual_scale), gap = T(e.gap), gap_scale = T(e.gap_scale))

status_of(code::Integer) = code in 0:3 ? STATUS[code + 1] :
    error("libcoolpdlp_hip: termination status $code has no TerminationStatus value")

function convergence_stats(::Type{T}, s::CStats, history) where {T}
    return ConvergenceStats(T; err = kkt_errors(T, s.err), time_elapsed = s.time_elapsed,
        kkt_passes = s.kkt_passes, termination_status = status_of(s.termination_status),
        error_history = history)
end

# solve(milp, sol, algo), src/algorithms/common.jl:186-201, for backend = HIPEngine()
function CoolPDLP.solve(milp::MILP, sol::PrimalDualSolution, algo::Algorithm{A, T, Ti, M, HIPEngine}) where {A, T, Ti, M}
    H = HostLP(milp)
    length(sol.x) == H.lp.n && length(sol.y) == H.lp.m ||
        throw(DimensionMismatch("starting point does not match the LP's size"))
    p = cparams(algo)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve H begin
        check(ccall((:cpdlp_create, libcpdlp), Cint, (Ref{CLP}, Ref{CParams}, Ptr{Ptr{Cvoid}}), H.lp, p, h))
    end
    try
        x0, y0 = Vector{Float64}(Array(sol.x)), Vector{Float64}(Array(sol.y))
        check(ccall((:cpdlp_solve, libcpdlp), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), h[], x0, y0))
        x, y = similar(x0), similar(y0)
        check(ccall((:cpdlp_get_solution, libcpdlp), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), h[], x, y))
        s = Ref{CStats}()
        check(ccall((:cpdlp_get_stats, libcpdlp), Cint, (Ptr{Cvoid}, Ref{CStats}), h[], s))
        count = Ref{Int64}(0)
        check(ccall((:cpdlp_get_history, libcpdlp), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{CKKT}, Int64, Ref{Int64}),
            h[], C_NULL, C_NULL, 0, count))
        passes = Vector{Int64}(undef, count[])
        errs = Vector{CKKT}(undef, count[])
        check(ccall((:cpdlp_get_history, libcpdlp), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{CKKT}, Int64, Ref{Int64}),
            h[], passes, errs, count[], count))
        history = Tuple{Int, KKTErrors{T}}[(Int(passes[i]), kkt_errors(T, errs[i])) for i in eachindex(passes)]
        return PrimalDualSolution(Vector{T}(x), Vector{T}(y)), convergence_stats(T, s[], history)
    finally
        ccall((:cpdlp_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), h[])
    end
end

# ---- 3. the loop in pieces (initialize / solve! yielding per chunk) --------------
"""
    HIPState

Handle of a solve run chunk by chunk: `initialize` (common.jl:176) -> `cpdlp_initialize`,
`solve!` (pdlp.jl:69-94) -> `cpdlp_run_checks` one chunk of `check_every` steps at a time,
yielding in between as the reference does (pdlp.jl:77).  Any precision: with
`HIPEngine(mixed = true)` the chunks run the FP32 phase, then the FP64 one, as `solve`.
"""
mutable struct HIPState{T}
    handle::Ptr{Cvoid}
    m::Int
    n::Int
end

function hip_initialize(milp::MILP, sol::PrimalDualSolution, algo::Algorithm{A, T, Ti, M, HIPEngine}) where {A, T, Ti, M}
    H = HostLP(milp)
    p = cparams(algo)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve H begin
        check(ccall((:cpdlp_create, libcpdlp), Cint, (Ref{CLP}, Ref{CParams}, Ptr{Ptr{Cvoid}}), H.lp, p, h))
    end
    state = HIPState{T}(h[], H.lp.m, H.lp.n)
    finalizer(s -> ccall((:cpdlp_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), s.handle), state)
    x0, y0 = Vector{Float64}(Array(sol.x)), Vector{Float64}(Array(sol.y))
    length(x0) == state.n && length(y0) == state.m ||
        throw(DimensionMismatch("starting point does not match the LP's size"))
    check(ccall((:cpdlp_initialize, libcpdlp), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
        state.handle, x0, y0))
    return state
end

function hip_solve!(state::HIPState)
    status = Ref{Int32}(3)   # STILL_RUNNING
    while status[] == 3
        check(ccall((:cpdlp_run_checks, libcpdlp), Cint, (Ptr{Cvoid}, Int64, Ref{Int32}),
            state.handle, 1, status))
        yield()
    end
    return status_of(status[])
end

# the loop's state between two checks (cpdlp_state): PDLPState, pdlp.jl:19-43, less its
# scratch and per-check statistics
struct CState
    eta::Float64
    omega::Float64
    eta_sum::Float64
    kkt_passes::Int64
    inner::Int64
    total::Int64
    restarts::Int64
    fp32_passes::Int64
end

"""
    HIPCheckpoint

A chunked solve's loop state (`hip_checkpoint`): the scalars, the five primal and dual
vectors (sol, sol_last, sol_avg, sol_avg_last, sol_restart; scaled, original order, one
column each) and the scaling they belong to.  `hip_resume` continues from it.
"""
struct HIPCheckpoint
    st::CState
    x::Matrix{Float64}
    y::Matrix{Float64}
    D2::Vector{Float64}
    D1::Vector{Float64}
end

function hip_checkpoint(state::HIPState)
    x, y = zeros(state.n, 5), zeros(state.m, 5)
    D2, D1 = zeros(state.n), zeros(state.m)
    st = Ref{CState}()
    check(ccall((:cpdlp_get_state, libcpdlp), Cint,
        (Ptr{Cvoid}, Ref{CState}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        state.handle, st, x, y, D2, D1))
    return HIPCheckpoint(st[], x, y, D2, D1)
end

"""
    hip_resume(milp, checkpoint, algo)

A new chunked solve of `milp` that continues the run `checkpoint` was taken from
(`cpdlp_resume`): with the same scaling the iterates are those of the run that never
stopped.  Then `hip_solve!` and `hip_result` as after `hip_initialize`.
"""
function hip_resume(milp::MILP, cpt::HIPCheckpoint, algo::Algorithm{A, T, Ti, M, HIPEngine}) where {A, T, Ti, M}
    H = HostLP(milp)
    p = cparams(algo)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve H begin
        check(ccall((:cpdlp_create, libcpdlp), Cint, (Ref{CLP}, Ref{CParams}, Ptr{Ptr{Cvoid}}), H.lp, p, h))
    end
    state = HIPState{T}(h[], H.lp.m, H.lp.n)
    finalizer(s -> ccall((:cpdlp_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), s.handle), state)
    size(cpt.x) == (state.n, 5) && size(cpt.y) == (state.m, 5) ||
        throw(DimensionMismatch("checkpoint does not match the LP's size"))
    check(ccall((:cpdlp_resume, libcpdlp), Cint,
        (Ptr{Cvoid}, Ref{CState}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        state.handle, cpt.st, cpt.x, cpt.y, cpt.D2, cpt.D1))
    return state
end

"The solution and stats of a chunked solve, as `solve` returns them."
function hip_result(state::HIPState{T}) where {T}
    x, y = zeros(state.n), zeros(state.m)
    check(ccall((:cpdlp_get_solution, libcpdlp), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
        state.handle, x, y))
    s = Ref{CStats}()
    check(ccall((:cpdlp_get_stats, libcpdlp), Cint, (Ptr{Cvoid}, Ref{CStats}), state.handle, s))
    return PrimalDualSolution(Vector{T}(x), Vector{T}(y)), convergence_stats(T, s[], Tuple{Int, KKTErrors{T}}[])
end

# ---- 4. many LPs at once (cpdlp_batch_*; no counterpart in the reference) ----------
"""
    solve(milps::Vector{<:MILP}, algo; hbm_bytes, checks_per_round = 20, sols = nothing)

Solve independent LPs together on one GPU (one workgroup per LP, `m + n <= 65536` each),
returning `[(PrimalDualSolution, ConvergenceStats)]` in input order.  `sols` optionally
holds one start point per LP (`nothing` entries start from zero), as `sol` in
`solve(milp, sol, algo)`.
"""
function CoolPDLP.solve(milps::Vector{<:MILP}, algo::Algorithm{A, T, Ti, M, HIPEngine};
        hbm_bytes::Integer = 0, checks_per_round::Integer = 20, sols = nothing) where {A, T, Ti, M}
    starts = sols === nothing ? fill(nothing, length(milps)) : sols
    length(starts) == length(milps) || throw(DimensionMismatch("one start point (or nothing) per LP"))
    p = cparams(algo)
    lps = map(HostLP, milps)
    if hbm_bytes == 0
        hbm_bytes = sum(lps) do H
            ccall((:cpdlp_batch_footprint, libcpdlp), Int64, (Int64, Int64, Int64), H.lp.m, H.lp.n, nnz(H.A))
        end + (64 << 20)
    end
    b = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cpdlp_batch_create, libcpdlp), Cint, (Ref{CParams}, Int64, Ptr{Ptr{Cvoid}}), p, hbm_bytes, b))
    try
        ids = map(zip(lps, starts)) do (H, sol)
            id = Ref{Int64}(0)
            x0 = sol === nothing ? Float64[] : Vector{Float64}(Array(sol.x))
            y0 = sol === nothing ? Float64[] : Vector{Float64}(Array(sol.y))
            sol === nothing || (length(x0) == H.lp.n && length(y0) == H.lp.m) ||
                throw(DimensionMismatch("start point does not match the LP's size"))
            GC.@preserve H x0 y0 begin
                check(ccall((:cpdlp_batch_add_start, libcpdlp), Cint,
                    (Ptr{Cvoid}, Ref{CLP}, Ptr{Float64}, Ptr{Float64}, Ref{Int64}), b[], H.lp,
                    sol === nothing ? C_NULL : pointer(x0), sol === nothing ? C_NULL : pointer(y0), id))
            end
            id[]
        end
        running = Ref{Int64}(length(ids))
        while running[] > 0
            check(ccall((:cpdlp_batch_run, libcpdlp), Cint, (Ptr{Cvoid}, Int32, Ref{Int64}),
                b[], checks_per_round, running))
            yield()
        end
        return map(zip(ids, lps)) do (id, H)
            x, y = zeros(H.lp.n), zeros(H.lp.m)
            s = Ref{CStats}()
            check(ccall((:cpdlp_batch_get, libcpdlp), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{CStats}),
                b[], id, x, y, s))
            check(ccall((:cpdlp_batch_release, libcpdlp), Cint, (Ptr{Cvoid}, Int64), b[], id))
            (PrimalDualSolution(Vector{T}(x), Vector{T}(y)), convergence_stats(T, s[], Tuple{Int, KKTErrors{T}}[]))
        end
    finally
        ccall((:cpdlp_batch_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), b[])
    end
end

# ---- 5. served requests (cpdlp_server_*; no counterpart in the reference) --------------
mutable struct CServerConfig
    batch_bytes::Int64
    plain_bytes::Int64
    batch_limit::Int32
    checks_per_round::Int32
    n_lanes::Int32
    lane_workers::NTuple{4, Int32}
    lane_max_nnz::NTuple{4, Float64}
    batch2_bytes::Int64
    batch2_limit::Int32
    batch2_checks::Int32
    CServerConfig() = new()
end

"""
    HIPServer(algo; batch_limit = 8192, batch2_limit = 0, checks_per_round = 20)

The library's continuous-batching server: `submit!` queues an LP (optionally with a start
point; the library copies it), `pump!` runs one scheduling round (or `start!` lets the
library pump on its own threads), `poll!` lists finished requests and `take!` hands one
over as `(PrimalDualSolution, ConvergenceStats, seconds)`.  LPs with m + n <= batch_limit
go to one batch engine, those with m + n <= batch2_limit to a second one with rounds of
its own, larger ones to plain engines.
"""
mutable struct HIPServer{T}
    ptr::Ptr{Cvoid}
    shapes::Dict{Int64, Tuple{Int, Int}}
end

function HIPServer(algo::Algorithm{A, T, Ti, M, HIPEngine}; batch_limit::Integer = 8192,
        batch2_limit::Integer = 0, checks_per_round::Integer = 20) where {A, T, Ti, M}
    p = cparams(algo)
    cfg = CServerConfig()
    check(ccall((:cpdlp_server_default_config, libcpdlp), Cint, (Ref{CServerConfig},), cfg))
    cfg.batch_limit = batch_limit
    cfg.batch2_limit = batch2_limit
    cfg.checks_per_round = checks_per_round
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cpdlp_server_create, libcpdlp), Cint, (Ref{CParams}, Ref{CServerConfig}, Ptr{Ptr{Cvoid}}),
        p, cfg, h))
    s = HIPServer{T}(h[], Dict{Int64, Tuple{Int, Int}}())
    finalizer(s) do s
        s.ptr == C_NULL || ccall((:cpdlp_server_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), s.ptr)
        s.ptr = C_NULL
    end
    return s
end

function submit!(s::HIPServer, milp::MILP, sol::Union{Nothing, PrimalDualSolution} = nothing)
    H = HostLP(milp)
    x0 = sol === nothing ? Float64[] : Vector{Float64}(Array(sol.x))
    y0 = sol === nothing ? Float64[] : Vector{Float64}(Array(sol.y))
    sol === nothing || (length(x0) == H.lp.n && length(y0) == H.lp.m) ||
        throw(DimensionMismatch("start point does not match the LP's size"))
    id = Ref{Int64}(0)
    GC.@preserve H x0 y0 begin   # copied by the library during the call
        check(ccall((:cpdlp_server_submit, libcpdlp), Cint,
            (Ptr{Cvoid}, Ref{CLP}, Ptr{Float64}, Ptr{Float64}, Ref{Int64}), s.ptr, H.lp,
            sol === nothing ? C_NULL : pointer(x0), sol === nothing ? C_NULL : pointer(y0), id))
    end
    s.shapes[id[]] = (Int(H.lp.m), Int(H.lp.n))
    return id[]
end

"The library pumps on its own thread from now on (cpdlp_server_start)."
start!(s::HIPServer) = (check(ccall((:cpdlp_server_start, libcpdlp), Cint, (Ptr{Cvoid},), s.ptr)); s)

function pump!(s::HIPServer)
    pending = Ref{Int64}(0)
    check(ccall((:cpdlp_server_pump, libcpdlp), Cint, (Ptr{Cvoid}, Ref{Int64}), s.ptr, pending))
    return pending[]
end

function poll!(s::HIPServer)
    ids = zeros(Int64, max(length(s.shapes), 1))
    cnt = Ref{Int64}(0)
    check(ccall((:cpdlp_server_poll, libcpdlp), Cint, (Ptr{Cvoid}, Ptr{Int64}, Int64, Ref{Int64}),
        s.ptr, ids, length(ids), cnt))
    return ids[1:cnt[]]
end

function Base.take!(s::HIPServer{T}, id::Integer) where {T}
    haskey(s.shapes, id) || throw(ArgumentError("unknown request id $id"))
    m, n = s.shapes[id]
    x, y = zeros(n), zeros(m)
    st = Ref{CStats}()
    secs = Ref{Float64}(0.0)
    route = Ref{Int32}(0)
    rc = ccall((:cpdlp_server_get, libcpdlp), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{CStats}, Ref{Float64}, Ref{Int32}),
        s.ptr, id, x, y, st, secs, route)
    # a request that has not finished stays queued under its id (CPDLP_ERR_STATE)
    still_running = rc == 5 && unsafe_string(ccall((:cpdlp_last_error, libcpdlp), Cstring, ())) == "request still running"
    still_running || delete!(s.shapes, id)
    check(rc)
    stats = convergence_stats(T, st[], Tuple{Int, KKTErrors{T}}[])
    return PrimalDualSolution(Vector{T}(x), Vector{T}(y)), stats, secs[]
end

end # module
