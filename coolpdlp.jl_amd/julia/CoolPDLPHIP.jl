"""
    CoolPDLPHIP

Julia binding of `libcoolpdlp_hip.so` (the MI355X PDLP/PDHG engine) for CoolPDLP.jl.

It plugs into the two extension points CoolPDLP already has, so `solve`,
`MOI.optimize!` and every option name stay as they are:

1. the matrix plug-in (`src/utils/mat_csr.jl:8-92`): `HIPSparseMatrixCSR(A)` builds the
   matrix in HBM and implements the 5-argument `mul!` on device vectors;
2. algorithm dispatch (`src/algorithms/common.jl:176,216`): `solve(milp, sol, algo)` with
   `backend = HIPEngine()` runs the whole PDLP/PDHG loop (scaling, step size, restarts,
   termination) on the GPU through `cpdlp_solve` and returns the reference's
   `(PrimalDualSolution, ConvergenceStats)`.

This file is not loaded by anything in this repository (there is no Julia in the
image); it is the binding a maintainer would add, written against
`include/coolpdlp_hip.h`.
"""
module CoolPDLPHIP

using CoolPDLP
using CoolPDLP: Algorithm, ConvergenceStats, KKTErrors, MILP, PrimalDualSolution,
    OPTIMAL, TIME_LIMIT, ITERATION_LIMIT, STILL_RUNNING
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
    pad::Int32
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
"""
    HIPSparseMatrixCSR(A::SparseMatrixCSC{Float64})

`GPUSparseMatrixCSR` equivalent stored in HBM by libcoolpdlp_hip
(`cpdlp_matrix_create`, replaces `GPUSparseMatrixCSR(A)`, `src/utils/mat_csr.jl:58-61`).
"""
mutable struct HIPSparseMatrixCSR <: AbstractSparseMatrix{Float64, Int32}
    handle::Ptr{Cvoid}
    m::Int
    n::Int
    nnz::Int
    function HIPSparseMatrixCSR(A::SparseMatrixCSC{Float64}; format::Integer = 1, device::Integer = 0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:cpdlp_matrix_create, libcpdlp), Cint,
            (Int64, Int64, Int32, Int32, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Int32, Int32, Int32, Ptr{Ptr{Cvoid}}),
            size(A, 1), size(A, 2), 1, 64, A.colptr, A.rowval, A.nzval, format, 0, device, h))
        M = new(h[], size(A, 1), size(A, 2), nnz(A))
        finalizer(x -> ccall((:cpdlp_matrix_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), x.handle), M)
        return M
    end
end

Base.size(A::HIPSparseMatrixCSR) = (A.m, A.n)
SparseArrays.nnz(A::HIPSparseMatrixCSR) = A.nnz

# mul!(c, A, b, α, β), src/utils/mat_csr.jl:81-92 -> cpdlp_matrix_mul on device pointers
# (here: any array type whose pointer is a HIP device pointer, e.g. an AMDGPU.ROCArray)
function LinearAlgebra.mul!(c::AbstractVector{Float64}, A::HIPSparseMatrixCSR,
        b::AbstractVector{Float64}, α::Number, β::Number)
    check(ccall((:cpdlp_matrix_mul, libcpdlp), Cint,
        (Ptr{Cvoid}, Float64, Ptr{Float64}, Float64, Ptr{Float64}, Ptr{Cvoid}),
        A.handle, Float64(α), pointer(b), Float64(β), pointer(c), C_NULL))
    return c
end

# ---- 2. the whole loop ------------------------------------------------------------
"""
    HIPEngine(; device = 0, precision = :fp64, format = :auto)

Backend selector: `PDLP(...; backend = HIPEngine())` makes `solve` run on libcoolpdlp_hip.
"""
Base.@kwdef struct HIPEngine
    device::Int = 0
    precision::Symbol = :fp64
    format::Symbol = :auto
end

function cparams(algo::Algorithm{A}, eng::HIPEngine) where {A}
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
    p.precision = Dict(:fp64 => 0, :fp32 => 1, :mixed => 2)[eng.precision]
    p.format = Dict(:auto => 0, :csr => 1, :ell => 2, :xslice => 3, :xslice_all => 4)[eng.format]
    p.device = eng.device
    return p
end

const STATUS = (OPTIMAL, TIME_LIMIT, ITERATION_LIMIT, STILL_RUNNING)

# solve(milp, sol, algo), src/algorithms/common.jl:186-201, for backend = HIPEngine()
function CoolPDLP.solve(milp::MILP, sol::PrimalDualSolution, algo::Algorithm{A, T, Ti, M, HIPEngine}) where {A, T, Ti, M}
    A_ = milp.A::SparseMatrixCSC
    c, lv, uv, lc, uc = (Vector{Float64}(v) for v in (milp.c, milp.lv, milp.uv, milp.lc, milp.uc))
    lp = CLP(size(A_, 1), size(A_, 2), 1, 64, pointer(A_.colptr), pointer(A_.rowval),
        pointer(A_.nzval), pointer(c), pointer(lv), pointer(uv), pointer(lc), pointer(uc))
    p = cparams(algo, algo.conversion.backend)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve A_ c lv uv lc uc begin
        check(ccall((:cpdlp_create, libcpdlp), Cint, (Ref{CLP}, Ref{CParams}, Ptr{Ptr{Cvoid}}), lp, p, h))
    end
    try
        x0, y0 = Vector{Float64}(sol.x), Vector{Float64}(sol.y)
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
        tokkt(e) = KKTErrors(; primal = e.primal, primal_scale = e.primal_scale, dual = e.dual,
            dual_scale = e.dual_scale, gap = e.gap, gap_scale = e.gap_scale)
        stats = ConvergenceStats(Float64; err = tokkt(s[].err), time_elapsed = s[].time_elapsed,
            kkt_passes = s[].kkt_passes, termination_status = STATUS[s[].termination_status + 1],
            error_history = [(passes[i], tokkt(errs[i])) for i in eachindex(passes)])
        return PrimalDualSolution(x, y), stats
    finally
        ccall((:cpdlp_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), h[])
    end
end

# ---- 3. the loop in pieces (initialize / solve! yielding per chunk) --------------
"""
    HIPState

Handle of a solve run chunk by chunk: `initialize` (common.jl:176) -> `cpdlp_initialize`,
`solve!` (pdlp.jl:69-94) -> `cpdlp_run_checks` one chunk of `check_every` steps at a time,
yielding in between as the reference does (pdlp.jl:77).  FP64 / FP32 only.
"""
mutable struct HIPState
    handle::Ptr{Cvoid}
    m::Int
    n::Int
end

function hip_initialize(milp::MILP, sol::PrimalDualSolution, algo::Algorithm{A, T, Ti, M, HIPEngine}) where {A, T, Ti, M}
    A_ = milp.A::SparseMatrixCSC
    c, lv, uv, lc, uc = (Vector{Float64}(v) for v in (milp.c, milp.lv, milp.uv, milp.lc, milp.uc))
    lp = CLP(size(A_, 1), size(A_, 2), 1, 64, pointer(A_.colptr), pointer(A_.rowval),
        pointer(A_.nzval), pointer(c), pointer(lv), pointer(uv), pointer(lc), pointer(uc))
    p = cparams(algo, algo.conversion.backend)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve A_ c lv uv lc uc begin
        check(ccall((:cpdlp_create, libcpdlp), Cint, (Ref{CLP}, Ref{CParams}, Ptr{Ptr{Cvoid}}), lp, p, h))
    end
    state = HIPState(h[], size(A_, 1), size(A_, 2))
    finalizer(s -> ccall((:cpdlp_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), s.handle), state)
    x0, y0 = Vector{Float64}(sol.x), Vector{Float64}(sol.y)
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
    return STATUS[status[] + 1]
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
    p = cparams(algo, algo.conversion.backend)
    lps = map(milps) do milp
        A_ = milp.A::SparseMatrixCSC
        (A_, (Vector{Float64}(v) for v in (milp.c, milp.lv, milp.uv, milp.lc, milp.uc))...)
    end
    if hbm_bytes == 0
        hbm_bytes = sum(lps) do (A_, _...)
            ccall((:cpdlp_batch_footprint, libcpdlp), Int64, (Int64, Int64, Int64),
                size(A_, 1), size(A_, 2), nnz(A_))
        end + (64 << 20)
    end
    b = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cpdlp_batch_create, libcpdlp), Cint, (Ref{CParams}, Int64, Ptr{Ptr{Cvoid}}), p, hbm_bytes, b))
    try
        ids = map(zip(lps, starts)) do ((A_, c, lv, uv, lc, uc), sol)
            lp = CLP(size(A_, 1), size(A_, 2), 1, 64, pointer(A_.colptr), pointer(A_.rowval),
                pointer(A_.nzval), pointer(c), pointer(lv), pointer(uv), pointer(lc), pointer(uc))
            id = Ref{Int64}(0)
            x0 = sol === nothing ? Float64[] : Vector{Float64}(Array(sol.x))
            y0 = sol === nothing ? Float64[] : Vector{Float64}(Array(sol.y))
            sol === nothing || (length(x0) == size(A_, 2) && length(y0) == size(A_, 1)) ||
                throw(DimensionMismatch("start point does not match the LP's size"))
            GC.@preserve A_ c lv uv lc uc x0 y0 begin
                check(ccall((:cpdlp_batch_add_start, libcpdlp), Cint,
                    (Ptr{Cvoid}, Ref{CLP}, Ptr{Float64}, Ptr{Float64}, Ref{Int64}), b[], lp,
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
        return map(zip(ids, lps)) do (id, (A_, _...))
            x, y = zeros(size(A_, 2)), zeros(size(A_, 1))
            s = Ref{CStats}()
            check(ccall((:cpdlp_batch_get, libcpdlp), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{CStats}),
                b[], id, x, y, s))
            check(ccall((:cpdlp_batch_release, libcpdlp), Cint, (Ptr{Cvoid}, Int64), b[], id))
            e = s[].err
            err = KKTErrors(; primal = e.primal, primal_scale = e.primal_scale, dual = e.dual,
                dual_scale = e.dual_scale, gap = e.gap, gap_scale = e.gap_scale)
            stats = ConvergenceStats(Float64; err, time_elapsed = s[].time_elapsed,
                kkt_passes = s[].kkt_passes, termination_status = STATUS[s[].termination_status + 1])
            (PrimalDualSolution(x, y), stats)
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
    CServerConfig() = new()
end

"""
    HIPServer(algo; batch_limit = 8192, checks_per_round = 20)

The library's continuous-batching server: `submit!` queues an LP (optionally with a start
point; the library copies it), `pump!` runs one scheduling round (or `start!` lets the
library pump on its own thread), `poll!` lists finished requests and `take!` hands one
over as `(PrimalDualSolution, ConvergenceStats, seconds)`.
"""
mutable struct HIPServer
    ptr::Ptr{Cvoid}
    shapes::Dict{Int64, Tuple{Int, Int}}
end

function HIPServer(algo::Algorithm{A, T, Ti, M, HIPEngine}; batch_limit::Integer = 8192,
        checks_per_round::Integer = 20) where {A, T, Ti, M}
    p = cparams(algo, algo.conversion.backend)
    cfg = CServerConfig()
    check(ccall((:cpdlp_server_default_config, libcpdlp), Cint, (Ref{CServerConfig},), cfg))
    cfg.batch_limit = batch_limit
    cfg.checks_per_round = checks_per_round
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:cpdlp_server_create, libcpdlp), Cint, (Ref{CParams}, Ref{CServerConfig}, Ptr{Ptr{Cvoid}}),
        p, cfg, h))
    s = HIPServer(h[], Dict{Int64, Tuple{Int, Int}}())
    finalizer(s) do s
        s.ptr == C_NULL || ccall((:cpdlp_server_destroy, libcpdlp), Cvoid, (Ptr{Cvoid},), s.ptr)
        s.ptr = C_NULL
    end
    return s
end

function submit!(s::HIPServer, milp::MILP, sol::Union{Nothing, PrimalDualSolution} = nothing)
    A_ = milp.A::SparseMatrixCSC
    c, lv, uv, lc, uc = (Vector{Float64}(v) for v in (milp.c, milp.lv, milp.uv, milp.lc, milp.uc))
    lp = CLP(size(A_, 1), size(A_, 2), 1, 64, pointer(A_.colptr), pointer(A_.rowval),
        pointer(A_.nzval), pointer(c), pointer(lv), pointer(uv), pointer(lc), pointer(uc))
    x0 = sol === nothing ? Float64[] : Vector{Float64}(Array(sol.x))
    y0 = sol === nothing ? Float64[] : Vector{Float64}(Array(sol.y))
    sol === nothing || (length(x0) == size(A_, 2) && length(y0) == size(A_, 1)) ||
        throw(DimensionMismatch("start point does not match the LP's size"))
    id = Ref{Int64}(0)
    GC.@preserve A_ c lv uv lc uc x0 y0 begin   # copied by the library during the call
        check(ccall((:cpdlp_server_submit, libcpdlp), Cint,
            (Ptr{Cvoid}, Ref{CLP}, Ptr{Float64}, Ptr{Float64}, Ref{Int64}), s.ptr, lp,
            sol === nothing ? C_NULL : pointer(x0), sol === nothing ? C_NULL : pointer(y0), id))
    end
    s.shapes[id[]] = (size(A_, 1), size(A_, 2))
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

function Base.take!(s::HIPServer, id::Integer)
    m, n = pop!(s.shapes, id)
    x, y = zeros(n), zeros(m)
    st = Ref{CStats}()
    secs = Ref{Float64}(0.0)
    route = Ref{Int32}(0)
    check(ccall((:cpdlp_server_get, libcpdlp), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{CStats}, Ref{Float64}, Ref{Int32}),
        s.ptr, id, x, y, st, secs, route))
    e = st[].err
    err = KKTErrors(; primal = e.primal, primal_scale = e.primal_scale, dual = e.dual,
        dual_scale = e.dual_scale, gap = e.gap, gap_scale = e.gap_scale)
    stats = ConvergenceStats(Float64; err, time_elapsed = st[].time_elapsed,
        kkt_passes = st[].kkt_passes, termination_status = STATUS[st[].termination_status + 1])
    return PrimalDualSolution(x, y), stats, secs[]
end

end # module
