# HPRSOCPHip.jl — the reference side of the MI355X backend: ccall bindings of
# include/hprsocp.h for the HPRSOCP Julia package.
#
# Included from the HPRSOCP module (include("HPRSOCPHip.jl")), it adds
#   * HPRSOCPHip.optimize_hip(model, params): optimize() (src/algorithm.jl:2538)
#     on the HIP backend, one call of hprsocp_optimize;
#   * HPRSOCPHip.optimize_batch_hip(models, params): optimize() of every model,
#     the instances iterated together (hprsocp_batch_*);
#   * HPRSOCPHip.HipWorkspace and the seam functions hip_update!,
#     hip_residuals!, hip_mnorm, hip_update_sigma!, hip_do_restart!,
#     hip_iterate!: the per-step seams of solve() (src/algorithm.jl:2765-2905),
#     for a host that keeps its own solve loop.  check_restart stays the
#     reference's own Julia function; its HPRSOCP_restart object is passed to
#     the backend by reference (hprsocp_restart_info has its C layout).
# The one dispatch line the reference adds in optimize:
#     params.use_gpu && HIP_BACKEND[] && return HPRSOCPHip.optimize_hip(model, params)
# The MOI Optimizer reaches the same path through optimize.
module HPRSOCPHip

using SparseArrays
import ..HPRSOCP: QP_info_cpu, HPRSOCP_parameters, HPRSOCP_results, HPRSOCP_restart, AbstractQOperatorCPU

const lib = get(ENV, "HPRSOCP_HIP_LIB", "libhprsocp.so")

# ---------------------------------------------------------------- C structs
# isbits structs that repeat the C layout field by field (Julia lays out isbits
# structs with C alignment rules).

struct CProblem
    m::Cint; n::Cint; A_nnz::Int64
    A_rowptr::Ptr{Cint}; A_colidx::Ptr{Cint}; A_val::Ptr{Cdouble}
    Q_kind::Cint; Q_nnz::Int64
    Q_rowptr::Ptr{Cint}; Q_colidx::Ptr{Cint}; Q_val::Ptr{Cdouble}
    F_rows::Cint; F_nnz::Int64
    F_rowptr::Ptr{Cint}; F_colidx::Ptr{Cint}; F_val::Ptr{Cdouble}; Q_diag::Ptr{Cdouble}
    c::Ptr{Cdouble}; AL::Ptr{Cdouble}; AU::Ptr{Cdouble}; l::Ptr{Cdouble}; u::Ptr{Cdouble}
    soc_rhs::Ptr{Cdouble}
    number_linear_con::Cint; number_lu_x::Cint
    n_soc_con::Cint; soc_con_idx::Ptr{Cint}
    n_soc_var::Cint; soc_var_idx::Ptr{Cint}
    obj_constant::Cdouble
end

mutable struct CParams          # field order of hprsocp_params
    stoptol::Cdouble; sigma::Cdouble; max_iter::Int64; time_limit::Cdouble
    check_iter::Cint; warm_up::Cint; print_frequency::Cint; device_number::Cint
    use_Ruiz_scaling::Cint; ruiz_iterations::Cint; use_bc_scaling::Cint
    bc_scaling_norm_type::Cint; use_l2_scaling::Cint; use_Pock_Chambolle_scaling::Cint
    soc_block_scaling_strategy::Cint
    initial_x::Ptr{Cdouble}; initial_y::Ptr{Cdouble}
    auto_save::Cint; save_filename::Cstring; verbose::Cint; use_gpu::Cint
    precision::Cint; use_graph::Cint; refine_tol::Cdouble; fp32_stall_checks::Cint
    reserved::NTuple{7,Cint}
    CParams() = new()            # filled by hprsocp_default_params
end

struct CResult                  # isbits: Ref{CResult}() for one result, Vector{CResult} for a batch
    iter::Int64; iter_4::Int64; iter_6::Int64
    time::Cdouble; time_4::Cdouble; time_6::Cdouble; power_time::Cdouble; setup_time::Cdouble
    primal_obj::Cdouble; dual_obj::Cdouble; residuals::Cdouble; gap::Cdouble
    err_Rp::Cdouble; err_Rp_linear::Cdouble; err_Rp_soc::Cdouble; err_Rd::Cdouble
    sigma::Cdouble; lambda_max_A::Cdouble; lambda_max_Q::Cdouble
    status::Cint; restarts::Cint; iter_fp32::Int64
end

mutable struct CResiduals       # hprsocp_residuals, mirrors HPRSOCP_residuals
    err_Rp::Cdouble; err_Rp_linear::Cdouble; err_Rp_soc::Cdouble; err_Rd::Cdouble
    kkt::Cdouble; primal_obj::Cdouble; rel_gap::Cdouble; dual_obj::Cdouble
    CResiduals() = new()
end

const BC = Dict(:l2 => 0, :rms => 1, :linf => 2, :l_inf => 2, :inf => 2)   # hprsocp_bc_norm
const STATUS = ("OPTIMAL", "MAX_ITER", "TIME_LIMIT", "CONTINUE")
# hprsocp_soc_strategy codes 1..20, in the order of scaling.jl:152-171
const SOC = (:max, :rms, :geom, :adaptive, :hybrid, :ruiz_geom, :role_split,
             :pock_safe_role_split, :size_split_role_split, :pre_pock_size_split,
             :pre_pock_huge_var_taper, :pre_pock_huge_var_taper_pock_geom, :pre_pock_band_split,
             :pre_pock_constraint_rms, :pre_pock_constraint_max, :pre_pock_constraint_taper,
             :small_cone_constraint_max, :count_aware_small_cone_bias, :phase_blend, :phase_taper)
soc_strategy_code(s::Symbol) = something(findfirst(==(s), SOC), 0)

const HPRSOCP_Q_CSR = Cint(0)
const HPRSOCP_Q_FACTORED = Cint(1)
const ITERATE_RESET = Cint(1)

last_error(h::Ptr{Cvoid}) = unsafe_string(ccall((:hprsocp_last_error, lib), Cstring, (Ptr{Cvoid},), h))
check(rc, h::Ptr{Cvoid} = C_NULL) = rc == 0 || error("hprsocp: " * last_error(h))

function cparams(p::HPRSOCP_parameters; precision::Symbol = :fp64)
    q = CParams()
    ccall((:hprsocp_default_params, lib), Cvoid, (Ref{CParams},), q)
    q.stoptol = p.stoptol; q.sigma = p.sigma; q.max_iter = p.max_iter
    q.time_limit = p.time_limit; q.check_iter = p.check_iter; q.warm_up = p.warm_up
    q.print_frequency = p.print_frequency; q.device_number = p.device_number
    q.use_Ruiz_scaling = p.use_Ruiz_scaling; q.ruiz_iterations = p.ruiz_iterations
    q.use_bc_scaling = p.use_bc_scaling; q.bc_scaling_norm_type = BC[p.bc_scaling_norm_type]
    q.use_Pock_Chambolle_scaling = p.use_Pock_Chambolle_scaling
    code = soc_strategy_code(p.soc_block_scaling_strategy)
    code == 0 && error("Unsupported soc_block_scaling_strategy=$(p.soc_block_scaling_strategy)")
    q.soc_block_scaling_strategy = code
    q.verbose = p.verbose; q.use_gpu = true
    q.precision = precision === :fp32_delta ? 2 : precision === :fp32_refine ? 1 : 0
    # the caller keeps p.save_filename / p.initial_x / p.initial_y alive (GC.@preserve)
    q.auto_save = p.auto_save
    q.save_filename = Base.unsafe_convert(Cstring, p.save_filename)
    q.initial_x = p.initial_x === nothing ? C_NULL : pointer(p.initial_x)
    q.initial_y = p.initial_y === nothing ? C_NULL : pointer(p.initial_y)
    return q
end

# ---------------------------------------------------------------- the problem
# An operator Q (AbstractQOperatorCPU, Q_operator_interface.jl:33-151) runs on the
# backend only in its factored form Q = F'F + Diagonal(d).  An operator type
# provides it with a method
#     HPRSOCPHip.hip_factored(Q::MyQOperatorCPU) -> (F::SparseMatrixCSC, d::Vector{Float64})
# (d may be `nothing`); an operator without one is rejected, never solved as Q = 0.
function hip_factored end

# Arrays of one hprsocp_problem: 0-based CSR (the CSC arrays of the transposes).
struct ProblemArrays
    rp::Vector{Cint}; ci::Vector{Cint}; av::Vector{Float64}
    qkind::Cint
    qrp::Vector{Cint}; qci::Vector{Cint}; qv::Vector{Float64}
    frows::Cint; frp::Vector{Cint}; fci::Vector{Cint}; fv::Vector{Float64}; qdiag::Vector{Float64}
    con::Vector{Cint}; var::Vector{Cint}
end

csr0(M::SparseMatrixCSC) = (Mt = sparse(transpose(M)); (Cint.(Mt.colptr .- 1), Cint.(Mt.rowval .- 1), Vector{Float64}(Mt.nzval)))

function problem_arrays(model::QP_info_cpu)
    AT = model.AT                                        # CSC of A' = CSR of A
    rp, ci, av = Cint.(AT.colptr .- 1), Cint.(AT.rowval .- 1), Vector{Float64}(AT.nzval)
    Q = model.Q
    empty = (Cint[0], Cint[], Float64[])
    if Q isa SparseMatrixCSC
        # CSC of Q = CSR of Q' ; the backend symmetrises (Q + Q')/2, so either is Q's operator
        qrp, qci, qv = nnz(Q) > 0 ? (Cint.(Q.colptr .- 1), Cint.(Q.rowval .- 1), Vector{Float64}(Q.nzval)) : empty
        fr = (Cint(0), Cint[0], Cint[], Float64[], Float64[])
        qkind = HPRSOCP_Q_CSR
    elseif Q isa AbstractQOperatorCPU
        hasmethod(hip_factored, Tuple{typeof(Q)}) ||
            error("HPRSOCPHip: Q operator $(typeof(Q)) has no factored form; define " *
                  "HPRSOCPHip.hip_factored(::$(typeof(Q))) -> (F, d) with Q = F'F + Diagonal(d)")
        F, d = hip_factored(Q)
        size(F, 2) == size(model.A, 2) || error("HPRSOCPHip: F must have n = $(size(model.A, 2)) columns")
        frp, fci, fv = csr0(SparseMatrixCSC{Float64}(F))
        fr = (Cint(size(F, 1)), frp, fci, fv, d === nothing ? Float64[] : Vector{Float64}(d))
        qrp, qci, qv = empty
        qkind = HPRSOCP_Q_FACTORED
    else
        error("HPRSOCPHip: unsupported Q type $(typeof(Q))")
    end
    ProblemArrays(rp, ci, av, qkind, qrp, qci, qv, fr[1], fr[2], fr[3], fr[4], fr[5],
                  Cint.(model.SOC_con_idx .- 1), Cint.(model.SOC_var_idx .- 1))
end

ptr_or_null(v::Vector) = isempty(v) ? Ptr{eltype(v)}(C_NULL) : pointer(v)

# Only valid inside GC.@preserve of `model` and `a`.
function cproblem(model::QP_info_cpu, a::ProblemArrays)
    m, n = size(model.A)
    CProblem(m, n, length(a.ci), pointer(a.rp), ptr_or_null(a.ci), ptr_or_null(a.av),
             a.qkind, length(a.qv), pointer(a.qrp), ptr_or_null(a.qci), ptr_or_null(a.qv),
             a.frows, length(a.fv), pointer(a.frp), ptr_or_null(a.fci), ptr_or_null(a.fv), ptr_or_null(a.qdiag),
             pointer(model.c), pointer(model.AL), pointer(model.AU),
             pointer(model.l), pointer(model.u), ptr_or_null(model.soc_rhs),
             model.number_linear_con, model.number_lu_x,
             length(a.con) - 1, pointer(a.con), length(a.var) - 1, pointer(a.var),
             model.obj_constant)
end

function results(res::CResult, x, y, z, w)
    r = HPRSOCP_results()
    r.iter, r.iter_4, r.iter_6 = res.iter, res.iter_4, res.iter_6
    r.time, r.time_4, r.time_6, r.power_time = res.time, res.time_4, res.time_6, res.power_time
    r.primal_obj, r.residuals, r.gap = res.primal_obj, res.residuals, res.gap
    r.status = STATUS[res.status + 1]
    r.x, r.y, r.z, r.w = x, y, z, w
    return r
end

# ---------------------------------------------------------------- optimize
function optimize_hip(model::QP_info_cpu, p::HPRSOCP_parameters; precision::Symbol = :fp64)
    a = problem_arrays(model)
    m, n = size(model.A)
    x = zeros(n); y = zeros(m); z = zeros(n); w = zeros(n)
    res = Ref{CResult}()
    GC.@preserve a model x y z w p begin
        prob = cproblem(model, a)
        rc = ccall((:hprsocp_optimize, lib), Cint,
                   (Ref{CProblem}, Ref{CParams}, Ref{CResult},
                    Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                   prob, cparams(p; precision), res, x, y, z, w)
        check(rc)
    end
    return results(res[], x, y, z, w)
end

# ---------------------------------------------------------------- batch
# optimize() of every model (each instance follows algorithm.jl:2538-2581), the
# instances packed block-diagonally and iterated together on the backend.
# x0s / y0s: optional per-instance initial points (`nothing` entries: cold start).
function optimize_batch_hip(models::AbstractVector{<:QP_info_cpu}, p::HPRSOCP_parameters;
                            precision::Symbol = :fp64, x0s = nothing, y0s = nothing)
    k = length(models)
    arrs = [problem_arrays(md) for md in models]
    res = Vector{CResult}(undef, k)
    out = HPRSOCP_results[]
    br = Ref{Ptr{Cvoid}}(C_NULL)
    bcheck(rc) = rc == 0 || error("hprsocp: " *
        unsafe_string(ccall((:hprsocp_batch_last_error, lib), Cstring, (Ptr{Cvoid},), br[])))
    GC.@preserve arrs models p x0s y0s begin
        # one Ref per CProblem, kept alive while create copies the problems
        refs = [Ref(cproblem(models[i], arrs[i])) for i in 1:k]
        GC.@preserve refs begin
            ptrs = [Base.unsafe_convert(Ptr{CProblem}, r) for r in refs]
            check(ccall((:hprsocp_batch_create, lib), Cint,
                        (Ptr{Ptr{CProblem}}, Cint, Ref{CParams}, Ref{Ptr{Cvoid}}),
                        ptrs, k, cparams(p; precision), br))
        end
        try
            bcheck(ccall((:hprsocp_batch_estimate_lambda, lib), Cint, (Ptr{Cvoid},), br[]))
            if x0s !== nothing || y0s !== nothing
                pick(v, i) = (v === nothing || v[i] === nothing) ? Ptr{Cdouble}(C_NULL) : pointer(v[i])
                xp = [pick(x0s, i) for i in 1:k]; yp = [pick(y0s, i) for i in 1:k]
                bcheck(ccall((:hprsocp_batch_set_initial, lib), Cint,
                             (Ptr{Cvoid}, Ptr{Ptr{Cdouble}}, Ptr{Ptr{Cdouble}}), br[], xp, yp))
            end
            bcheck(ccall((:hprsocp_batch_solve, lib), Cint, (Ptr{Cvoid}, Ptr{CResult}), br[], res))
            for i in 1:k
                m, n = size(models[i].A)
                x = zeros(n); y = zeros(m); z = zeros(n); w = zeros(n)
                bcheck(ccall((:hprsocp_batch_get_solution, lib), Cint,
                             (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                             br[], i - 1, x, y, z, w))
                push!(out, results(res[i], x, y, z, w))
            end
        finally
            ccall((:hprsocp_batch_destroy, lib), Cvoid, (Ptr{Cvoid},), br[])
        end
    end
    return out
end

# ---------------------------------------------------------------- seams
# A handle on the device: the scaled problem, its workspace and stream.
mutable struct HipWorkspace
    h::Ptr{Cvoid}
    m::Int; n::Int
    function HipWorkspace(model::QP_info_cpu, p::HPRSOCP_parameters; precision::Symbol = :fp64)
        a = problem_arrays(model)
        hr = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve a model p begin
            prob = cproblem(model, a)
            check(ccall((:hprsocp_create, lib), Cint, (Ref{CProblem}, Ref{CParams}, Ref{Ptr{Cvoid}}),
                        prob, cparams(p; precision), hr))
        end
        m, n = size(model.A)
        ws = new(hr[], m, n)
        finalizer(ws -> ccall((:hprsocp_destroy, lib), Cvoid, (Ptr{Cvoid},), ws.h), ws)
        return ws
    end
end

estimate_eigenvalues!(ws::HipWorkspace) = check(ccall((:hprsocp_estimate_lambda, lib), Cint, (Ptr{Cvoid},), ws.h), ws.h)

# main_update_gpu! (algorithm.jl:2284-2309): Halpern factor from restart_info.inner
hip_update!(ws::HipWorkspace, restart_info::HPRSOCP_restart, to_check::Bool) =
    check(ccall((:hprsocp_update, lib), Cint, (Ptr{Cvoid}, Int64, Cint), ws.h, restart_info.inner, to_check), ws.h)

# compute_residuals! (algorithm.jl:113-196)
function hip_residuals!(ws::HipWorkspace, iter::Int)
    r = CResiduals()
    check(ccall((:hprsocp_compute_residuals, lib), Cint, (Ptr{Cvoid}, Int64, Ref{CResiduals}), ws.h, iter, r), ws.h)
    return r
end

# compute_M_norm! (algorithm.jl:1302-1359)
function hip_mnorm(ws::HipWorkspace)
    v = Ref{Cdouble}(0.0)
    check(ccall((:hprsocp_mnorm, lib), Cint, (Ptr{Cvoid}, Ref{Cdouble}), ws.h, v), ws.h)
    return v[]
end

# update_sigma! (algorithm.jl:729-990) and do_restart! (:1130-1166) on the
# reference's own restart_info object
hip_update_sigma!(ws::HipWorkspace, restart_info::HPRSOCP_restart, r::CResiduals) =
    check(ccall((:hprsocp_update_sigma, lib), Cint, (Ptr{Cvoid}, Ref{HPRSOCP_restart}, Ref{CResiduals}),
                ws.h, restart_info, r), ws.h)
hip_do_restart!(ws::HipWorkspace, restart_info::HPRSOCP_restart) =
    check(ccall((:hprsocp_restart, lib), Cint, (Ptr{Cvoid}, Ref{HPRSOCP_restart}), ws.h, restart_info), ws.h)

# k bodies of the solve loop on the backend (algorithm.jl:2806-2904)
function hip_iterate!(ws::HipWorkspace, k::Integer; reset::Bool = false)
    res = Ref{CResult}()
    check(ccall((:hprsocp_iterate, lib), Cint, (Ptr{Cvoid}, Int64, Cint, Ref{CResult}),
                ws.h, k, reset ? ITERATE_RESET : Cint(0), res), ws.h)
    return res[]
end

function hip_solution(ws::HipWorkspace)
    x = zeros(ws.n); y = zeros(ws.m); z = zeros(ws.n); w = zeros(ws.n)
    check(ccall((:hprsocp_get_solution, lib), Cint,
                (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}), ws.h, x, y, z, w), ws.h)
    return x, y, z, w
end

end # module
