// ctl.hip — the device controller of a batch solve: the window-edge step of the
// solve loop (algorithm.jl:2806-2867) for every instance at once, queued on the
// stream after the residual (and, at checks, sigma-dot and M-norm) reductions,
// so that no decision needs a host round trip:
//   residual block -> deferred M-norms (checks) -> termination status ->
//   check_restart -> update_sigma! -> do_restart! flags and counters ->
//   milestones -> retirement -> Halpern counter base.
// The arithmetic is ctl.h's, the same functions the host loop runs, so a batch
// instance takes the decisions of its single-problem solve bit for bit.
#include <hip/hip_runtime.h>

#include "ctl.h"
#include "ctlk.h"
#include "kernels.h"

namespace hpr {

__global__ __launch_bounds__(256) void k_ctl_event(CtlArgs a) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nb; i += gridDim.x * blockDim.x) {
        ctl::CtlInst &C = a.ci[i];
        InstScal &S = a.isc[i];
        hprsocp_restart_info &R = C.R;
        const double *v = a.slots + (int64_t)i * a.slot_stride;
        const ctl::InstConst &H = a.ic[i];
        const bool skip_q = a.skip_q != 0, qd = a.q_diag != 0, hm = H.has_m != 0;
        if (a.phase != 2) {
            a.flag[i] = 0;
            if (!S.active) {
                a.status[i] = C.status;
                continue;
            }
            R.inner += a.iter - C.last_event;  // updates since the previous event
            C.last_event = a.iter;
            ctl::residuals(v + a.sl_rn, v + a.sl_rm, v[a.sl_bv], S.b_scale, S.c_scale, H, skip_q, a.iter == 0, C.rs);
            if (a.is_check) {  // deferred M-norms (perform_iteration_step!, algorithm.jl:2731-2741)
                if (C.pend[0])
                    R.last_gap = ctl::jmin(
                        ctl::mnorm(v + a.sl_mn0, C.mn_sig[0], C.mn_lamA[0], C.mn_lamQ[0], !skip_q, hm, qd), R.last_gap);
                if (C.pend[1])
                    R.current_gap = ctl::mnorm(v + a.sl_mn1, C.mn_sig[1], C.mn_lamA[1], C.mn_lamQ[1], !skip_q, hm, qd);
                C.pend[0] = C.pend[1] = 0;
            }
            int32_t st = HPRSOCP_STATUS_CONTINUE;  // check_break, algorithm.jl:1169-1187
            if (C.rs.kkt < a.stoptol) st = HPRSOCP_STATUS_OPTIMAL;
            else if (a.iter == a.max_iter) st = HPRSOCP_STATUS_MAX_ITER;
            else if (a.time_up) st = HPRSOCP_STATUS_TIME_LIMIT;
            C.st_pend = st;
            ctl::check_restart(R, a.iter, a.check_iter, S.sigma, C.rs.kkt);
            C.gold = 0;
            if (R.restart_flag >= 1) {  // update_sigma!, from this event's dots
                double sa, sb, sc, sd;
                ctl::sigma_terms(v + a.sl_sig, skip_q, hm, qd, S.lambda_A, S.lambda_Q, sa, sb, sc, sd);
                if (!skip_q && hm && qd) {  // unified_golden_Q_diag over the device vectors, between the phases
                    gw_init(a.gw[i], sa, sb, 1e-12, 1e12, 1e-13, kGoldenL);
                    C.gold = 1;
                } else {
                    C.sm = ctl::sigma_model(sa, sb, sc, sd, skip_q, hm, qd);
                }
            }
            if (a.phase == 1) continue;
        } else if (!S.active) {
            continue;
        }
        if (R.restart_flag >= 1) {
            double sm = C.sm;
            if (C.gold) {
                sm = gw_result(a.gw[i]);
                a.gw[i].active = 0;
                C.gold = 0;
            }
            S.sigma = ctl::sigma_next(R, C.rs, sm, S.sigma);
        }
        if (R.restart_flag > 0) {  // do_restart!: the copies run next, flag-driven
            a.flag[i] = 1;
            R.last_gap = R.current_gap;
            R.save_gap = INFINITY;
            R.times += 1;
            R.inner = 0;
        }
        if (C.rs.kkt < 1e-4 && C.first4) {
            C.it4 = a.iter;
            C.first4 = 0;
        }
        if (C.rs.kkt < 1e-6 && C.first6) {
            C.it6 = a.iter;
            C.first6 = 0;
        }
        const int32_t st = C.st_pend;
        if (st != HPRSOCP_STATUS_CONTINUE) {
            C.status = st;
            C.iter_end = a.iter;
            C.rs_end = C.rs;
            C.sigma_end = S.sigma;
            C.restarts_end = R.times;
            S.active = 0;
            a.status[i] = st;
            continue;
        }
        S.inner0 = (double)R.inner;
        if (a.is_check) {  // the M-norm after this check's update feeds last_gap at the next check
            C.pend[0] = R.restart_flag > 0;
            C.mn_sig[0] = S.sigma;
            C.mn_lamA[0] = S.lambda_A;
            C.mn_lamQ[0] = S.lambda_Q;
        }
        if (a.next_is_check) {
            C.pend[1] = 1;
            C.mn_sig[1] = S.sigma;
            C.mn_lamA[1] = S.lambda_A;
            C.mn_lamQ[1] = S.lambda_Q;
        }
        a.status[i] = HPRSOCP_STATUS_CONTINUE;
    }
}

void launch_ctl_event(hipStream_t st, const CtlArgs &a) {
    if (a.nb <= 0) return;
    const int blocks = (a.nb + 255) / 256;
    hipLaunchKernelGGL(k_ctl_event, dim3((unsigned)blocks), dim3(256), 0, st, a);
}

static_assert(kGoldenMaxPts <= kMaxGolden, "a walk's points fit its row of xs");

__global__ void k_gold_step(GoldenWalk *gw, int32_t nb, const double *slots, int32_t stride, int32_t sl_gold,
                            double *xs, int32_t *np) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    GoldenWalk &g = gw[i];
    if (np[i] > 0) gw_consume(g, slots + (int64_t)i * stride + sl_gold);
    np[i] = g.active && !gw_done(g) ? gw_next_points(g, xs + (int64_t)i * kMaxGolden) : 0;
}

void launch_gold_step(hipStream_t st, GoldenWalk *gw, int32_t nb, const double *slots, int32_t stride,
                      int32_t sl_gold, double *xs, int32_t *np) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_gold_step, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, gw, nb, slots, stride,
                       sl_gold, xs, np);
}

// InstScal::fp32 of every instance, leaving the other fields (which the
// controller owns on the device) alone.
__global__ void k_isc_fp32(InstScal *isc, int32_t nb, int32_t f32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) isc[i].fp32 = f32;
}

void launch_isc_fp32(hipStream_t st, InstScal *isc, int32_t nb, int32_t f32) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_isc_fp32, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, isc, nb, f32);
}

// nflag[i] = !flag[i]
__global__ void k_flag_not(const int32_t *flag, int32_t *nflag, int32_t nb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) nflag[i] = flag[i] == 0;
}

void launch_flag_not(hipStream_t st, const int32_t *flag, int32_t *nflag, int32_t nb) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_flag_not, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, flag, nflag, nb);
}

// The shared control arithmetic on the device, for the determinism check:
// out[5k .. 5k+4] = exp(x), log(|x|), sqrt(|x|), the general-Q golden section
// of (|x|, 1 / (1 + |x|), x^2, 0.5 |x|), and sigma_next of a fixed restart
// state with sm = |x| + 1e-3.
__global__ void k_ctl_probe(const double *in, int32_t n, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    ctl_probe_eval(in[k], out + 5 * (int64_t)k);
}

void launch_ctl_probe(hipStream_t st, const double *in, int32_t n, double *out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ctl_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, n, out);
}

}  // namespace hpr
