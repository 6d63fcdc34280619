// ctlk.h — launch interface of the device controller (ctl.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ctl.h"
#include "golden.h"
#include "hotpath.h"

namespace hpr {

// Arguments of one window-edge step (an event: a residual check or print step).
struct CtlArgs {
    ctl::CtlInst *ci;           // nb
    InstScal *isc;              // nb: sigma, Halpern base and retirement are written here
    const ctl::InstConst *ic;   // nb
    int32_t *flag;              // nb: restart flags for the flag-driven copies
    int32_t *status;            // nb: hprsocp_status after this event (CONTINUE while running)
    const double *slots;        // per-instance reduction slots, slot_stride apart
    int32_t slot_stride, sl_rn, sl_rm, sl_bv, sl_mn0, sl_mn1, sl_sig;
    int32_t nb;
    int64_t iter, check_iter, max_iter;
    double stoptol;
    int32_t is_check, next_is_check, time_up, skip_q, q_diag;
    // 0: the whole step; with a diagonal Q the step runs in two phases around the
    // golden-section passes: 1 up to the sigma model (walks started in gw), 2 from
    // the sigma update on
    int32_t phase;
    GoldenWalk *gw;  // nb
};

void launch_ctl_event(hipStream_t st, const CtlArgs &a);
// One pass of the instances' golden-section walks: consume the g values of the
// previous pass (slots[i * stride + sl_gold ..], when np[i] > 0), then write the
// next points to xs[i * kMaxGolden ..] and their number to np[i] (0: no search
// or done).  kGoldenPasses evaluations between kGoldenPasses + 1 steps finish
// every walk and leave np = 0.
void launch_gold_step(hipStream_t st, GoldenWalk *gw, int32_t nb, const double *slots, int32_t stride,
                      int32_t sl_gold, double *xs, int32_t *np);
void launch_isc_fp32(hipStream_t st, InstScal *isc, int32_t nb, int32_t f32);
void launch_flag_not(hipStream_t st, const int32_t *flag, int32_t *nflag, int32_t nb);

// The determinism probe's evaluation, the same on host and device.
__host__ __device__ inline void ctl_probe_eval(double x, double *o) {
    HPR_NOFMA
    const double ax = fabs(x);
    o[0] = ctl::ctl_exp(x);
    o[1] = ctl::ctl_log(ax);
    o[2] = sqrt(ax);
    o[3] = ctl::golden_scalar(ax, 1.0 / (1.0 + ax), x * x, 0.5 * ax, 1e-12, 1e12, 1e-13);
    hprsocp_restart_info r;
    ctl::init_restart(r);
    r.first_restart = 0;
    r.best_kkt = 1e-3;
    r.best_gap = 1.0;
    r.current_gap = 1.0 + 0.5 * ax;
    r.best_sigma = 2.0;
    hprsocp_residuals res{};
    res.kkt = 1e-3 * (1.0 + ax);
    res.err_Rp = 1e-9 * (1.0 + ax);
    res.err_Rd = 3e-8;
    o[4] = ctl::sigma_next(r, res, ax + 1e-3, 0.7);
}
void launch_ctl_probe(hipStream_t st, const double *in, int32_t n, double *out);

}  // namespace hpr
