// scale_rules.h — SOC block aggregation rule of the equilibration passes.
// Restates _soc_block_scale_value_scalar (reference src/utils/scaling.jl:253-477):
// given a cone block's length L, the max and rms of its per-row norms, the
// strategy code, the phase (generic/ruiz/pre_pock/pock) and the location
// (generic/constraint/variable), return the one factor shared by the block.
#pragma once
#include <math.h>
#include <stdint.h>

#ifndef HPR_HD
#if defined(__HIPCC__)
#define HPR_HD __host__ __device__
#else
#define HPR_HD
#endif
#endif

namespace hpr {

enum ScalePhase : int32_t { PHASE_GENERIC = 0, PHASE_RUIZ = 1, PHASE_PRE_POCK = 2, PHASE_POCK = 3 };
enum ScaleLoc : int32_t { LOC_GENERIC = 0, LOC_CONSTRAINT = 1, LOC_VARIABLE = 2 };

constexpr int32_t kAdaptiveSocBlock = 8;      // ADAPTIVE_SOC_BLOCK_THRESHOLD
constexpr int32_t kPockRmsSocBlock = 512;     // POCK_RMS_SOC_BLOCK_THRESHOLD
constexpr int32_t kManySocCones = 100000;     // MANY_SOC_CONSTRAINT_CONES_THRESHOLD

// strategies that run two extra max-passes between Ruiz and Pock (uses_pre_pock_stage)
HPR_HD inline bool uses_pre_pock(int32_t s) { return s >= 10 && s <= 18; }

HPR_HD inline double soc_block_value(int32_t L, double bm, double rms, int32_t s, int32_t ph,
                                     int32_t loc, int32_t cone_count) {
    if (L <= 0) return 1.0;
    if (s == 1) return bm;  // :max
    const double geom = sqrt(bm * rms);
    const bool small = L <= kAdaptiveSocBlock;
    const bool ruiz = ph == PHASE_RUIZ, pre = ph == PHASE_PRE_POCK, pock = ph == PHASE_POCK;
    const bool con = loc == LOC_CONSTRAINT;
    // variable-side rule shared by the pre-Pock family (strategies 10-18)
    auto pre_family_var = [&](double pre_large) -> double {
        if (ruiz) return small ? bm : rms;
        if (pre) return L <= kPockRmsSocBlock ? rms : pre_large;
        if (pock) return bm;
        return geom;
    };
    switch (s) {
    case 2: return rms;
    case 3: return geom;
    case 4: return small ? bm : rms;
    case 5: return ruiz ? (small ? bm : rms) : (pock ? rms : geom);
    case 6: return ruiz ? (small ? bm : geom) : (pock ? rms : geom);
    case 7:
        if (con) return geom;
        return ruiz ? (small ? bm : rms) : (pock ? rms : geom);
    case 8:
        if (pock) return bm;
        if (con) return geom;
        return ruiz ? (small ? bm : rms) : geom;
    case 9:
        if (con) return pock ? bm : geom;
        if (ruiz) return small ? bm : rms;
        if (pock) return L <= kPockRmsSocBlock ? rms : bm;
        return geom;
    case 10:
        if (con) return pock ? bm : geom;
        return pre_family_var(bm);
    case 11:
        if (con) return pock ? bm : geom;
        return pre_family_var(sqrt(sqrt(bm * rms * rms * rms)));
    case 12:
        if (con) return pock ? bm : geom;
        if (ruiz) return small ? bm : rms;
        if (pre) return L <= kPockRmsSocBlock ? rms : sqrt(sqrt(bm * rms * rms * rms));
        if (pock) return L <= kPockRmsSocBlock ? bm : geom;
        return geom;
    case 13:
        if (con) return pock ? bm : geom;
        if (ruiz) return small ? bm : rms;
        if (pre) return small ? bm : (L <= kPockRmsSocBlock ? rms : bm);
        if (pock) return bm;
        return geom;
    case 14:
        if (con) return pock ? bm : (pre ? rms : geom);
        return pre_family_var(bm);
    case 15:
        if (con) return (pock || pre) ? bm : geom;
        return pre_family_var(bm);
    case 16:
        if (con) return pock ? bm : (pre ? sqrt(sqrt(bm * rms * rms * rms)) : geom);
        return pre_family_var(bm);
    case 17:
        if (con) {
            if (pock) return bm;
            if (ruiz || pre) return small ? bm : geom;
            return geom;
        }
        return pre_family_var(bm);
    case 18:
        if (con) {
            const bool many_small = cone_count >= kManySocCones && small;
            if (pock) return bm;
            if (many_small) {
                if (ruiz) return sqrt(sqrt(bm * bm * bm * rms));
                if (pre) return bm;
            }
            return geom;
        }
        return pre_family_var(bm);
    case 19:
    case 20:  // :phase_blend and :phase_taper share one rule
        if (ruiz) return small ? bm : rms;
        if (pock) return small ? rms : geom;
        return geom;
    default: return bm;
    }
}

}  // namespace hpr
