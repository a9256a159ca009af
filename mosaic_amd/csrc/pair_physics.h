// Per-pair scattering factors of the exact kernels (k_sum_exact in
// kinematic.hip, k_scatter in scatter.hip), in FP32 once the geometry is
// known:
//   * the Waasmaier-Kirfel form factor as a function of omc = 1 - cos 2theta
//     (Q = k sqrt(2 omc), s = Q 0.25e-10 / pi; Beam.py:669-679);
//   * the polarization factor sqrt(P), P = (1-p)(1+cos^2 2theta)/2 +
//     p |e x r_hat|^2 clipped to [0, 1] (Beam.py:785-800; the cross-product
//     form of 1 - (e.r_hat)^2 keeps its relative precision where e.r_hat -> 1,
//     as the reference's general kernel does, Beam.py:2329-2343);
//   * the analyser acceptance of the general kernel (Beam.py:2350-2378):
//     top-hat (kind 1, zero outside the cone of half-angle accept about the
//     analyser axis) or Darwin roll-off (kind 2, 1 / (1 + (delta / hw)^2)),
//     with delta the angle between the atom->pixel ray and the axis.
#pragma once

#include <hip/hip_runtime.h>

namespace mosaic {

constexpr double LOG2E_D = 1.4426950408889634;
constexpr double S_PER_Q = 0.25e-10 / 3.14159265358979323846;   // s [1/A] per Q [1/m]

// A WK row prepared for evaluation in omc:
//   f0 = c + sum_m a_m 2^(nb_m omc),  nb_m = -b_m (2 k^2 S_PER_Q^2) log2(e),
// plus f0(0) and the anomalous terms.
struct FormRow {
    float a[5];
    float nb[5];
    float c, f0z, fp, fpp;
};

inline FormRow make_form_row(const float* wk, float f0z, float fp, float fpp, double k) {
    FormRow r;
    const double cq = 2.0 * k * k * S_PER_Q * S_PER_Q * LOG2E_D;
    for (int m = 0; m < 5; ++m) {
        r.a[m] = wk[m];
        r.nb[m] = static_cast<float>(-static_cast<double>(wk[6 + m]) * cq);
    }
    r.c = wk[5];
    r.f0z = f0z;
    r.fp = fp;
    r.fpp = fpp;
    return r;
}

__device__ __forceinline__ float form_f0(const float (&a)[5], const float (&nb)[5], float c,
                                         float omc) {
    float f = c;
#pragma unroll
    for (int m = 0; m < 5; ++m) f = fmaf(a[m], __builtin_amdgcn_exp2f(nb[m] * omc), f);
    return f;
}

// sqrt(P) from cos 2theta and |e x r_hat|^2.
__device__ __forceinline__ float pol_factor(float cos2t, float ecross2, float p) {
    float P = fmaf(0.5f * (1.0f - p), fmaf(cos2t, cos2t, 1.0f), p * ecross2);
    P = fminf(1.0f, fmaxf(0.0f, P));
    return __builtin_sqrtf(P);
}

// Analyser weight from sin and cos of the ray's angle to the analyser axis.
// kind 1: 1 inside the cone delta <= accept, else 0; kind 2: Darwin roll-off
// with inv_hw = 1 / half-width (0: no roll-off).
template <int KIND>
__device__ __forceinline__ float analyser_weight(float sin_d, float cos_d, float accept,
                                                 float inv_hw) {
    const float delta = atan2f(sin_d, cos_d);
    if (KIND == 1) return delta <= accept ? 1.0f : 0.0f;
    if (KIND == 2) {
        const float r = delta * inv_hw;
        return 1.0f / fmaf(r, r, 1.0f);
    }
    return 1.0f;
}

}  // namespace mosaic
