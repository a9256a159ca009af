// Shared helpers for the gfx950 kernels: status plumbing, double-double
// arithmetic and workspace carving.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "mosaic_amd.h"

#define MOSAIC_HIP_CHECK(expr)                                   \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return MOSAIC_ERR_HIP;             \
    } while (0)

namespace mosaic {

// 1 / (2 pi) as an unevaluated double-double hi + lo.
constexpr double INV_2PI_HI = 0.15915494309189535;
constexpr double INV_2PI_LO = -9.839338337591243e-18;

__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    e = fma(a, b, -p);
}

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// frac(K |T|) in double-double (K = k / 2 pi), with |T| = r + r_lo.
__device__ __forceinline__ void origin_phase(double tx, double ty, double tz, double k,
                                             double& f, double& r, double& r_lo) {
    // |T|^2 in double-double
    double a, ae, b, be, c, ce, s, se, s2, se2;
    two_prod(tx, tx, a, ae);
    two_prod(ty, ty, b, be);
    two_prod(tz, tz, c, ce);
    two_sum(a, b, s, se);
    two_sum(s, c, s2, se2);
    const double sq_hi = s2, sq_lo = (se + se2) + (ae + be + ce);
    r = sqrt(sq_hi);
    double rr, rre;
    two_prod(r, r, rr, rre);
    const double resid = ((sq_hi - rr) - rre) + sq_lo;
    r_lo = (r > 0.0) ? resid / (2.0 * r) : 0.0;
    double K_hi, K_lo;
    two_prod(k, INV_2PI_HI, K_hi, K_lo);
    K_lo = fma(k, INV_2PI_LO, K_lo);
    double th, the;
    two_prod(K_hi, r, th, the);
    the += K_hi * r_lo + K_lo * r;
    const double n = floor(th);
    f = (th - n) + the;
    f -= floor(f);
}

// Carves aligned sub-buffers out of one caller-provided workspace.  With a
// null base it only measures.
struct Carver {
    char* base;
    size_t off = 0;
    explicit Carver(void* b) : base(static_cast<char*>(b)) {}
    template <typename T>
    T* take(size_t count, size_t align = 256) {
        off = (off + align - 1) / align * align;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
};

inline int blocks_for(int64_t n, int threads) {
    return static_cast<int>((n + threads - 1) / threads);
}

}  // namespace mosaic
