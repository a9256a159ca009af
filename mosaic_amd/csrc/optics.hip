// Optics elements and detector post-processing on gfx950, beside the thin
// lens and free space of propagate.hip:
//  * two-bounce Bragg magnifier, a geometric resampling (Optics.py:132-209);
//  * angular filter, an amplitude mask on the field's 2-D spectrum
//    (Optics.py:211-336), over a hipFFT round trip;
//  * Fresnel zone plate, a real-space binary (or ideal-lens) transmission
//    (Optics.py:338-425);
//  * aperture, square or circular (Optics.py:427-471);
//  * detector point-spread function: the intensity (or amplitude) convolved
//    with a normalised kernel under scipy.ndimage's boundary modes, phase kept
//    (Detector.py:873-1087).
// Fields are complex64 (rows, cols) row-major.  Each element follows the
// reference's float32 arithmetic in its operation order (its NumPy/CuPy
// expressions), so results agree with it to float32 rounding; the PSF sums in
// FP64 as scipy does.
#include <hipfft/hipfft.h>

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

#include "common.h"

namespace mosaic {
namespace {

constexpr int OPT_THREADS = 256;

__global__ void k_magnifier(const float2* __restrict__ F, float2* __restrict__ out, int Ny,
                            int Nx, float cx, float cy, float Mx, float My, int order,
                            int pad_edge, float sr, float si) {
#pragma clang fp contract(off)
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(Ny) * Nx) return;
    const int iy = static_cast<int>(idx / Nx), ix = static_cast<int>(idx % Nx);
    const float Xi = cx + (static_cast<float>(ix) - cx) / Mx;
    const float Yi = cy + (static_cast<float>(iy) - cy) / My;
    float2 v;
    if (order == 0) {
        const int xn = static_cast<int>(rintf(Xi)), yn = static_cast<int>(rintf(Yi));
        const bool valid = xn >= 0 && xn < Nx && yn >= 0 && yn < Ny;
        v = valid ? F[static_cast<int64_t>(yn) * Nx + xn] : make_float2(0.f, 0.f);
    } else if (pad_edge) {
        // the reference's weights here are float64 (float32 minus int32), the
        // products complex128, cast to complex64 at the end
        const int x0 = static_cast<int>(floorf(Xi)), y0 = static_cast<int>(floorf(Yi));
        const int x0c = min(max(x0, 0), Nx - 1), x1c = min(max(x0 + 1, 0), Nx - 1);
        const int y0c = min(max(y0, 0), Ny - 1), y1c = min(max(y0 + 1, 0), Ny - 1);
        const double wx = static_cast<double>(Xi) - x0c, wy = static_cast<double>(Yi) - y0c;
        const float2 f00 = F[static_cast<int64_t>(y0c) * Nx + x0c];
        const float2 f10 = F[static_cast<int64_t>(y0c) * Nx + x1c];
        const float2 f01 = F[static_cast<int64_t>(y1c) * Nx + x0c];
        const float2 f11 = F[static_cast<int64_t>(y1c) * Nx + x1c];
        v.x = static_cast<float>(((f00.x * (1.0 - wx) * (1.0 - wy) + f10.x * wx * (1.0 - wy)) +
                                  f01.x * (1.0 - wx) * wy) + f11.x * wx * wy);
        v.y = static_cast<float>(((f00.y * (1.0 - wx) * (1.0 - wy) + f10.y * wx * (1.0 - wy)) +
                                  f01.y * (1.0 - wx) * wy) + f11.y * wx * wy);
    } else {
        const int x0 = static_cast<int>(floorf(Xi)), y0 = static_cast<int>(floorf(Yi));
        const float wx = Xi - static_cast<float>(x0), wy = Yi - static_cast<float>(y0);
        const bool valid = Xi >= 0.f && Xi < static_cast<float>(Nx - 1) && Yi >= 0.f &&
                           Yi < static_cast<float>(Ny - 1);
        const int x0c = min(max(x0, 0), max(Nx - 2, 0)), x1c = x0c + 1;
        const int y0c = min(max(y0, 0), max(Ny - 2, 0)), y1c = y0c + 1;
        if (valid) {
            const float2 f00 = F[static_cast<int64_t>(y0c) * Nx + x0c];
            const float2 f10 = F[static_cast<int64_t>(y0c) * Nx + x1c];
            const float2 f01 = F[static_cast<int64_t>(y1c) * Nx + x0c];
            const float2 f11 = F[static_cast<int64_t>(y1c) * Nx + x1c];
            const float ux = 1.f - wx, uy = 1.f - wy;
            v.x = ((f00.x * ux * uy + f10.x * wx * uy) + f01.x * ux * wy) + f11.x * wx * wy;
            v.y = ((f00.y * ux * uy + f10.y * wx * uy) + f01.y * ux * wy) + f11.y * wx * wy;
        } else {
            v = make_float2(0.f, 0.f);
        }
    }
    out[idx] = make_float2(v.x * sr - v.y * si, v.x * si + v.y * sr);
}

struct FilterArgs {
    int Ny, Nx;
    float k0;
    float cx, cy, hx, hy, cph, sph;
    int mode_1d, elliptical, rolloff, order;
    float dx, dy;
    float scale;               // 1 / (Ny Nx), the inverse transform's normalisation
};

__global__ void k_angular_mask(float2* __restrict__ S, const FilterArgs A, double dxd,
                               double dyd) {
#pragma clang fp contract(off)
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(A.Ny) * A.Nx) return;
    const int iy = static_cast<int>(idx / A.Nx), ix = static_cast<int>(idx % A.Nx);
    const int hx_ = (A.Nx - 1) / 2 + 1, hy_ = (A.Ny - 1) / 2 + 1;
    const int kxi = ix < hx_ ? ix : ix - A.Nx, kyi = iy < hy_ ? iy : iy - A.Ny;
    // numpy.fft.fftfreq: index times 1 / (n d) in float64, then float32
    const float fx = static_cast<float>(static_cast<double>(kxi) * (1.0 / (A.Nx * dxd)));
    const float fy = static_cast<float>(static_cast<double>(kyi) * (1.0 / (A.Ny * dyd)));
    const float kx = 6.2831855f * fx, ky = 6.2831855f * fy;
    const float tx = kx / A.k0, ty = ky / A.k0;
    const float tX = tx - A.cx, tY = ty - A.cy;
    const float tpar = A.cph * tX + A.sph * tY;
    const float tperp = -A.sph * tX + A.cph * tY;
    const float eps = 1e-30f;
    float H;
    if (A.mode_1d) {
        const float s = tpar / fmaxf(A.hx, eps);
        if (A.rolloff == 1) H = 1.f / (1.f + s * s);
        else if (A.rolloff == 0) H = fabsf(s) <= 1.f ? 1.f : 0.f;
        else H = 1.f / sqrtf(1.f + powf(fabsf(s), 2.f * A.order));
    } else if (A.elliptical) {
        const float u = tpar / fmaxf(A.hx, eps), v = tperp / fmaxf(A.hy, eps);
        const float r2 = u * u + v * v;
        if (A.rolloff == 1) H = 1.f / (1.f + r2);
        else if (A.rolloff == 0) H = r2 <= 1.f ? 1.f : 0.f;
        else H = 1.f / sqrtf(1.f + powf(r2, static_cast<float>(A.order)));
    } else {
        const float delta = sqrtf(tX * tX + tY * tY);
        const float hw = fmaxf(A.hx, eps);
        if (A.rolloff == 1) H = 1.f / (1.f + (delta / hw) * (delta / hw));
        else if (A.rolloff == 0) H = delta <= hw ? 1.f : 0.f;
        else H = 1.f / sqrtf(1.f + powf(delta / hw, 2.f * A.order));
    }
    const float g = H * A.scale;
    const float2 s = S[idx];
    S[idx] = make_float2(s.x * g, s.y * g);
}

__global__ void k_scale_c(float2* __restrict__ E, int64_t n, float sr, float si) {
#pragma clang fp contract(off)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = E[i];
    E[i] = make_float2(v.x * sr - v.y * si, v.x * si + v.y * sr);
}

struct ZoneArgs {
    int Ny, Nx;
    float cx, cy, dx, dy;
    float lf;              // lambda f: the zone index is n(r) = R2 / (lambda f)
    float R2max, Rstop2, R, r_inner, taper_den;
    float amp;
    float half_kf;         // -0.5 k / f_work
    int type;              // 0 amplitude, 1 phase, 2 ideal
    int stop, taper;
};

__global__ void k_zone_plate(float2* __restrict__ E, const ZoneArgs A) {
#pragma clang fp contract(off)
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(A.Ny) * A.Nx) return;
    const int iy = static_cast<int>(idx / A.Nx), ix = static_cast<int>(idx % A.Nx);
    const float x = (static_cast<float>(ix) - A.cx) * A.dx;
    const float y = (static_cast<float>(iy) - A.cy) * A.dy;
    const float R2 = x * x + y * y;
    bool inside = R2 <= A.R2max;
    if (A.stop) inside = inside && (R2 >= A.Rstop2);
    float tr = 0.f, ti = 0.f;
    if (inside) {
        if (A.type == 2) {
            const float ph = A.half_kf * R2;
            float sn, cs;
            sincosf(ph, &sn, &cs);
            tr = A.amp * cs;
            ti = A.amp * sn;
        } else {
            const float n_zone = R2 / A.lf;
            const int parity = static_cast<int>(floorf(n_zone)) % 2;
            if (parity == 0) tr = A.amp;
            else if (A.type == 1) tr = -A.amp;
        }
    }
    if (A.taper) {
        const float r = sqrtf(R2);
        if (r > A.r_inner && r <= A.R) {
            const float w = 0.5f * (1.f + cosf(3.14159265f * (r - A.r_inner) / A.taper_den));
            tr *= w;
            ti *= w;
        }
    }
    const float2 e = E[idx];
    E[idx] = make_float2(e.x * tr - e.y * ti, e.x * ti + e.y * tr);
}

__global__ void k_aperture(float2* __restrict__ E, int Ny, int Nx, float cx, float cy, float dx,
                           float dy, float half, float half2, int circular) {
#pragma clang fp contract(off)
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(Ny) * Nx) return;
    const int iy = static_cast<int>(idx / Nx), ix = static_cast<int>(idx % Nx);
    const float x = (static_cast<float>(ix) - cx) * dx;
    const float y = (static_cast<float>(iy) - cy) * dy;
    const bool keep = circular ? (x * x + y * y) <= half2
                               : (fabsf(x) <= half && fabsf(y) <= half);
    if (!keep) E[idx] = make_float2(0.f, 0.f);
}

// scipy.ndimage boundary modes: 0 constant, 1 reflect, 2 nearest, 3 mirror, 4 wrap;
// returns -1 for a constant (zero) sample
__device__ __forceinline__ int bmap(int j, int n, int mode) {
    if (j >= 0 && j < n) return j;
    switch (mode) {
        case 1: {           // reflect: d c b a | a b c d | d c b a
            const int p = 2 * n;
            int m = ((j % p) + p) % p;
            return m < n ? m : p - 1 - m;
        }
        case 2: return j < 0 ? 0 : n - 1;
        case 3: {           // mirror: d c b | a b c d | c b a
            if (n == 1) return 0;
            const int p = 2 * n - 2;
            int m = ((j % p) + p) % p;
            return m < n ? m : p - m;
        }
        case 4: return ((j % n) + n) % n;
        default: return -1;
    }
}

__global__ void k_psf_data(const float2* __restrict__ E, int64_t n, int amplitude,
                           double* __restrict__ D) {
#pragma clang fp contract(off)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = hypotf(E[i].x, E[i].y);
    D[i] = amplitude ? static_cast<double>(a) : static_cast<double>(a * a);
}

__global__ void k_psf_conv(const double* __restrict__ D, const float2* __restrict__ E,
                           float2* __restrict__ out, int rows, int cols,
                           const double* __restrict__ K, int S, int mode_r, int mode_c,
                           int amplitude) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(rows) * cols) return;
    const int i = static_cast<int>(idx / cols), j = static_cast<int>(idx % cols);
    const int hw = S / 2;
    double s = 0.0;
    for (int a = 0; a < S; ++a) {
        const int ii = bmap(i + hw - a, rows, mode_r);
        if (ii < 0) continue;
        const double* Drow = D + static_cast<int64_t>(ii) * cols;
        const double* Krow = K + static_cast<int64_t>(a) * S;
        for (int b = 0; b < S; ++b) {
            const int jj = bmap(j + hw - b, cols, mode_c);
            if (jj < 0) continue;
            s += Krow[b] * Drow[jj];
        }
    }
    s = fmax(s, 0.0);
    const float amp = amplitude ? static_cast<float>(s) : static_cast<float>(sqrt(s));
    const float2 e = E[idx];
    const float ph = atan2f(e.y, e.x);
    float sn, cs;
    sincosf(ph, &sn, &cs);
    out[idx] = make_float2(amp * cs, amp * sn);
}

// hipFFT plans per (device, rows, cols); the caller's workspace is the work area
struct OptPlans {
    std::mutex mu;
    std::map<std::tuple<int, int, int>, std::pair<hipfftHandle, size_t>> plans;
};
OptPlans& opt_plans() {
    static OptPlans c;
    return c;
}

int opt_plan(int Ny, int Nx, hipfftHandle* h, size_t* work) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return MOSAIC_ERR_HIP;
    OptPlans& c = opt_plans();
    std::lock_guard<std::mutex> lk(c.mu);
    auto key = std::make_tuple(dev, Ny, Nx);
    auto it = c.plans.find(key);
    if (it == c.plans.end()) {
        hipfftHandle p;
        if (hipfftCreate(&p) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
        if (hipfftSetAutoAllocation(p, 0) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
        size_t ws = 0;
        if (hipfftMakePlan2d(p, Ny, Nx, HIPFFT_C2C, &ws) != HIPFFT_SUCCESS) {
            hipfftDestroy(p);
            return MOSAIC_ERR_HIP;
        }
        it = c.plans.emplace(key, std::make_pair(p, ws)).first;
    }
    *h = it->second.first;
    *work = it->second.second;
    return MOSAIC_OK;
}

}  // namespace
}  // namespace mosaic

using namespace mosaic;

extern "C" int mosaic_bragg_magnifier(const float* field_in, float* field_out, int Ny, int Nx,
                                      double Mx, double My, double reflectivity,
                                      double phase_shift, int order, int pad_edge,
                                      int conserve_energy, mosaic_stream_t stream) {
    if (!field_in || !field_out || Ny < 1 || Nx < 1 || Mx == 0.0 || My == 0.0)
        return MOSAIC_ERR_ARGS;
    if (field_in == field_out) return MOSAIC_ERR_ARGS;
    double amp = std::sqrt(std::max(reflectivity, 0.0));
    if (conserve_energy) {
        const double jac = std::fabs(Mx * My);
        if (jac > 0.0) amp *= 1.0 / std::sqrt(jac);
    }
    const float sr = static_cast<float>(amp * std::cos(phase_shift));
    const float si = static_cast<float>(amp * std::sin(phase_shift));
    hipLaunchKernelGGL(k_magnifier, dim3(blocks_for(static_cast<int64_t>(Ny) * Nx, OPT_THREADS)),
                       dim3(OPT_THREADS), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const float2*>(field_in),
                       reinterpret_cast<float2*>(field_out), Ny, Nx,
                       static_cast<float>((Nx - 1) / 2.0), static_cast<float>((Ny - 1) / 2.0),
                       static_cast<float>(Mx), static_cast<float>(My), order == 0 ? 0 : 1,
                       pad_edge ? 1 : 0, sr, si);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_angular_filter_workspace(int Ny, int Nx, size_t* bytes) {
    if (!bytes || Ny < 1 || Nx < 1) return MOSAIC_ERR_ARGS;
    hipfftHandle h;
    size_t work = 0;
    const int rc = opt_plan(Ny, Nx, &h, &work);
    if (rc != MOSAIC_OK) return rc;
    *bytes = work + 256;
    return MOSAIC_OK;
}

extern "C" int mosaic_angular_filter(float* field, int Ny, int Nx, double dx, double dy,
                                     double wavelength, const mosaic_angular_filter_t* f,
                                     void* workspace, size_t workspace_bytes,
                                     mosaic_stream_t stream) {
    if (!field || !f || Ny < 1 || Nx < 1 || !(dx > 0.0) || !(dy > 0.0) || !(wavelength > 0.0) ||
        f->rolloff < 0 || f->rolloff > 2)
        return MOSAIC_ERR_ARGS;
    hipfftHandle plan;
    size_t work = 0;
    int rc = opt_plan(Ny, Nx, &plan, &work);
    if (rc != MOSAIC_OK) return rc;
    if (work > 0 && (!workspace || workspace_bytes < work)) return MOSAIC_ERR_WORKSPACE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    FilterArgs A;
    A.Ny = Ny;
    A.Nx = Nx;
    A.k0 = static_cast<float>(2.0 * 3.141592653589793 / wavelength);
    A.cx = static_cast<float>(f->center_x_mrad * 1e-3);
    A.cy = static_cast<float>(f->center_y_mrad * 1e-3);
    A.hx = static_cast<float>(f->half_angle_x_mrad * 1e-3);
    A.hy = static_cast<float>(f->half_angle_y_mrad * 1e-3);
    const double phi = f->roll_deg * 3.141592653589793 / 180.0;
    A.cph = static_cast<float>(std::cos(phi));
    A.sph = static_cast<float>(std::sin(phi));
    A.mode_1d = f->mode_1d ? 1 : 0;
    A.elliptical = f->elliptical ? 1 : 0;
    A.rolloff = f->rolloff;
    A.order = std::max(1, f->order);
    A.dx = static_cast<float>(dx);
    A.dy = static_cast<float>(dy);
    A.scale = static_cast<float>(1.0 / (static_cast<double>(Ny) * Nx));
    if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
    if (work > 0 && hipfftSetWorkArea(plan, workspace) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
    hipfftComplex* F = reinterpret_cast<hipfftComplex*>(field);
    if (hipfftExecC2C(plan, F, F, HIPFFT_FORWARD) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
    const int64_t n = static_cast<int64_t>(Ny) * Nx;
    hipLaunchKernelGGL(k_angular_mask, dim3(blocks_for(n, OPT_THREADS)), dim3(OPT_THREADS), 0, st,
                       reinterpret_cast<float2*>(field), A, dx, dy);
    MOSAIC_HIP_CHECK(hipGetLastError());
    if (hipfftExecC2C(plan, F, F, HIPFFT_BACKWARD) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
    const double amp = std::sqrt(std::max(f->transmission, 0.0));
    hipLaunchKernelGGL(k_scale_c, dim3(blocks_for(n, OPT_THREADS)), dim3(OPT_THREADS), 0, st,
                       reinterpret_cast<float2*>(field), n,
                       static_cast<float>(amp * std::cos(f->phase_shift)),
                       static_cast<float>(amp * std::sin(f->phase_shift)));
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_zone_plate(float* field, int Ny, int Nx, double dx, double dy,
                                 double wavelength, const mosaic_zone_plate_t* z,
                                 mosaic_stream_t stream) {
    if (!field || !z || Ny < 1 || Nx < 1 || !(wavelength > 0.0) ||
        !(z->outermost_zone_width_nm > 0.0) || !(z->diameter_um > 0.0) || z->type < 0 ||
        z->type > 2)
        return MOSAIC_ERR_ARGS;
    const double drN = z->outermost_zone_width_nm * 1e-9, D = z->diameter_um * 1e-6;
    const int order = std::max(1, z->order);
    const double f1 = D * drN / wavelength, f_work = f1 / order, R = D / 2.0;
    ZoneArgs A;
    A.Ny = Ny;
    A.Nx = Nx;
    A.cx = static_cast<float>((Nx - 1) / 2.0);
    A.cy = static_cast<float>((Ny - 1) / 2.0);
    A.dx = static_cast<float>(dx);
    A.dy = static_cast<float>(dy);
    A.lf = static_cast<float>(wavelength * f1);
    A.R2max = static_cast<float>(R * R);
    A.stop = z->central_stop_fraction > 0.0 ? 1 : 0;
    const double Rs = R * z->central_stop_fraction;
    A.Rstop2 = static_cast<float>(Rs * Rs);
    A.R = static_cast<float>(R);
    A.taper = z->edge_taper > 0.0 ? 1 : 0;
    A.r_inner = static_cast<float>(R * (1.0 - z->edge_taper));
    A.taper_den = static_cast<float>(R * z->edge_taper);
    A.amp = static_cast<float>(std::sqrt(std::max(z->efficiency, 0.0)));
    A.half_kf = static_cast<float>(-0.5 * ((2.0 * 3.141592653589793 / wavelength) / f_work));
    A.type = z->type;
    hipLaunchKernelGGL(k_zone_plate, dim3(blocks_for(static_cast<int64_t>(Ny) * Nx, OPT_THREADS)),
                       dim3(OPT_THREADS), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<float2*>(field), A);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_aperture(float* field, int Ny, int Nx, double dx, double dy,
                               double width_mm, int circular, mosaic_stream_t stream) {
    if (!field || Ny < 1 || Nx < 1) return MOSAIC_ERR_ARGS;
    const double half = 0.5 * (width_mm * 1e-3);
    hipLaunchKernelGGL(k_aperture, dim3(blocks_for(static_cast<int64_t>(Ny) * Nx, OPT_THREADS)),
                       dim3(OPT_THREADS), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<float2*>(field), Ny, Nx,
                       static_cast<float>((Nx - 1) / 2.0), static_cast<float>((Ny - 1) / 2.0),
                       static_cast<float>(dx), static_cast<float>(dy), static_cast<float>(half),
                       static_cast<float>(half * half), circular ? 1 : 0);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_psf_workspace(int rows, int cols, int kernel_size, size_t* bytes) {
    if (!bytes || rows < 1 || cols < 1 || kernel_size < 1 || kernel_size % 2 == 0)
        return MOSAIC_ERR_ARGS;
    Carver c(nullptr);
    c.take<double>(static_cast<size_t>(rows) * cols);
    c.take<double>(static_cast<size_t>(kernel_size) * kernel_size);
    c.take<float2>(static_cast<size_t>(rows) * cols);
    *bytes = c.off + 256;
    return MOSAIC_OK;
}

extern "C" int mosaic_psf_apply(float* field, int rows, int cols, const double* kernel,
                                int kernel_size, int mode_rows, int mode_cols, int amplitude,
                                void* workspace, size_t workspace_bytes,
                                mosaic_stream_t stream) {
    size_t need = 0;
    int rc = mosaic_psf_workspace(rows, cols, kernel_size, &need);
    if (rc != MOSAIC_OK) return rc;
    if (!field || !kernel || mode_rows < 0 || mode_rows > 4 || mode_cols < 0 || mode_cols > 4)
        return MOSAIC_ERR_ARGS;
    if (!workspace || workspace_bytes < need) return MOSAIC_ERR_WORKSPACE;
    Carver c(workspace);
    const int64_t n = static_cast<int64_t>(rows) * cols;
    double* D = c.take<double>(n);
    double* K = c.take<double>(static_cast<size_t>(kernel_size) * kernel_size);
    float2* out = c.take<float2>(n);
    hipStream_t st = static_cast<hipStream_t>(stream);
    MOSAIC_HIP_CHECK(hipMemcpyAsync(K, kernel, sizeof(double) * kernel_size * kernel_size,
                                    hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_psf_data, dim3(blocks_for(n, OPT_THREADS)), dim3(OPT_THREADS), 0, st,
                       reinterpret_cast<const float2*>(field), n, amplitude ? 1 : 0, D);
    hipLaunchKernelGGL(k_psf_conv, dim3(blocks_for(n, OPT_THREADS)), dim3(OPT_THREADS), 0, st, D,
                       reinterpret_cast<const float2*>(field), out, rows, cols, K, kernel_size,
                       mode_rows, mode_cols, amplitude ? 1 : 0);
    MOSAIC_HIP_CHECK(hipGetLastError());
    MOSAIC_HIP_CHECK(hipMemcpyAsync(field, out, sizeof(float2) * n, hipMemcpyDeviceToDevice, st));
    // the host kernel table must stay valid until its copy has run
    MOSAIC_HIP_CHECK(hipStreamSynchronize(st));
    return MOSAIC_OK;
}
