// Angular-spectrum free-space propagation and the thin lens on gfx950.
//
// One propagation step (Beam.py:7829-7943, 7945-8022): pad the (Nz, Ny)
// field to (Nz2, Ny2) -- edge replication or a constant -- FFT, multiply by
//     H = exp(-i z kt^2 / (sqrt(k^2 - kt^2) + k))        (propagating)
//     H = exp(-|z| sqrt(kt^2 - k^2)) exp(-i z k)         (evanescent)
// (prop_mul_kernel / prop_mul_cpu, Beam.py:4739-4900), inverse FFT, crop the
// centre back to (Nz, Ny).  The thin lens exp(-i k r^2 / 2f) with the CRL
// absorption factors (Optics.py:49-130) is fused into the pad kernel when it
// precedes a free-space step, so the lensed field is never written at its
// own size.
//
// The FFTs are hipFFT plans (cached per size and device, work area supplied by
// the caller).  The pad, transfer-function and crop passes are hand-written:
//   k_pad_lens -- reads the small field once, writes the padded grid;
//   k_tf       -- one read-modify-write of the spectrum; the transfer
//                 function phase is formed in FP64 and reduced mod 2 pi
//                 before the FP32 sincos;
//   k_crop     -- reads only the crop window, folds in the 1/(Ny2 Nz2) of the
//                 inverse transform.

#include <hipfft/hipfft.h>

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

#include "common.h"

namespace mosaic {
namespace {

constexpr double TWO_PI = 6.283185307179586476925286766559;

struct LensArgs {
    int active;
    double half_k_over_f;   // 0.5 k / f   [rad / m^2]
    double parab_coef;      // mu N / (2 R_apex)  [1/m^2], 0 when off
    float uniform_amp;      // exp(-N t / sigma), 1 when off
};

__global__ void k_pad_lens(const float2* __restrict__ in, int Ny, int Nz, int Ny2, int Nz2,
                           int y0, int z0, int edge, float2 pad_value, double dx, double dy,
                           LensArgs lens, float2* __restrict__ out) {
    const int iy2 = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz2 = blockIdx.y;
    if (iy2 >= Ny2 || iz2 >= Nz2) return;
    int iy = iy2 - y0, iz = iz2 - z0;
    float2 v;
    const bool inside = iy >= 0 && iy < Ny && iz >= 0 && iz < Nz;
    if (!inside && !edge) {
        v = pad_value;
    } else {
        iy = min(max(iy, 0), Ny - 1);
        iz = min(max(iz, 0), Nz - 1);
        v = in[static_cast<int64_t>(iz) * Ny + iy];
        if (lens.active) {
            // centred coordinates of the source pixel (x along columns)
            const double x = (iy - 0.5 * (Ny - 1)) * dx;
            const double y = (iz - 0.5 * (Nz - 1)) * dy;
            const double r2 = x * x + y * y;
            double ph = -lens.half_k_over_f * r2;
            ph -= TWO_PI * rint(ph / TWO_PI);
            float s, c;
            sincosf(static_cast<float>(ph), &s, &c);
            float a = lens.uniform_amp;
            if (lens.parab_coef > 0.0) a *= static_cast<float>(exp(-lens.parab_coef * r2));
            v = make_float2(a * (v.x * c - v.y * s), a * (v.x * s + v.y * c));
        }
    }
    out[static_cast<int64_t>(iz2) * Ny2 + iy2] = v;
}

// Standalone thin lens in place (no following free-space step).
__global__ void k_lens(float2* __restrict__ E, int Ny, int Nz, double dx, double dy,
                       LensArgs lens) {
    const int iy = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (iy >= Ny || iz >= Nz) return;
    const double x = (iy - 0.5 * (Ny - 1)) * dx;
    const double y = (iz - 0.5 * (Nz - 1)) * dy;
    const double r2 = x * x + y * y;
    double ph = -lens.half_k_over_f * r2;
    ph -= TWO_PI * rint(ph / TWO_PI);
    float s, c;
    sincosf(static_cast<float>(ph), &s, &c);
    float a = lens.uniform_amp;
    if (lens.parab_coef > 0.0) a *= static_cast<float>(exp(-lens.parab_coef * r2));
    const int64_t i = static_cast<int64_t>(iz) * Ny + iy;
    const float2 v = E[i];
    E[i] = make_float2(a * (v.x * c - v.y * s), a * (v.x * s + v.y * c));
}

// Spatial frequency of FFT bin i of n at pitch d, in rad/m (numpy fftfreq order).
__device__ __forceinline__ double kfreq(int i, int n, double d) {
    const int m = (i < (n + 1) / 2) ? i : i - n;
    return TWO_PI * static_cast<double>(m) / (static_cast<double>(n) * d);
}

__global__ void k_tf(float2* __restrict__ F, int Ny2, int Nz2, double dy, double dz, double k,
                     double z) {
    const int iy = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (iy >= Ny2 || iz >= Nz2) return;
    const double ky = kfreq(iy, Ny2, dy), kz = kfreq(iz, Nz2, dz);
    const double kt2 = ky * ky + kz * kz;
    const double kx2 = k * k - kt2;
    double ph, amp;
    if (kx2 >= 0.0) {
        ph = -z * kt2 / (sqrt(kx2) + k);
        amp = 1.0;
    } else {
        ph = -z * k;
        amp = exp(-fabs(z) * sqrt(-kx2));
    }
    ph -= TWO_PI * rint(ph / TWO_PI);
    float s, c;
    sincosf(static_cast<float>(ph), &s, &c);
    const float a = static_cast<float>(amp);
    const int64_t i = static_cast<int64_t>(iz) * Ny2 + iy;
    const float2 g = F[i];
    F[i] = make_float2(a * (g.x * c - g.y * s), a * (g.x * s + g.y * c));
}

__global__ void k_crop(const float2* __restrict__ F, int Ny2, int y0, int z0, int Ny, int Nz,
                       float scale, float2* __restrict__ out) {
    const int iy = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (iy >= Ny || iz >= Nz) return;
    const float2 v = F[static_cast<int64_t>(iz + z0) * Ny2 + (iy + y0)];
    out[static_cast<int64_t>(iz) * Ny + iy] = make_float2(v.x * scale, v.y * scale);
}

// hipFFT plans cached per (device, Nz2, Ny2); the caller's workspace is the
// work area, so no plan allocates.
struct PlanCache {
    std::mutex mu;
    std::map<std::tuple<int, int, int>, std::pair<hipfftHandle, size_t>> plans;
};
PlanCache& cache() {
    static PlanCache c;
    return c;
}

int get_plan(int Nz2, int Ny2, hipfftHandle* h, size_t* work) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return MOSAIC_ERR_HIP;
    PlanCache& c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto key = std::make_tuple(dev, Nz2, Ny2);
    auto it = c.plans.find(key);
    if (it == c.plans.end()) {
        hipfftHandle p;
        if (hipfftCreate(&p) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
        if (hipfftSetAutoAllocation(p, 0) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
        size_t ws = 0;
        if (hipfftMakePlan2d(p, Nz2, Ny2, HIPFFT_C2C, &ws) != HIPFFT_SUCCESS) {
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

extern "C" int mosaic_asp_workspace(int Ny2, int Nz2, size_t* bytes) {
    if (Ny2 < 1 || Nz2 < 1 || !bytes) return MOSAIC_ERR_ARGS;
    hipfftHandle h;
    size_t work = 0;
    int rc = get_plan(Nz2, Ny2, &h, &work);
    if (rc != MOSAIC_OK) return rc;
    Carver c(nullptr);
    c.take<float2>(static_cast<size_t>(Ny2) * Nz2);
    c.take<char>(work);
    *bytes = c.off + 256;
    return MOSAIC_OK;
}

extern "C" int mosaic_asp_propagate(const float* field_in, int Ny, int Nz, double dy, double dz,
                                    double wavelength, double z, int Ny2, int Nz2, int pad_edge,
                                    float pad_re, float pad_im, const mosaic_lens_t* lens,
                                    float* field_out, void* workspace, size_t workspace_bytes,
                                    mosaic_stream_t stream) {
    if (!field_in || !field_out || Ny < 1 || Nz < 1 || Ny2 < Ny || Nz2 < Nz ||
        !(dy > 0.0) || !(dz > 0.0) || !(wavelength > 0.0) || !workspace)
        return MOSAIC_ERR_ARGS;
    hipfftHandle plan;
    size_t work = 0;
    int rc = get_plan(Nz2, Ny2, &plan, &work);
    if (rc != MOSAIC_OK) return rc;
    Carver c(workspace);
    float2* F = c.take<float2>(static_cast<size_t>(Ny2) * Nz2);
    void* wa = c.take<char>(work);
    if (c.off > workspace_bytes) return MOSAIC_ERR_WORKSPACE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int y0 = (Ny2 - Ny) / 2, z0 = (Nz2 - Nz) / 2;
    const double k = TWO_PI / wavelength;
    LensArgs L{0, 0.0, 0.0, 1.0f};
    if (lens) {
        if (!(lens->focal_length_m != 0.0)) return MOSAIC_ERR_ARGS;
        L.active = 1;
        L.half_k_over_f = 0.5 * k / lens->focal_length_m;
        L.parab_coef = lens->parab_coef;
        L.uniform_amp = lens->uniform_amp;
    }
    dim3 blk(256);
    hipLaunchKernelGGL(k_pad_lens, dim3(blocks_for(Ny2, 256), Nz2), blk, 0, st,
                       reinterpret_cast<const float2*>(field_in), Ny, Nz, Ny2, Nz2, y0, z0,
                       pad_edge ? 1 : 0, make_float2(pad_re, pad_im), dy, dz, L, F);
    MOSAIC_HIP_CHECK(hipGetLastError());
    {
        std::lock_guard<std::mutex> lk(cache().mu);
        if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
        if (hipfftSetWorkArea(plan, wa) != HIPFFT_SUCCESS) return MOSAIC_ERR_HIP;
        if (hipfftExecC2C(plan, reinterpret_cast<hipfftComplex*>(F),
                          reinterpret_cast<hipfftComplex*>(F), HIPFFT_FORWARD) != HIPFFT_SUCCESS)
            return MOSAIC_ERR_HIP;
        hipLaunchKernelGGL(k_tf, dim3(blocks_for(Ny2, 256), Nz2), blk, 0, st, F, Ny2, Nz2, dy, dz,
                           k, z);
        if (hipGetLastError() != hipSuccess) return MOSAIC_ERR_HIP;
        if (hipfftExecC2C(plan, reinterpret_cast<hipfftComplex*>(F),
                          reinterpret_cast<hipfftComplex*>(F), HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
            return MOSAIC_ERR_HIP;
    }
    const float scale = static_cast<float>(1.0 / (static_cast<double>(Ny2) * Nz2));
    hipLaunchKernelGGL(k_crop, dim3(blocks_for(Ny, 256), Nz), blk, 0, st, F, Ny2, y0, z0, Ny, Nz,
                       scale, reinterpret_cast<float2*>(field_out));
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_lens_apply(float* field, int Ny, int Nz, double dx, double dy,
                                 double wavelength, const mosaic_lens_t* lens,
                                 mosaic_stream_t stream) {
    if (!field || Ny < 1 || Nz < 1 || !lens || !(wavelength > 0.0) ||
        !(lens->focal_length_m != 0.0))
        return MOSAIC_ERR_ARGS;
    const double k = TWO_PI / wavelength;
    LensArgs L{1, 0.5 * k / lens->focal_length_m, lens->parab_coef, lens->uniform_amp};
    hipLaunchKernelGGL(k_lens, dim3(blocks_for(Ny, 256), Nz), dim3(256), 0,
                       static_cast<hipStream_t>(stream), reinterpret_cast<float2*>(field), Ny,
                       Nz, dx, dy, L);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}
