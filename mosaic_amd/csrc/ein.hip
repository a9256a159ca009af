// Depth-dependent entrance field on gfx950: the reference's E0 * tau * phi
// sampler (ein_bilinear_kernel / _ein_bilinear_cpu, Beam.py:956-1045,
// 2439-2578) with the beam-column transmission map it samples
// (_compute_beam_column_A_map_cpu/_gpu, Beam.py:7118-7282) and the sample's
// depth window (_compute_global_depth_bounds, Beam.py:5583-5620).
//
// For a +x beam with transverse basis (e1, e2) = (y, z):
//  * depth window: s_min, s_max = min and max of the lab x of every atom;
//  * column map: every atom inside the beam grid deposits its forward
//    scattering factor (f0(0) + f', f'') onto the 3 x 3 grid nodes around its
//    (u, v) with triangular-shaped-cloud weights; the column sums give the
//    phase phi = -k C sum_re and attenuation tau = max(k C sum_im, 0),
//    C = r_e lambda^2 / (2 pi du dv), optionally blurred by a Gaussian of
//    radius R pixels (circular convolution with the kernel at the grid's
//    corner, as an FFT product with a zero-padded kernel forms it), and
//    A = exp(-tau + i phi); the sampler reads tau' = -ln(|A| + 1e-20) and
//    phi' = arg A (wrapped to (-pi, pi], as the reference re-derives them);
//  * entrance field per atom: E0, tau', phi' sampled bilinearly at (u, v),
//    f = clip((x - s_min) / (s_max - s_min), 0, 1), and
//    Ein = E0 exp(-f tau') exp(i f phi'); zero outside the grid.
// Float32 throughout in the reference's operation order; the deposit adds
// with atomics, so the column sums differ from a sequential sum by float32
// rounding (~1e-7 relative).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"

namespace mosaic {
namespace {

constexpr int EIN_THREADS = 256;
constexpr int EIN_MAX_SPECIES = 16;
constexpr int EIN_MAX_RADIUS = 16;

struct Pose {
    float R[9];
    float T[3];
};

struct EinLayout {
    float* sum_re;
    float* sum_im;
    float* tau;      // sampled maps tau', phi'
    float* phi;
    float* tmp0;     // blur scratch
    double* kern;    // blur kernel
    unsigned int* bounds;
    size_t total;
};

EinLayout ein_layout(void* base, int NyB, int NzB) {
    Carver c(base);
    const size_t g = static_cast<size_t>(NyB) * NzB;
    EinLayout L;
    L.sum_re = c.take<float>(g);
    L.sum_im = c.take<float>(g);
    L.tau = c.take<float>(g);
    L.phi = c.take<float>(g);
    L.tmp0 = c.take<float>(g);
    L.kern = c.take<double>((2 * EIN_MAX_RADIUS + 1) * (2 * EIN_MAX_RADIUS + 1));
    L.bounds = c.take<unsigned int>(2);
    L.total = c.off + 256;
    return L;
}

__device__ __forceinline__ unsigned int ord_f(float f) {
    const unsigned int b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// lab = R s + T in float32, the reference's (pos @ R.T) + T
__device__ __forceinline__ void lab_of(const float* __restrict__ pos, int64_t i, const Pose& P,
                                       float& lx, float& ly, float& lz) {
#pragma clang fp contract(off)
    const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    lx = ((x * P.R[0] + y * P.R[1]) + z * P.R[2]) + P.T[0];
    ly = ((x * P.R[3] + y * P.R[4]) + z * P.R[5]) + P.T[1];
    lz = ((x * P.R[6] + y * P.R[7]) + z * P.R[8]) + P.T[2];
}

__global__ void k_zero(float* __restrict__ a, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0f;
}

__global__ void k_bounds_init(unsigned int* b) {
    b[0] = 0xFFFFFFFFu;
    b[1] = 0u;
}

__global__ __launch_bounds__(EIN_THREADS) void k_depth_bounds(const float* __restrict__ pos,
                                                              int64_t n, Pose P,
                                                              unsigned int* __restrict__ b) {
    unsigned int lo = 0xFFFFFFFFu, hi = 0u;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float lx, ly, lz;
        lab_of(pos, i, P, lx, ly, lz);
        const unsigned int o = ord_f(lx);
        lo = min(lo, o);
        hi = max(hi, o);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, static_cast<unsigned int>(__shfl_xor(static_cast<int>(lo), o, 64)));
        hi = max(hi, static_cast<unsigned int>(__shfl_xor(static_cast<int>(hi), o, 64)));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&b[0], lo);
        atomicMax(&b[1], hi);
    }
}

struct DepositArgs {
    const float* pos;
    const uint8_t* species;
    int64_t n;
    Pose P;
    int NyB, NzB;
    float du, dv, uc, vc;
    float fr[EIN_MAX_SPECIES];   // f0(0) + f'
    float fi[EIN_MAX_SPECIES];   // f''
    float* sum_re;
    float* sum_im;
};

__device__ __forceinline__ float tsc_w(float d) {
    // 1-D triangular-shaped-cloud weight at distance d (pixels) from a node
    if (d <= 0.5f) return 0.75f - d * d;
    if (d <= 1.5f) {
        const float t = 1.5f - d;
        return 0.5f * t * t;
    }
    return 0.0f;
}

__global__ __launch_bounds__(EIN_THREADS) void k_tsc_deposit(const DepositArgs A) {
#pragma clang fp contract(off)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    float lx, ly, lz;
    lab_of(A.pos, i, A.P, lx, ly, lz);
    const float iu = ly / A.du + A.uc;
    const float iv = lz / A.dv + A.vc;
    if (!(iu >= 0.f && iu <= static_cast<float>(A.NyB - 1) && iv >= 0.f &&
          iv <= static_cast<float>(A.NzB - 1)))
        return;
    const int sp = A.species[i];
    const float fr = A.fr[sp], fi = A.fi[sp];
    const int ic = static_cast<int>(floorf(iu + 0.5f));
    const int jc = static_cast<int>(floorf(iv + 0.5f));
    float wu[3], wv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        wu[d] = tsc_w(fabsf(iu - static_cast<float>(ic - 1 + d)));
        wv[d] = tsc_w(fabsf(iv - static_cast<float>(jc - 1 + d)));
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int ii = ic - 1 + a;
        if (ii < 0 || ii >= A.NyB) continue;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int jj = jc - 1 + b;
            const float fac = wu[a] * wv[b];
            if (jj < 0 || jj >= A.NzB || !(fac > 0.f)) continue;
            const int64_t idx = static_cast<int64_t>(ii) * A.NzB + jj;
            atomicAdd(&A.sum_re[idx], fr * fac);
            atomicAdd(&A.sum_im[idx], fi * fac);
        }
    }
}

// phi = -k C sum_re, tau = max(k C sum_im, 0) in float32
__global__ void k_column_phase(const float* __restrict__ sre, const float* __restrict__ sim,
                               int64_t g, float negkC_scale, float C, float kA,
                               float* __restrict__ phi, float* __restrict__ tau) {
#pragma clang fp contract(off)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= g) return;
    phi[i] = negkC_scale * (C * sre[i]);
    tau[i] = fmaxf(kA * (C * sim[i]), 0.0f);
}

// circular convolution with a (2R+1)^2 Gaussian placed at the grid's corner:
// out[i][j] = sum_ab k[a][b] in[(i - a) mod NyB][(j - b) mod NzB], in FP64
__global__ void k_blur(const float* __restrict__ in, float* __restrict__ out, int NyB, int NzB,
                       int R, const double* __restrict__ kern, int clamp0) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(NyB) * NzB) return;
    const int i = static_cast<int>(idx / NzB), j = static_cast<int>(idx % NzB);
    const int w = 2 * R + 1;
    double s = 0.0;
    for (int a = 0; a < w; ++a) {
        const int ii = ((i - a) % NyB + NyB) % NyB;
        for (int b = 0; b < w; ++b) {
            const int jj = ((j - b) % NzB + NzB) % NzB;
            s += kern[a * w + b] * static_cast<double>(in[static_cast<int64_t>(ii) * NzB + jj]);
        }
    }
    float v = static_cast<float>(s);
    if (clamp0) v = fmaxf(v, 0.0f);
    out[idx] = v;
}

// A = exp(-tau + i phi) (complex64), then the sampled maps tau' = -ln(|A| +
// 1e-20) and phi' = arg A
__global__ void k_column_maps(float* __restrict__ tau, float* __restrict__ phi, int64_t g,
                              float2* __restrict__ A_out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= g) return;
    const float m = expf(-tau[i]);
    float sn, cs;
    sincosf(phi[i], &sn, &cs);
    const float re = m * cs, im = m * sn;
    if (A_out) A_out[i] = make_float2(re, im);
    tau[i] = -logf(hypotf(re, im) + 1e-20f);
    phi[i] = atan2f(im, re);
}

struct EinDepthArgs {
    const float* pos;
    int64_t n;
    Pose P;
    const float2* E0;
    const float* tau;
    const float* phi;
    int NyB, NzB;
    float du, dv, uc, vc;
    float s_min, denom;
    float2* out;
};

__global__ __launch_bounds__(EIN_THREADS) void k_ein_depth(const EinDepthArgs A) {
#pragma clang fp contract(off)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    float lx, ly, lz;
    lab_of(A.pos, i, A.P, lx, ly, lz);
    const float iu = ly / A.du + A.uc;
    const float iv = lz / A.dv + A.vc;
    float2 res = make_float2(0.f, 0.f);
    if (iu >= 0.f && iu <= static_cast<float>(A.NyB - 1) && iv >= 0.f &&
        iv <= static_cast<float>(A.NzB - 1)) {
        const int i0 = static_cast<int>(floorf(iu));
        const int j0 = static_cast<int>(floorf(iv));
        const int i1 = min(i0 + 1, A.NyB - 1);
        const int j1 = min(j0 + 1, A.NzB - 1);
        const float fu = iu - static_cast<float>(i0);
        const float fv = iv - static_cast<float>(j0);
        const float w00 = (1.f - fu) * (1.f - fv);
        const float w01 = (1.f - fu) * fv;
        const float w10 = fu * (1.f - fv);
        const float w11 = fu * fv;
        const int r00 = i0 * A.NzB + j0, r01 = i0 * A.NzB + j1;
        const int r10 = i1 * A.NzB + j0, r11 = i1 * A.NzB + j1;
        const float ts = ((A.tau[r00] * w00 + A.tau[r01] * w01) + A.tau[r10] * w10) +
                         A.tau[r11] * w11;
        const float ps = ((A.phi[r00] * w00 + A.phi[r01] * w01) + A.phi[r10] * w10) +
                         A.phi[r11] * w11;
        const float2 a = A.E0[r00], b = A.E0[r01], c = A.E0[r10], d = A.E0[r11];
        const float er = ((a.x * w00 + b.x * w01) + c.x * w10) + d.x * w11;
        const float ei = ((a.y * w00 + b.y * w01) + c.y * w10) + d.y * w11;
        const float f = fminf(fmaxf((lx - A.s_min) / A.denom, 0.f), 1.f);
        const float amp = expf(-f * ts);
        float sn, cs;
        sincosf(f * ps, &sn, &cs);
        res.x = (er * cs - ei * sn) * amp;
        res.y = (er * sn + ei * cs) * amp;
    }
    A.out[i] = res;
}

// M-channel incident amplitude (the general kernel's lattice-phase sum,
// Beam.py:2283-2309): amp_j = sum_m E[j][m] e^{2 pi i g_m . r_j}, r_j the lab
// position R s_j + T in metres (FP64), the phase reduced modulo one cycle in
// FP64 before the float32 sine and cosine.
constexpr int MBEAM_MAX = 64;
struct MbeamArgs {
    double RT[12];              // R row-major, then T [A]
    double g[3 * MBEAM_MAX];    // reciprocal vectors [cycles per metre]
    int M;
};

__global__ __launch_bounds__(EIN_THREADS) void k_mbeam(const float* __restrict__ pos, int64_t n,
                                                       const float2* __restrict__ E,
                                                       const MbeamArgs A,
                                                       float2* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const double lx = (fma(A.RT[0], x, fma(A.RT[1], y, A.RT[2] * z)) + A.RT[9]) * 1e-10;
    const double ly = (fma(A.RT[3], x, fma(A.RT[4], y, A.RT[5] * z)) + A.RT[10]) * 1e-10;
    const double lz = (fma(A.RT[6], x, fma(A.RT[7], y, A.RT[8] * z)) + A.RT[11]) * 1e-10;
    double ar = 0.0, ai = 0.0;
    for (int m = 0; m < A.M; ++m) {
        double t = fma(A.g[3 * m], lx, fma(A.g[3 * m + 1], ly, A.g[3 * m + 2] * lz));
        t -= rint(t);
        const float tf = static_cast<float>(t);
        const float cs = __builtin_amdgcn_cosf(tf), sn = __builtin_amdgcn_sinf(tf);
        const float2 e = E[i * A.M + m];
        ar += static_cast<double>(e.x * cs - e.y * sn);
        ai += static_cast<double>(e.x * sn + e.y * cs);
    }
    out[i] = make_float2(static_cast<float>(ar), static_cast<float>(ai));
}

Pose pose_of(const float* rot, const float* trans) {
    Pose P;
    for (int q = 0; q < 9; ++q) P.R[q] = rot[q];
    for (int d = 0; d < 3; ++d) P.T[d] = trans[d];
    return P;
}

}  // namespace
}  // namespace mosaic

using namespace mosaic;

extern "C" int mosaic_ein_depth_workspace(int NyB, int NzB, size_t* bytes) {
    if (!bytes || NyB < 1 || NzB < 1) return MOSAIC_ERR_ARGS;
    *bytes = ein_layout(nullptr, NyB, NzB).total;
    return MOSAIC_OK;
}

extern "C" int mosaic_ein_depth_bounds(int64_t atom_count, const float* pos, const float* rot,
                                       const float* trans, int NyB, int NzB, double* bounds,
                                       void* workspace, size_t workspace_bytes,
                                       mosaic_stream_t stream) {
    if (atom_count < 0 || !rot || !trans || !bounds || NyB < 1 || NzB < 1) return MOSAIC_ERR_ARGS;
    bounds[0] = 0.0;
    bounds[1] = 1.0;
    if (atom_count == 0) return MOSAIC_OK;
    if (!pos) return MOSAIC_ERR_ARGS;
    const EinLayout L = ein_layout(workspace, NyB, NzB);
    if (!workspace || workspace_bytes < L.total) return MOSAIC_ERR_WORKSPACE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(1), 0, st, L.bounds);
    const int blocks = static_cast<int>(std::min<int64_t>(2048, (atom_count + EIN_THREADS - 1) / EIN_THREADS));
    hipLaunchKernelGGL(k_depth_bounds, dim3(blocks), dim3(EIN_THREADS), 0, st, pos, atom_count,
                       pose_of(rot, trans), L.bounds);
    MOSAIC_HIP_CHECK(hipGetLastError());
    unsigned int hb[2];
    MOSAIC_HIP_CHECK(hipMemcpyAsync(hb, L.bounds, sizeof(hb), hipMemcpyDeviceToHost, st));
    MOSAIC_HIP_CHECK(hipStreamSynchronize(st));
    // decode on the host: the order-preserving encoding of a float
    auto dec = [](unsigned int u) {
        const unsigned int b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
        float f;
        std::memcpy(&f, &b, sizeof(f));
        return f;
    };
    const float lo = dec(hb[0]), hi = dec(hb[1]);
    if (std::isfinite(lo) && std::isfinite(hi) && hi > lo) {
        bounds[0] = lo;
        bounds[1] = hi;
    }
    return MOSAIC_OK;
}

extern "C" int mosaic_ein_column_map(int64_t atom_count, const float* pos, const uint8_t* species,
                                     int n_species, const float* fwd, const float* rot,
                                     const float* trans, int NyB, int NzB, double du, double dv,
                                     double uc, double vc, double wavelength_A, int kernel_radius,
                                     float* A_out, void* workspace, size_t workspace_bytes,
                                     mosaic_stream_t stream) {
    if (atom_count < 0 || !rot || !trans || !fwd || NyB < 1 || NzB < 1 || n_species < 1 ||
        n_species > EIN_MAX_SPECIES || !(du > 0.0) || !(dv > 0.0) || !(wavelength_A > 0.0) ||
        kernel_radius < 0 || kernel_radius > EIN_MAX_RADIUS)
        return MOSAIC_ERR_ARGS;
    if (atom_count > 0 && (!pos || !species)) return MOSAIC_ERR_ARGS;
    const EinLayout L = ein_layout(workspace, NyB, NzB);
    if (!workspace || workspace_bytes < L.total) return MOSAIC_ERR_WORKSPACE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t g = static_cast<int64_t>(NyB) * NzB;
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(g, 256)), dim3(256), 0, st, L.sum_re, g);
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(g, 256)), dim3(256), 0, st, L.sum_im, g);
    if (atom_count > 0) {
        DepositArgs D;
        D.pos = pos;
        D.species = species;
        D.n = atom_count;
        D.P = pose_of(rot, trans);
        D.NyB = NyB;
        D.NzB = NzB;
        D.du = static_cast<float>(du);
        D.dv = static_cast<float>(dv);
        D.uc = static_cast<float>(uc);
        D.vc = static_cast<float>(vc);
        for (int s = 0; s < EIN_MAX_SPECIES; ++s) {
            D.fr[s] = s < n_species ? fwd[2 * s] : 0.f;
            D.fi[s] = s < n_species ? fwd[2 * s + 1] : 0.f;
        }
        D.sum_re = L.sum_re;
        D.sum_im = L.sum_im;
        hipLaunchKernelGGL(k_tsc_deposit, dim3(blocks_for(atom_count, EIN_THREADS)),
                           dim3(EIN_THREADS), 0, st, D);
        MOSAIC_HIP_CHECK(hipGetLastError());
    }
    // C = r_e lambda^2 / (2 pi du dv) (dimensionless), k = 2 pi / lambda, in Angstrom
    const double r_e_A = 2.81794092e-5, two_pi = 6.283185307179586;
    const double C = (r_e_A * (wavelength_A * wavelength_A)) / (two_pi * (du * dv));
    const double kA = two_pi / wavelength_A;
    const int gb = blocks_for(g, 256);
    hipLaunchKernelGGL(k_column_phase, dim3(gb), dim3(256), 0, st, L.sum_re, L.sum_im, g,
                       static_cast<float>(-kA), static_cast<float>(C), static_cast<float>(kA),
                       L.phi, L.tau);
    MOSAIC_HIP_CHECK(hipGetLastError());
    if (kernel_radius > 0) {
        const int R = kernel_radius, w = 2 * R + 1;
        const double sig = R / 2.0;
        double kern[(2 * EIN_MAX_RADIUS + 1) * (2 * EIN_MAX_RADIUS + 1)];
        double tot = 0.0;
        for (int a = 0; a < w; ++a)
            for (int b = 0; b < w; ++b) {
                const double y = a - R, x = b - R;
                const double v = static_cast<double>(
                    static_cast<float>(std::exp(-(x * x + y * y) / (2.0 * sig * sig))));
                kern[a * w + b] = v;
                tot += v;
            }
        for (int q = 0; q < w * w; ++q) kern[q] /= tot;
        double* dk = L.kern;
        MOSAIC_HIP_CHECK(hipMemcpyAsync(dk, kern, sizeof(double) * w * w, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_blur, dim3(gb), dim3(256), 0, st, L.phi, L.tmp0, NyB, NzB, R, dk, 0);
        MOSAIC_HIP_CHECK(hipMemcpyAsync(L.phi, L.tmp0, sizeof(float) * g, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_blur, dim3(gb), dim3(256), 0, st, L.tau, L.tmp0, NyB, NzB, R, dk, 1);
        MOSAIC_HIP_CHECK(hipMemcpyAsync(L.tau, L.tmp0, sizeof(float) * g, hipMemcpyDeviceToDevice, st));
        // the host kernel table must outlive the copy
        MOSAIC_HIP_CHECK(hipStreamSynchronize(st));
    }
    hipLaunchKernelGGL(k_column_maps, dim3(gb), dim3(256), 0, st, L.tau, L.phi, g,
                       reinterpret_cast<float2*>(A_out));
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_ein_depth(int64_t atom_count, const float* pos, const float* rot,
                                const float* trans, const float* E0, int NyB, int NzB, double du,
                                double dv, double uc, double vc, double s_min, double s_max,
                                float* amp_out, const void* workspace, size_t workspace_bytes,
                                mosaic_stream_t stream) {
    if (atom_count < 0 || !rot || !trans || NyB < 1 || NzB < 1 || !(du > 0.0) || !(dv > 0.0))
        return MOSAIC_ERR_ARGS;
    if (atom_count == 0) return MOSAIC_OK;
    if (!pos || !E0 || !amp_out) return MOSAIC_ERR_ARGS;
    const EinLayout L = ein_layout(const_cast<void*>(workspace), NyB, NzB);
    if (!workspace || workspace_bytes < L.total) return MOSAIC_ERR_WORKSPACE;
    EinDepthArgs A;
    A.pos = pos;
    A.n = atom_count;
    A.P = pose_of(rot, trans);
    A.E0 = reinterpret_cast<const float2*>(E0);
    A.tau = L.tau;
    A.phi = L.phi;
    A.NyB = NyB;
    A.NzB = NzB;
    A.du = static_cast<float>(du);
    A.dv = static_cast<float>(dv);
    A.uc = static_cast<float>(uc);
    A.vc = static_cast<float>(vc);
    A.s_min = static_cast<float>(s_min);
    const double den = s_max - s_min;
    A.denom = static_cast<float>(den > 0.0 ? den : 1.0);
    A.out = reinterpret_cast<float2*>(amp_out);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_ein_depth, dim3(blocks_for(atom_count, EIN_THREADS)), dim3(EIN_THREADS),
                       0, st, A);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_mbeam_amplitude(int64_t atom_count, const float* pos, const float* rot,
                                      const float* trans, const float* E, int M,
                                      const double* g, float* amp_out, mosaic_stream_t stream) {
    if (atom_count < 0 || !rot || !trans || M < 1 || M > MBEAM_MAX || !g) return MOSAIC_ERR_ARGS;
    if (atom_count == 0) return MOSAIC_OK;
    if (!pos || !E || !amp_out) return MOSAIC_ERR_ARGS;
    MbeamArgs A;
    for (int q = 0; q < 9; ++q) A.RT[q] = rot[q];
    for (int d = 0; d < 3; ++d) A.RT[9 + d] = trans[d];
    for (int q = 0; q < 3 * MBEAM_MAX; ++q) A.g[q] = q < 3 * M ? g[q] : 0.0;
    A.M = M;
    hipLaunchKernelGGL(k_mbeam, dim3(blocks_for(atom_count, EIN_THREADS)), dim3(EIN_THREADS), 0,
                       static_cast<hipStream_t>(stream), pos, atom_count,
                       reinterpret_cast<const float2*>(E), A, reinterpret_cast<float2*>(amp_out));
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}
