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
// Two implementations: the pruned three-pass transform below (the default
// where its size rules hold: power-of-two padded sizes, data lengths up to
// 8192), and the padded 2-D hipFFT round trip (plans cached per size and
// device, work area supplied by the caller) with hand-written passes:
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

// ---------------------------------------------------------------------------
// Pruned transforms.  Along one axis the padded line of length M (a power of
// two) holds the data x_n (n < N) at offset o = (M - N) / 2 and, outside it,
// the edge values a (left) and b (right) -- x_0 and x_{N-1} in edge mode, the
// pad constant otherwise.  With L = 2^ceil(log2 N) (at least 64), R = M / L
// and w = e^{-2 pi i / M}, the forward transform splits as
//     X_{Rm+r} = w^{o (Rm+r)} FFT_L(x_n w^{n r})[m] + E(Rm+r),
//     E(k) = [(a - b) - a w^{o k} + b w^{(o+N) k}] / (1 - w^k)   (k != 0),
//     E(0) = a o + b (M - o - N)
// (the two constant runs are geometric series), and the inverse transform,
// needed only on the crop window n < N, splits as
//     y_n = sum_r w^{-(o+n) r} IFFT_L(X_{R.+r})[(o + n) mod L].
// So a step never forms the padded grid: one pass transforms the N_z data
// rows along y, one pass per y-frequency column transforms along z, applies
// the transfer function and inverts back to the N_z crop rows, and one pass
// inverts the crop rows along y.  HBM traffic per step is 16 N_y N_z +
// 32 M_y N_z bytes (the intermediate is M_y x N_z, column-major), against
// 64 M_y M_z for a padded 2-D FFT round trip.
//
// Each line's length-L transforms run in LDS: Stockham autosort, radix 8
// (then one radix-4 or radix-2 stage), natural order in and out, twiddles
// from an LDS table e^{-2 pi i t / L}; one float2 of padding per eight
// spreads the stride-8 stores over the banks.

constexpr int FFT_MIN = 64;
constexpr int FFT_MAX = 8192;

__host__ __device__ constexpr int fft_threads(int L) { return L / 8 < 64 ? 64 : L / 8; }
__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

__device__ __forceinline__ int lp(int e) { return e + (e >> 3); }

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// w^e = e^{-2 pi i e / M} for a power-of-two M, the exponent reduced exactly
__device__ __forceinline__ float2 wpow(long long e, int M) {
    const int q = static_cast<int>(e & (M - 1));
    float s, c;
    sincospif(-2.0f * static_cast<float>(q) / static_cast<float>(M), &s, &c);
    return make_float2(c, s);
}

template <bool INV>
__device__ __forceinline__ void fft4(float2& v0, float2& v1, float2& v2, float2& v3) {
    const float2 c0 = cadd(v0, v2), c1 = csub(v0, v2), c2 = cadd(v1, v3),
                 c3 = mul_mi<INV>(csub(v1, v3));
    v0 = cadd(c0, c2);
    v2 = csub(c0, c2);
    v1 = cadd(c1, c3);
    v3 = csub(c1, c3);
}

template <bool INV>
__device__ __forceinline__ void fft8(float2* v) {
    constexpr float H = 0.70710678118654752f;
    float2 a[4], b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        a[n] = cadd(v[n], v[n + 4]);
        b[n] = csub(v[n], v[n + 4]);
    }
    // b_n w8^n
    b[1] = INV ? make_float2(H * (b[1].x - b[1].y), H * (b[1].x + b[1].y))
               : make_float2(H * (b[1].x + b[1].y), H * (b[1].y - b[1].x));
    b[2] = mul_mi<INV>(b[2]);
    b[3] = INV ? make_float2(-H * (b[3].x + b[3].y), H * (b[3].x - b[3].y))
               : make_float2(H * (b[3].y - b[3].x), -H * (b[3].x + b[3].y));
    fft4<INV>(a[0], a[1], a[2], a[3]);
    fft4<INV>(b[0], b[1], b[2], b[3]);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        v[2 * m] = a[m];
        v[2 * m + 1] = b[m];
    }
}

// One Stockham stage of radix RD on butterfly j (j < L / RD), stride Ns.
// Twiddles of a stage (stride Ns, radix RD) sit at tw[off + (r - 1) Ns + k],
// k = j mod Ns: consecutive lanes read consecutive entries.
template <int L, int RD, bool INV>
__device__ __forceinline__ void stage_load(const float2* buf, const float2* tw, int off, int Ns,
                                           int j, float2* v) {
#pragma unroll
    for (int r = 0; r < RD; ++r) v[r] = buf[lp(j + r * (L / RD))];
    const int k = j & (Ns - 1);
    if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < RD; ++r) {
            const float2 w = tw[off + (r - 1) * Ns + k];
            v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
        }
    }
    if (RD == 8) {
        fft8<INV>(v);
    } else if (RD == 4) {
        fft4<INV>(v[0], v[1], v[2], v[3]);
    } else {
        const float2 t = csub(v[0], v[1]);
        v[0] = cadd(v[0], v[1]);
        v[1] = t;
    }
}

template <int L, int RD>
__device__ __forceinline__ void stage_store(float2* buf, int Ns, int j, const float2* v) {
    const int k = j & (Ns - 1);
    const int d = (j / Ns) * Ns * RD + k;
#pragma unroll
    for (int r = 0; r < RD; ++r) buf[lp(d + r * Ns)] = v[r];
}

// In-place length-L transform of buf (natural order), all threads of the
// block taking part; begins and ends with the block synchronised.
template <int L, bool INV>
__device__ void fft_lds(float2* buf, const float2* tw) {
    constexpr int T = fft_threads(L);
    constexpr int LG = ilog2c(L);
    const int t = threadIdx.x;
    int Ns = 1, off = 0;
#pragma unroll
    for (int s = 0; s < LG / 3; ++s) {
        float2 v[8];
        const bool act = t < L / 8;
        if (act) stage_load<L, 8, INV>(buf, tw, off, Ns, t, v);
        __syncthreads();
        if (act) stage_store<L, 8>(buf, Ns, t, v);
        __syncthreads();
        off += 7 * Ns;
        Ns *= 8;
    }
    if (LG % 3 == 2) {
        constexpr int NB = L / 4, PER = (NB + T - 1) / T;
        float2 v[PER][4];
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (t + q * T < NB) stage_load<L, 4, INV>(buf, tw, off, Ns, t + q * T, v[q]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (t + q * T < NB) stage_store<L, 4>(buf, Ns, t + q * T, v[q]);
        __syncthreads();
    } else if (LG % 3 == 1) {
        constexpr int NB = L / 2, PER = (NB + T - 1) / T;
        float2 v[PER][2];
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (t + q * T < NB) stage_load<L, 2, INV>(buf, tw, off, Ns, t + q * T, v[q]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (t + q * T < NB) stage_store<L, 2>(buf, Ns, t + q * T, v[q]);
        __syncthreads();
    }
}

// Per-axis tables of a step, computed once in FP64 (k_axis_tables):
//   w[e]  = w^e = e^{-2 pi i e / M}            (every twiddle of the axis)
//   c1[k] = (1 - w^{ok}) / (1 - w^k),  c1[0] = o
//   c2[k] = (w^{(o+N)k} - 1) / (1 - w^k),  c2[0] = M - o - N
// so that the constant runs contribute a c1[k] + b c2[k] to bin k.
//   q[k]  = e^{-i z kf^2 / (2 k0)}, kf the bin's signed frequency (rad/m)
//           (the paraxial part of the transfer function along this axis)
struct AxisTab {
    const float2* w;
    const float2* c1;
    const float2* c2;
    const float2* q;
    int M, N, o;
};

__global__ void k_axis_tables(int M, int N, double sfreq, double z, double k0,
                              float2* __restrict__ w, float2* __restrict__ c1,
                              float2* __restrict__ c2, float2* __restrict__ q) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    const int o = (M - N) / 2;
    double s, c;
    sincospi(-2.0 * k / M, &s, &c);
    w[k] = make_float2(static_cast<float>(c), static_cast<float>(s));
    {
        const double kf = sfreq * (k < (M + 1) / 2 ? k : k - M);
        double ph = -z * kf * kf / (2.0 * k0);
        ph -= TWO_PI * rint(ph / TWO_PI);
        double qs, qc;
        sincos(ph, &qs, &qc);
        q[k] = make_float2(static_cast<float>(qc), static_cast<float>(qs));
    }
    if (k == 0) {
        c1[0] = make_float2(static_cast<float>(o), 0.f);
        c2[0] = make_float2(static_cast<float>(M - o - N), 0.f);
        return;
    }
    double so, co, sn, cn;
    sincospi(-2.0 * ((static_cast<long long>(o) * k) % M) / M, &so, &co);
    sincospi(-2.0 * ((static_cast<long long>(o + N) * k) % M) / M, &sn, &cn);
    // 1 / (1 - w^k)
    const double dr = 1.0 - c, di = -s, inv = 1.0 / (dr * dr + di * di);
    const double ir = dr * inv, ii = -di * inv;
    const double ar = 1.0 - co, ai = -so;          // 1 - w^{ok}
    const double br = cn - 1.0, bi = sn;           // w^{(o+N)k} - 1
    c1[k] = make_float2(static_cast<float>(ar * ir - ai * ii), static_cast<float>(ar * ii + ai * ir));
    c2[k] = make_float2(static_cast<float>(br * ir - bi * ii), static_cast<float>(br * ii + bi * ir));
}

// The stage-major twiddle table of fft_lds (L - 1 entries): for each stage
// (Ns, RD) and r = 1..RD-1, k < Ns, w_L^{k r L / (Ns RD)} = w^{k r L R / (Ns RD)}.
template <int L>
__device__ __forceinline__ void fill_twiddles(float2* tw, const AxisTab& X) {
    constexpr int LG = ilog2c(L), T = fft_threads(L);
    const int R = X.M / L;
    int Ns = 1, off = 0;
#pragma unroll
    for (int s = 0; s < LG / 3 + (LG % 3 ? 1 : 0); ++s) {
        const int RD = s < LG / 3 ? 8 : (LG % 3 == 2 ? 4 : 2);
        for (int i = threadIdx.x; i < (RD - 1) * Ns; i += T) {
            const int r = 1 + i / Ns, k = i - (r - 1) * Ns;
            tw[off + i] = X.w[k * r * (L / (Ns * RD)) * R];
        }
        off += (RD - 1) * Ns;
        Ns *= RD;
    }
}

__device__ __forceinline__ float2 wtab(const AxisTab& X, long long e) {
    return X.w[static_cast<int>(e & (X.M - 1))];
}

// Bin k of the padded line's transform from FFT_L output value f of branch
// r = k mod R: w^{ok} f + a c1[k] + b c2[k].
__device__ __forceinline__ float2 bin_value(const AxisTab& X, float2 f, int k, float2 a, float2 b) {
    return cadd(cadd(cmul(f, wtab(X, static_cast<long long>(X.o) * k)), cmul(a, X.c1[k])),
                cmul(b, X.c2[k]));
}

// Row r of a pass that streams rows through B[k * Nz + row]: groups of 16
// adjacent rows run on one XCD (workgroups go round-robin over the 8 XCDs),
// so the 8-byte column accesses of a group share 128-byte lines in its L2.
__device__ __forceinline__ int xcd_row(int b, int n_rows) {
    if (n_rows & 127) return b;
    const int xcd = b & 7, slot = b >> 3;
    return ((slot >> 4) << 7) | (xcd << 4) | (slot & 15);
}

__device__ __forceinline__ float2 lens_phase(float2 v, int iy, int iz, int Ny, int Nz, double dx,
                                             double dy, const LensArgs& lens) {
    const double x = (iy - 0.5 * (Ny - 1)) * dx;
    const double y = (iz - 0.5 * (Nz - 1)) * dy;
    const double r2 = x * x + y * y;
    double ph = -lens.half_k_over_f * r2;
    ph -= TWO_PI * rint(ph / TWO_PI);
    float s, c;
    sincosf(static_cast<float>(ph), &s, &c);
    float a = lens.uniform_amp;
    if (lens.parab_coef > 0.0) a *= static_cast<float>(exp(-lens.parab_coef * r2));
    return make_float2(a * (v.x * c - v.y * s), a * (v.x * s + v.y * c));
}

// Pass 1: forward transform along y of each data row (lens applied first),
// written column-major: B[ky * Nz + row].
template <int L>
__global__ __launch_bounds__(fft_threads(L)) void k_asp_rows_fwd(
    const float2* __restrict__ in, int Nz, const AxisTab Y, int edge, float2 padc, double dx,
    double dy, LensArgs lens, float2* __restrict__ B) {
    constexpr int T = fft_threads(L), E = L / T;
    __shared__ float2 buf[L + L / 8];
    __shared__ float2 tw[L];
    fill_twiddles<L>(tw, Y);
    const int Ny = Y.N, My = Y.M;
    const int row = xcd_row(blockIdx.x, Nz);
    const int t = threadIdx.x;
    const float2* src = in + static_cast<int64_t>(row) * Ny;
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + e * T;
        x[e] = make_float2(0.f, 0.f);
        if (n < Ny) {
            x[e] = src[n];
            if (lens.active) x[e] = lens_phase(x[e], n, row, Ny, Nz, dx, dy, lens);
        }
    }
    float2 a = padc, b = padc;
    if (edge) {
        a = src[0];
        b = src[Ny - 1];
        if (lens.active) {
            a = lens_phase(a, 0, row, Ny, Nz, dx, dy, lens);
            b = lens_phase(b, Ny - 1, row, Ny, Nz, dx, dy, lens);
        }
    }
    const int R = My / L;
    for (int r = 0; r < R; ++r) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int n = t + e * T;
            buf[lp(n)] = n < Ny ? cmul(x[e], wtab(Y, static_cast<long long>(n) * r))
                                : make_float2(0.f, 0.f);
        }
        __syncthreads();
        fft_lds<L, false>(buf, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = t + e * T;
            const int ky = R * m + r;
            B[static_cast<int64_t>(ky) * Nz + row] = bin_value(Y, buf[lp(m)], ky, a, b);
        }
    }
}

// Pass 2: per y-frequency column, forward along z, transfer function,
// inverse along z onto the Nz crop rows; in place in B.  sy, sz convert a
// signed bin index to rad/m (2 pi / (M d)).
//
// Transfer function.  With u = kt^2 / k0^2 the phase -z kt^2 / (k0 + sqrt(k0^2 - kt^2))
// is -z ky^2/(2 k0) - z kz^2/(2 k0) - z kt^2/(2 k0) u / (1 + sqrt(1 - u))^2:
// two paraxial terms (per column, and the z table q, both in FP64) and a
// remainder that the host bounds by 1e-3 rad for EXACT = false, formed in
// FP32 and exponentiated to third order.  EXACT = true (high numerical
// aperture, evanescent bins) forms the whole phase in FP64.
template <int L, bool EXACT>
__global__ __launch_bounds__(fft_threads(L)) __attribute__((amdgpu_waves_per_eu(4)))
void k_asp_cols(float2* __restrict__ B, const AxisTab Z, int My, int edge, float2 padc, double sy,
                double sz, double kw, double z) {
    constexpr int T = fft_threads(L), E = L / T;
    __shared__ float2 buf[L + L / 8];
    __shared__ float2 tw[L];
    fill_twiddles<L>(tw, Z);
    const int Nz = Z.N, Mz = Z.M, oz = Z.o;
    const int ky = blockIdx.x;
    const int t = threadIdx.x;
    float2* col = B + static_cast<int64_t>(ky) * Nz;
    float2 x[E], y[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + e * T;
        x[e] = n < Nz ? col[n] : make_float2(0.f, 0.f);
        y[e] = make_float2(0.f, 0.f);
    }
    float2 a, b;
    if (edge) {
        a = col[0];
        b = col[Nz - 1];
    } else {
        // a constant row transforms to M_y times the constant at ky = 0
        a = ky == 0 ? make_float2(padc.x * My, padc.y * My) : make_float2(0.f, 0.f);
        b = a;
    }
    const double kyf = sy * (ky < (My + 1) / 2 ? ky : ky - My);
    const double kw2 = kw * kw - kyf * kyf;
    float2 cy;
    {
        double ph = -z * kyf * kyf / (2.0 * kw);
        ph -= TWO_PI * rint(ph / TWO_PI);
        float sn, cs;
        sincosf(static_cast<float>(ph), &sn, &cs);
        cy = make_float2(cs, sn);
    }
    const float zk = static_cast<float>(z / (2.0 * kw)), ik2 = static_cast<float>(1.0 / (kw * kw));
    const float kyf2 = static_cast<float>(kyf * kyf), szf = static_cast<float>(sz);
    const int R = Mz / L;
    for (int r = 0; r < R; ++r) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int n = t + e * T;
            buf[lp(n)] = n < Nz ? cmul(x[e], wtab(Z, static_cast<long long>(n) * r))
                                : make_float2(0.f, 0.f);
        }
        __syncthreads();
        fft_lds<L, false>(buf, tw);
#pragma unroll 2
        for (int e = 0; e < E; ++e) {
            const int m = t + e * T;
            const int kz = R * m + r;
            const float2 V = bin_value(Z, buf[lp(m)], kz, a, b);
            const int kzs = kz < (Mz + 1) / 2 ? kz : kz - Mz;
            float2 H;
            if (EXACT) {
                const double kzf = sz * kzs;
                const double kx2 = kw2 - kzf * kzf;
                double ph, amp = 1.0;
                if (kx2 >= 0.0) {
                    ph = -z * (kyf * kyf + kzf * kzf) / (sqrt(kx2) + kw);
                } else {
                    ph = -z * kw;
                    amp = exp(-fabs(z) * sqrt(-kx2));
                }
                ph -= TWO_PI * rint(ph / TWO_PI);
                float sn, cs;
                sincosf(static_cast<float>(ph), &sn, &cs);
                const float am = static_cast<float>(amp);
                H = make_float2(am * cs, am * sn);
            } else {
                const float kzf = szf * static_cast<float>(kzs);
                const float kt2 = kyf2 + kzf * kzf;
                const float u = kt2 * ik2;
                const float sq = 1.f + sqrtf(1.f - u);
                const float rem = -zk * kt2 * u / (sq * sq);
                const float r2 = rem * rem;
                H = cmul(cmul(cy, Z.q[kz]),
                         make_float2(1.f - 0.5f * r2, rem * (1.f - r2 * (1.f / 6.f))));
            }
            buf[lp(m)] = cmul(V, H);
        }
        __syncthreads();
        fft_lds<L, true>(buf, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int n = t + e * T;
            if (n < Nz)
                y[e] = cadd(y[e], cmulc(buf[lp((oz + n) & (L - 1))],
                                        wtab(Z, static_cast<long long>(oz + n) * r)));
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + e * T;
        if (n < Nz) col[n] = y[e];
    }
}

// Pass 3: inverse along y of each crop row, scaled, into the output field.
template <int L>
__global__ __launch_bounds__(fft_threads(L)) void k_asp_rows_inv(const float2* __restrict__ B,
                                                                int Nz, const AxisTab Y,
                                                                float scale,
                                                                float2* __restrict__ out) {
    constexpr int T = fft_threads(L), E = L / T;
    __shared__ float2 buf[L + L / 8];
    __shared__ float2 tw[L];
    fill_twiddles<L>(tw, Y);
    const int Ny = Y.N, My = Y.M, oy = Y.o;
    const int row = xcd_row(blockIdx.x, Nz);
    const int t = threadIdx.x;
    float2 acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = make_float2(0.f, 0.f);
    const int R = My / L;
    for (int r = 0; r < R; ++r) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int m = t + e * T;
            buf[lp(m)] = B[static_cast<int64_t>(R * m + r) * Nz + row];
        }
        __syncthreads();
        fft_lds<L, true>(buf, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = t + e * T;
            if (i < Ny)
                acc[e] = cadd(acc[e], cmulc(buf[lp((oy + i) & (L - 1))],
                                            wtab(Y, static_cast<long long>(oy + i) * r)));
        }
    }
    float2* dst = out + static_cast<int64_t>(row) * Ny;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = t + e * T;
        if (i < Ny) dst[i] = make_float2(acc[e].x * scale, acc[e].y * scale);
    }
}

int fft_len(int n) {
    int L = FFT_MIN;
    while (L < n) L *= 2;
    return L;
}

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <int L>
void launch_rows_fwd(hipStream_t st, const float2* in, int Nz, const AxisTab& Y, int edge,
                     float2 padc, double dx, double dy, LensArgs lens, float2* B) {
    hipLaunchKernelGGL(k_asp_rows_fwd<L>, dim3(Nz), dim3(fft_threads(L)), 0, st, in, Nz, Y, edge,
                       padc, dx, dy, lens, B);
}
template <int L>
void launch_cols(hipStream_t st, bool exact, float2* B, const AxisTab& Z, int My, int edge,
                 float2 padc, double sy, double sz, double kw, double z) {
    if (exact)
        hipLaunchKernelGGL((k_asp_cols<L, true>), dim3(My), dim3(fft_threads(L)), 0, st, B, Z, My,
                           edge, padc, sy, sz, kw, z);
    else
        hipLaunchKernelGGL((k_asp_cols<L, false>), dim3(My), dim3(fft_threads(L)), 0, st, B, Z,
                           My, edge, padc, sy, sz, kw, z);
}
template <int L>
void launch_rows_inv(hipStream_t st, const float2* B, int Nz, const AxisTab& Y, float scale,
                     float2* out) {
    hipLaunchKernelGGL(k_asp_rows_inv<L>, dim3(Nz), dim3(fft_threads(L)), 0, st, B, Nz, Y, scale,
                       out);
}

#define MOSAIC_FFT_DISPATCH(L_, CALL) \
    switch (L_) {                     \
        case 64: CALL(64); break;     \
        case 128: CALL(128); break;   \
        case 256: CALL(256); break;   \
        case 512: CALL(512); break;   \
        case 1024: CALL(1024); break; \
        case 2048: CALL(2048); break; \
        case 4096: CALL(4096); break; \
        default: CALL(8192); break;   \
    }

// The pruned step's size rules.
bool pruned_ok(int Ny, int Nz, int Ny2, int Nz2) {
    if (!pow2(Ny2) || !pow2(Nz2)) return false;
    const int Ly = fft_len(Ny), Lz = fft_len(Nz);
    return Ly <= Ny2 && Lz <= Nz2 && Ly <= FFT_MAX && Lz <= FFT_MAX;
}

// Table space of a pruned step: w, c1, c2, q per axis.
size_t pruned_table_count(int Ny2, int Nz2) { return 4 * (static_cast<size_t>(Ny2) + Nz2); }

int pruned_step(hipStream_t st, const float2* in, int Ny, int Nz, double dy, double dz, double k,
                double z, int Ny2, int Nz2, int edge, float2 padc, const LensArgs& lens,
                float2* B, float2* tabs, float2* out) {
    const int Ly = fft_len(Ny), Lz = fft_len(Nz);
    const double sy = TWO_PI / (static_cast<double>(Ny2) * dy);
    const double sz = TWO_PI / (static_cast<double>(Nz2) * dz);
    AxisTab Y{tabs, tabs + Ny2, tabs + 2 * Ny2, tabs + 3 * Ny2, Ny2, Ny, (Ny2 - Ny) / 2};
    float2* tz = tabs + 4 * static_cast<size_t>(Ny2);
    AxisTab Zt{tz, tz + Nz2, tz + 2 * Nz2, tz + 3 * Nz2, Nz2, Nz, (Nz2 - Nz) / 2};
    hipLaunchKernelGGL(k_axis_tables, dim3(blocks_for(Ny2, 256)), dim3(256), 0, st, Ny2, Ny, sy,
                       z, k, tabs, tabs + Ny2, tabs + 2 * Ny2, tabs + 3 * Ny2);
    hipLaunchKernelGGL(k_axis_tables, dim3(blocks_for(Nz2, 256)), dim3(256), 0, st, Nz2, Nz, sz,
                       z, k, tz, tz + Nz2, tz + 2 * Nz2, tz + 3 * Nz2);
    MOSAIC_HIP_CHECK(hipGetLastError());
    // the split transfer function applies where its remainder stays below
    // 1e-3 rad on the whole grid (no evanescent bins)
    const double kt2max = (0.5 * Ny2 * sy) * (0.5 * Ny2 * sy) + (0.5 * Nz2 * sz) * (0.5 * Nz2 * sz);
    const double umax = kt2max / (k * k);
    bool exact = umax >= 0.5;
    if (!exact) {
        const double sq = 1.0 + std::sqrt(1.0 - umax);
        exact = std::fabs(z) * kt2max / (2.0 * k) * umax / (sq * sq) > 1e-3;
    }
#define CALL_RF(L) launch_rows_fwd<L>(st, in, Nz, Y, edge, padc, dy, dz, lens, B)
    MOSAIC_FFT_DISPATCH(Ly, CALL_RF)
#undef CALL_RF
    MOSAIC_HIP_CHECK(hipGetLastError());
#define CALL_C(L) launch_cols<L>(st, exact, B, Zt, Ny2, edge, padc, sy, sz, k, z)
    MOSAIC_FFT_DISPATCH(Lz, CALL_C)
#undef CALL_C
    MOSAIC_HIP_CHECK(hipGetLastError());
    const float scale = static_cast<float>(1.0 / (static_cast<double>(Ny2) * Nz2));
#define CALL_RI(L) launch_rows_inv<L>(st, B, Nz, Y, scale, out)
    MOSAIC_FFT_DISPATCH(Ly, CALL_RI)
#undef CALL_RI
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
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
    c.take<float2>(pruned_table_count(Ny2, Nz2));
    *bytes = c.off + 256;
    return MOSAIC_OK;
}

extern "C" int mosaic_asp_propagate(const float* field_in, int Ny, int Nz, double dy, double dz,
                                    double wavelength, double z, int Ny2, int Nz2, int pad_edge,
                                    float pad_re, float pad_im, const mosaic_lens_t* lens,
                                    int method, float* field_out, void* workspace,
                                    size_t workspace_bytes, mosaic_stream_t stream) {
    if (!field_in || !field_out || Ny < 1 || Nz < 1 || Ny2 < Ny || Nz2 < Nz ||
        !(dy > 0.0) || !(dz > 0.0) || !(wavelength > 0.0) || !workspace || method < 0 ||
        method > 2)
        return MOSAIC_ERR_ARGS;
    const bool pruned = pruned_ok(Ny, Nz, Ny2, Nz2);
    if (method == 2 && !pruned) return MOSAIC_ERR_UNSUPPORTED;
    hipfftHandle plan;
    size_t work = 0;
    int rc = get_plan(Nz2, Ny2, &plan, &work);
    if (rc != MOSAIC_OK) return rc;
    Carver c(workspace);
    float2* F = c.take<float2>(static_cast<size_t>(Ny2) * Nz2);
    void* wa = c.take<char>(work);
    float2* tabs = c.take<float2>(pruned_table_count(Ny2, Nz2));
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
    if (pruned && method != 1)
        return pruned_step(st, reinterpret_cast<const float2*>(field_in), Ny, Nz, dy, dz, k, z,
                           Ny2, Nz2, pad_edge ? 1 : 0, make_float2(pad_re, pad_im), L, F,
                           tabs, reinterpret_cast<float2*>(field_out));
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
