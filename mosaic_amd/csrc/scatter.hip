// compute_scattering_cffi on the GPU: the reference's CPU kernel
// (Beam.py:629-816) with the same arguments, each array a device pointer,
// ordered on a HIP stream.
//
// Per pixel p and atom j (lab frame, metres):
//     d = T_p - r_j,  r = |d|,  cos 2theta = d_x / r,
//     S = f0(Q) + f' + i f''  (minus f0(0) + f' + i f'' where remove_forward
//                              and Q < Q_cut[p]),
//     out[p] += amp_j S e^{i k (x_j + r)} r_e [R0 / r] [sqrt P]
// with R0 = |T_p| and Q_cut half the pixel's diagonal in Q (Beam.py:718-746).
// Pairs with r == 0 are skipped (Beam.py:765).
//
// MI355X formulation: one lane owns SC_PIX pixels; a workgroup walks one
// slice of the atoms, staged 256 at a time in LDS (positions in FP64, the WK
// row prepared for evaluation in 1 - cos 2theta).  The pair geometry is FP64
// and cancellation-free:
//     phase = Theta0(p) + K (x_j + (r - R0)),   r - R0 = (|r_j|^2 - 2 T.r_j) / (r + R0),
// with Theta0 = frac(K R0) in double-double (K = 1 / lambda), so the phase is
// exact to ~1e-12 revolutions where the reference's fmod(k (x + r), 2 pi)
// carries ~1e-5 rad of rounding; 1 - cos 2theta = sin^2 2theta / (1 + cos 2theta)
// in the forward hemisphere.  The form factor (five v_exp_f32), polarization
// and the phasor (v_sin/v_cos on the reduced phase, in revolutions) are FP32;
// the pair terms accumulate in FP64.  Slice partials (FP64) are summed in
// slice order and added, as float32, into out_r / out_i, as the reference
// adds its FP64 accumulators once per call.
#include <algorithm>
#include <mutex>

#include "common.h"
#include "pair_physics.h"

namespace mosaic {
namespace {

constexpr int SC_THREADS = 256;
constexpr int SC_PIX = 4;
constexpr int SC_CHUNK = 256;         // atoms staged per LDS step
constexpr int SC_TARGET_BLOCKS = 2048;
constexpr int SC_MAX_SLICES = 256;
constexpr double SC_RE = 2.81794092e-15;

struct ScArgs {
    const double* pos;
    const float* f0p;
    const float* f0z;
    const float* anr;
    const float* ani;
    const float* amr;
    const float* ami;
    const double* cx;
    const double* cy;
    const double* cz;
    double2* partial;
    int n_atoms;
    int Ny, Nz, npix;
    int n_slices;
    int atoms_per_slice;
    double k;
    double K;        // 1 / lambda
    double cq;       // 2 k^2 S_PER_Q^2 log2(e)
    double pol_deg;
    double pol_e[3];
    int remove_forward;
    int apply_pol;
    int apply_decay;
};

// Q of the pixel's right and upper neighbours, as the reference forms its
// cut-off (Beam.py:718-746).
__device__ double pixel_qcut(const double* cx, const double* cy, const double* cz, int p,
                             int Ny, int Nz, double k) {
    const int ix = p % Ny, iy = p / Ny;
    const double tx = cx[p], ty = cy[p], tz = cz[p];
    const double R0 = sqrt(tx * tx + ty * ty + tz * tz);
    double ux = 0.0, uy = 0.0, uz = 0.0;
    if (R0 > 0.0) { ux = tx / R0; uy = ty / R0; uz = tz / R0; }
    const int nb[2] = {(ix + 1 < Ny) ? (p + 1) : ((ix > 0) ? (p - 1) : p),
                       (iy + 1 < Nz) ? (p + Ny) : ((iy > 0) ? (p - Ny) : p)};
    double q2[2];
    for (int t = 0; t < 2; ++t) {
        const double rx = cx[nb[t]], ry = cy[nb[t]], rz = cz[nb[t]];
        const double Rr = sqrt(rx * rx + ry * ry + rz * rz);
        double cd = (Rr > 0.0) ? (ux * rx + uy * ry + uz * rz) / Rr : 1.0;
        cd = fmin(1.0, fmax(-1.0, cd));
        q2[t] = k * sqrt(fmax(0.0, 2.0 * (1.0 - cd)));
    }
    return 0.5 * sqrt(q2[0] * q2[0] + q2[1] * q2[1]);
}

__global__ __launch_bounds__(SC_THREADS) void k_scatter(const ScArgs A) {
    __shared__ double4 s_pos[SC_CHUNK];          // x, y, z, |r|^2
    __shared__ float s_row[SC_CHUNK][16];        // a[5], nb[5], c, f0z, f', f'', amp re, im

    const int slice = blockIdx.y;
    const int a_begin = slice * A.atoms_per_slice;
    const int a_end = min(A.n_atoms, a_begin + A.atoms_per_slice);

    int pix[SC_PIX];
    bool valid[SC_PIX];
    double tx[SC_PIX], ty[SC_PIX], tz[SC_PIX], th0[SC_PIX], R0[SC_PIX];
    float ocut[SC_PIX];
    double acc_r[SC_PIX], acc_i[SC_PIX];
#pragma unroll
    for (int i = 0; i < SC_PIX; ++i) {
        const int p = blockIdx.x * (SC_THREADS * SC_PIX) + i * SC_THREADS + threadIdx.x;
        valid[i] = p < A.npix;
        pix[i] = valid[i] ? p : A.npix - 1;
        tx[i] = A.cx[pix[i]];
        ty[i] = A.cy[pix[i]];
        tz[i] = A.cz[pix[i]];
        double f, r, r_lo;
        origin_phase(tx[i], ty[i], tz[i], A.k, f, r, r_lo);
        th0[i] = f;
        R0[i] = r + r_lo;
        ocut[i] = -1.0f;
        if (A.remove_forward) {
            // Q < Q_cut  <=>  1 - cos 2theta < (Q_cut / k)^2 / 2
            const double qc = pixel_qcut(A.cx, A.cy, A.cz, pix[i], A.Ny, A.Nz, A.k) / A.k;
            ocut[i] = static_cast<float>(0.5 * qc * qc);
        }
        acc_r[i] = 0.0;
        acc_i[i] = 0.0;
    }
    const float pdeg = static_cast<float>(A.pol_deg);
    const double ex = A.pol_e[0], ey = A.pol_e[1], ez = A.pol_e[2];

    for (int base = a_begin; base < a_end; base += SC_CHUNK) {
        const int n = min(SC_CHUNK, a_end - base);
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < n) {
            const int64_t a = base + threadIdx.x;
            const double x = A.pos[3 * a], y = A.pos[3 * a + 1], z = A.pos[3 * a + 2];
            s_pos[threadIdx.x] = make_double4(x, y, z, fma(x, x, fma(y, y, z * z)));
            const float* w = A.f0p + 11 * a;
            float* row = s_row[threadIdx.x];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                row[m] = w[m];
                row[5 + m] = static_cast<float>(-static_cast<double>(w[6 + m]) * A.cq);
            }
            row[10] = w[5];
            row[11] = A.f0z[a];
            row[12] = A.anr[a];
            row[13] = A.ani[a];
            row[14] = A.amr[a];
            row[15] = A.ami[a];
        }
        __syncthreads();
        for (int j = 0; j < n; ++j) {
            const double4 P = s_pos[j];
            const float* row = s_row[j];
            float fa[5], fnb[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) { fa[m] = row[m]; fnb[m] = row[5 + m]; }
            const float fc = row[10], f0z = row[11], fp = row[12], fpp = row[13];
            const float amr = row[14], ami = row[15];
#pragma unroll
            for (int i = 0; i < SC_PIX; ++i) {
                const double dx = tx[i] - P.x, dy = ty[i] - P.y, dz = tz[i] - P.z;
                const double syz = fma(dy, dy, dz * dz);
                const double r2 = fma(dx, dx, syz);
                if (r2 == 0.0) continue;
                const double r = sqrt(r2);
                const double ir2 = 1.0 / r2;
                const double ir = r * ir2;
                // phase in revolutions, exact to FP64 rounding of small terms
                const double ta = fma(tx[i], P.x, fma(ty[i], P.y, tz[i] * P.z));
                const double dR = fma(-2.0, ta, P.w) / (r + R0[i]);
                double th = fma(A.K, P.x + dR, th0[i]);
                th -= rint(th);
                const float phs = static_cast<float>(th);
                const float c2t = static_cast<float>(dx * ir);
                const float s2t = static_cast<float>(syz * ir2);
                const float omc = c2t > 0.0f ? s2t / (1.0f + c2t) : 1.0f - c2t;
                float sr = form_f0(fa, fnb, fc, omc) + fp, si = fpp;
                if (A.remove_forward && omc < ocut[i]) {
                    sr -= f0z + fp;
                    si -= fpp;
                }
                float scale = 1.0f;
                if (A.apply_decay && R0[i] > 0.0) scale = static_cast<float>(R0[i] * ir);
                if (A.apply_pol) {
                    const double cx_ = ey * dz - ez * dy, cy_ = ez * dx - ex * dz,
                                 cz_ = ex * dy - ey * dx;
                    const float ec2 = static_cast<float>(fma(cx_, cx_, fma(cy_, cy_, cz_ * cz_)) * ir2);
                    scale *= pol_factor(c2t, ec2, pdeg);
                }
                const float tr = (amr * sr - ami * si) * scale;
                const float ti = (amr * si + ami * sr) * scale;
                const float cs = __builtin_amdgcn_cosf(phs);
                const float sn = __builtin_amdgcn_sinf(phs);
                acc_r[i] += static_cast<double>(fmaf(tr, cs, -ti * sn));
                acc_i[i] += static_cast<double>(fmaf(tr, sn, ti * cs));
            }
        }
    }
    double2* out = A.partial + static_cast<int64_t>(slice) * A.npix;
#pragma unroll
    for (int i = 0; i < SC_PIX; ++i)
        if (valid[i]) out[pix[i]] = make_double2(acc_r[i], acc_i[i]);
}

// Sum the slice partials in slice order, scale by r_e and add into the
// float32 outputs.
__global__ void k_scatter_reduce(const double2* __restrict__ partial, int n_slices, int npix,
                                 float* __restrict__ out_r, float* __restrict__ out_i) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    double r = 0.0, im = 0.0;
    for (int s = 0; s < n_slices; ++s) {
        const double2 v = partial[static_cast<int64_t>(s) * npix + p];
        r += v.x;
        im += v.y;
    }
    out_r[p] += static_cast<float>(r * SC_RE);
    out_i[p] += static_cast<float>(im * SC_RE);
}

__global__ void k_sincos_rev(const float* __restrict__ x, int64_t n, float* __restrict__ s,
                             float* __restrict__ c) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    s[i] = __builtin_amdgcn_sinf(x[i]);
    c[i] = __builtin_amdgcn_cosf(x[i]);
}

struct ScPlan {
    int pix_blocks;
    int n_slices;
    int atoms_per_slice;
};

// Atom slices over gridDim.y fill the chip when the detector is small: about
// SC_TARGET_BLOCKS workgroups, slices of whole LDS steps.
ScPlan sc_plan(int atom_count, int npix, int max_slices) {
    ScPlan P;
    P.pix_blocks = blocks_for(npix, SC_THREADS * SC_PIX);
    const int steps = std::max(1, (atom_count + SC_CHUNK - 1) / SC_CHUNK);
    int ns = std::max(1, SC_TARGET_BLOCKS / P.pix_blocks);
    ns = std::min(ns, std::min(steps, std::min(SC_MAX_SLICES, max_slices)));
    const int steps_per = (steps + ns - 1) / ns;
    P.atoms_per_slice = steps_per * SC_CHUNK;
    P.n_slices = (steps + steps_per - 1) / steps_per;
    return P;
}

bool sc_args_ok(int atom_count, int Ny, int Nz, double k_val) {
    return atom_count >= 0 && Ny >= 1 && Nz >= 1 && k_val > 0.0 &&
           static_cast<int64_t>(Ny) * Nz <= (int64_t(1) << 30);
}

int sc_launch(int atom_count, const double* positions, const float* f0_params,
              const float* f0_zero, int remove_forward, const float* s_anom_real,
              const float* s_anom_imag, const float* initial_amp_r, const float* initial_amp_i,
              int Ny, int Nz, const double* coords_x, const double* coords_y,
              const double* coords_z, double k_val, int apply_pol, double pol_deg, double pol_ex,
              double pol_ey, double pol_ez, int apply_spherical_decay, float* out_r,
              float* out_i, double2* partial, const ScPlan& P, hipStream_t st) {
    const int npix = Ny * Nz;
    if (atom_count == 0) return MOSAIC_OK;
    ScArgs A;
    A.pos = positions;
    A.f0p = f0_params;
    A.f0z = f0_zero;
    A.anr = s_anom_real;
    A.ani = s_anom_imag;
    A.amr = initial_amp_r;
    A.ami = initial_amp_i;
    A.cx = coords_x;
    A.cy = coords_y;
    A.cz = coords_z;
    A.partial = partial;
    A.n_atoms = atom_count;
    A.Ny = Ny;
    A.Nz = Nz;
    A.npix = npix;
    A.n_slices = P.n_slices;
    A.atoms_per_slice = P.atoms_per_slice;
    A.k = k_val;
    A.K = k_val * INV_2PI_HI;
    A.cq = 2.0 * k_val * k_val * S_PER_Q * S_PER_Q * LOG2E_D;
    A.pol_deg = pol_deg;
    A.pol_e[0] = pol_ex;
    A.pol_e[1] = pol_ey;
    A.pol_e[2] = pol_ez;
    A.remove_forward = remove_forward ? 1 : 0;
    A.apply_pol = apply_pol ? 1 : 0;
    A.apply_decay = apply_spherical_decay ? 1 : 0;
    hipLaunchKernelGGL(k_scatter, dim3(P.pix_blocks, P.n_slices), dim3(SC_THREADS), 0, st, A);
    MOSAIC_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_scatter_reduce, dim3(blocks_for(npix, 256)), dim3(256), 0, st,
                       partial, P.n_slices, npix, out_r, out_i);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

// Scratch of the allocation-free entry point: its slice partials live in
// this file-scope device array instead of a caller workspace.  Each GPU the
// library runs on holds its own copy until the process ends.  In host code
// the name is only a stand-in; the device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's scratch address, looked up on the device's first call.
int scratch_address(void** out) {
    int dev = -1;
    MOSAIC_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return MOSAIC_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MOSAIC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return MOSAIC_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MOSAIC_OK;
}

}  // namespace
}  // namespace mosaic

using namespace mosaic;

extern "C" int mosaic_scatter_workspace(int atom_count, int Ny, int Nz, size_t* bytes) {
    if (!bytes || !sc_args_ok(atom_count, Ny, Nz, 1.0)) return MOSAIC_ERR_ARGS;
    const ScPlan P = sc_plan(atom_count, Ny * Nz, SC_MAX_SLICES);
    *bytes = static_cast<size_t>(P.n_slices) * static_cast<size_t>(Ny) * Nz * sizeof(double2);
    return MOSAIC_OK;
}

extern "C" int mosaic_scatter(int atom_count, const double* positions, const float* f0_params,
                              const float* f0_zero, int remove_forward,
                              const float* s_anom_real, const float* s_anom_imag,
                              const float* initial_amp_r, const float* initial_amp_i, int Ny,
                              int Nz, const double* coords_x, const double* coords_y,
                              const double* coords_z, double k_val, int apply_pol,
                              double pol_deg, double pol_ex, double pol_ey, double pol_ez,
                              int apply_spherical_decay, float* out_r, float* out_i,
                              void* workspace, size_t workspace_bytes, mosaic_stream_t stream) {
    if (!sc_args_ok(atom_count, Ny, Nz, k_val)) return MOSAIC_ERR_ARGS;
    if (!coords_x || !coords_y || !coords_z || !out_r || !out_i) return MOSAIC_ERR_ARGS;
    if (atom_count > 0 && (!positions || !f0_params || !f0_zero || !s_anom_real ||
                           !s_anom_imag || !initial_amp_r || !initial_amp_i))
        return MOSAIC_ERR_ARGS;
    if (atom_count == 0) return MOSAIC_OK;
    const ScPlan P = sc_plan(atom_count, Ny * Nz, SC_MAX_SLICES);
    const size_t need = static_cast<size_t>(P.n_slices) * static_cast<size_t>(Ny) * Nz *
                        sizeof(double2);
    if (!workspace || workspace_bytes < need) return MOSAIC_ERR_WORKSPACE;
    return sc_launch(atom_count, positions, f0_params, f0_zero, remove_forward, s_anom_real,
                     s_anom_imag, initial_amp_r, initial_amp_i, Ny, Nz, coords_x, coords_y,
                     coords_z, k_val, apply_pol, pol_deg, pol_ex, pol_ey, pol_ez,
                     apply_spherical_decay, out_r, out_i, static_cast<double2*>(workspace), P,
                     static_cast<hipStream_t>(stream));
}

// mosaic_scatter without a workspace: the slice partials (16 B per pixel per
// atom slice) go into the library's per-device scratch, so the call
// allocates nothing.  Every partial the reduction reads is written first by
// the same call, so nothing needs clearing; calls that overlap on one GPU
// share the scratch and race.
extern "C" int mosaic_scatter_noalloc(int atom_count, const double* positions,
                                      const float* f0_params, const float* f0_zero,
                                      int remove_forward, const float* s_anom_real,
                                      const float* s_anom_imag, const float* initial_amp_r,
                                      const float* initial_amp_i, int Ny, int Nz,
                                      const double* coords_x, const double* coords_y,
                                      const double* coords_z, double k_val, int apply_pol,
                                      double pol_deg, double pol_ex, double pol_ey,
                                      double pol_ez, int apply_spherical_decay, float* out_r,
                                      float* out_i, mosaic_stream_t stream) {
    if (!sc_args_ok(atom_count, Ny, Nz, k_val)) return MOSAIC_ERR_ARGS;
    if (!coords_x || !coords_y || !coords_z || !out_r || !out_i) return MOSAIC_ERR_ARGS;
    if (atom_count > 0 && (!positions || !f0_params || !f0_zero || !s_anom_real ||
                           !s_anom_imag || !initial_amp_r || !initial_amp_i))
        return MOSAIC_ERR_ARGS;
    const size_t per_slice = static_cast<size_t>(Ny) * Nz * sizeof(double2);
    if (per_slice > NOALLOC_SCRATCH_BYTES) return MOSAIC_ERR_TOO_LARGE;
    if (atom_count == 0) return MOSAIC_OK;
    void* base = nullptr;
    const int rc = scratch_address(&base);
    if (rc != MOSAIC_OK) return rc;
    const int max_slices = static_cast<int>(NOALLOC_SCRATCH_BYTES / per_slice);
    const ScPlan P = sc_plan(atom_count, Ny * Nz, max_slices);
    return sc_launch(atom_count, positions, f0_params, f0_zero, remove_forward, s_anom_real,
                     s_anom_imag, initial_amp_r, initial_amp_i, Ny, Nz, coords_x, coords_y,
                     coords_z, k_val, apply_pol, pol_deg, pol_ex, pol_ey, pol_ez,
                     apply_spherical_decay, out_r, out_i, static_cast<double2*>(base), P,
                     static_cast<hipStream_t>(stream));
}

extern "C" int mosaic_sincos_rev(int64_t n, const float* x, float* sin_out, float* cos_out,
                                 mosaic_stream_t stream) {
    if (n < 0) return MOSAIC_ERR_ARGS;
    if (n == 0) return MOSAIC_OK;
    if (!x || !sin_out || !cos_out) return MOSAIC_ERR_ARGS;
    hipLaunchKernelGGL(k_sincos_rev, dim3(blocks_for(n, 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), x, n, sin_out, cos_out);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}
