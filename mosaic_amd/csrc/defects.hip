// Defect displacement fields applied on the device, for generating the
// surrogate's training images: one dislocation (isotropic elasticity, mixed
// screw/edge), a homogeneous hydrostatic strain and one intrinsic stacking
// fault, added to crystal-frame lattice positions and rotated into the sample
// frame.  The reference builds these fields on the host per chunk
// (Defects.py: dislocation_displacement, apply_stacking_fault_chunk at
// 3951-3990); here one launch displaces a whole crystal.

#include "common.h"

namespace mosaic {
namespace {

struct DefectArgs {
    float core[3];      // a point on the dislocation line [A]
    float line[3];      // unit line direction
    float e1[3], e2[3]; // unit vectors normal to the line; e1 along the edge part of b
    float b_screw;      // b . line [A]
    float b_edge;       // |b - (b . line) line| [A]
    float nu;           // Poisson ratio
    float strain;       // hydrostatic strain about the origin
    float fault_n[3];   // unit normal of the fault plane
    float fault_c;      // plane offset: fault_n . r = fault_c
    float fault_shift[3];  // displacement of atoms on the positive side [A]
    float R[9];         // crystal -> sample rotation
};

__global__ void k_defects(const float* __restrict__ base, int64_t n, DefectArgs D,
                          float* __restrict__ out) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const float r[3] = {base[3 * i], base[3 * i + 1], base[3 * i + 2]};
    float u[3] = {0.f, 0.f, 0.f};
    // dislocation
    float d[3] = {r[0] - D.core[0], r[1] - D.core[1], r[2] - D.core[2]};
    const float x = d[0] * D.e1[0] + d[1] * D.e1[1] + d[2] * D.e1[2];
    const float y = d[0] * D.e2[0] + d[1] * D.e2[1] + d[2] * D.e2[2];
    const float r2 = x * x + y * y + 1.0f;          // 1 A^2 core regularisation
    const float th = atan2f(y, x);
    const float inv2pi = 0.15915494309189535f;
    const float us = D.b_screw * th * inv2pi;
    const float k1 = 1.f / (2.f * (1.f - D.nu)), k2 = 1.f / (4.f * (1.f - D.nu));
    const float ux = D.b_edge * inv2pi * (th + x * y * k1 / r2);
    const float uy = -D.b_edge * inv2pi * ((1.f - 2.f * D.nu) * k2 * logf(r2) +
                                           (x * x - y * y) * k2 / r2);
    for (int c = 0; c < 3; ++c) u[c] += us * D.line[c] + ux * D.e1[c] + uy * D.e2[c];
    // homogeneous strain
    for (int c = 0; c < 3; ++c) u[c] += D.strain * r[c];
    // stacking fault
    const float side =
        r[0] * D.fault_n[0] + r[1] * D.fault_n[1] + r[2] * D.fault_n[2] - D.fault_c;
    if (side > 0.f)
        for (int c = 0; c < 3; ++c) u[c] += D.fault_shift[c];
    const float p[3] = {r[0] + u[0], r[1] + u[1], r[2] + u[2]};
    for (int c = 0; c < 3; ++c)
        out[3 * i + c] = D.R[3 * c] * p[0] + D.R[3 * c + 1] * p[1] + D.R[3 * c + 2] * p[2];
}

}  // namespace
}  // namespace mosaic

using namespace mosaic;

extern "C" int mosaic_defect_positions(int64_t n, const float* base, const float* params,
                                       float* out, mosaic_stream_t stream) {
    if (n < 0 || !params) return MOSAIC_ERR_ARGS;
    if (n == 0) return MOSAIC_OK;
    if (!base || !out) return MOSAIC_ERR_ARGS;
    DefectArgs D;
    const float* q = params;
    for (int c = 0; c < 3; ++c) {
        D.core[c] = q[c];
        D.line[c] = q[3 + c];
        D.e1[c] = q[6 + c];
        D.e2[c] = q[9 + c];
        D.fault_n[c] = q[16 + c];
        D.fault_shift[c] = q[19 + c];
    }
    D.b_screw = q[12];
    D.b_edge = q[13];
    D.nu = q[14];
    D.strain = q[15];
    for (int c = 0; c < 9; ++c) D.R[c] = q[22 + c];
    D.fault_c = q[31];
    hipLaunchKernelGGL(k_defects, dim3(blocks_for(n, 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), base, n, D, out);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}
