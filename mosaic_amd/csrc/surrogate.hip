// Kernels of the learned rocking-scan surrogate: a BF16 MFMA GEMM with FP32
// accumulation (forward and both backward products of every layer, batched
// over ensemble members), and the elementwise passes around it (bias +
// activation and its backward, bias gradients, the 2x2 transposed-convolution
// pixel shuffle, loss, Adam).  Activations and master weights are FP32 in
// HBM; tiles are rounded to BF16 (round to nearest even) as they are staged in
// LDS, so no BF16 copies are kept.
//
// GEMM: C[m][n] = alpha * sum_k op(A)[m][k] op(B)[k][n] + beta * C[m][n]
//   op(A)[m][k] = A[m*lda + k] (TA = 0) or A[k*lda + m] (TA = 1)
//   op(B)[k][n] = B[k*ldb + n] (TB = 0) or B[n*ldb + k] (TB = 1)
// A 64x64 output tile per 256-thread workgroup; each wave owns a 32x32
// quadrant held in one v_mfma_f32_32x32x8_bf16 accumulator (16 FP32 per
// lane); K advances 32 per LDS stage (4 MFMAs).  grid.z indexes the batch
// (ensemble member) with element strides sA, sB, sC.

#include "common.h"

namespace mosaic {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int GBM = 64, GBN = 64, GBK = 32;

__device__ __forceinline__ unsigned short to_bf16(float x) {
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<unsigned short>(u >> 16);  // inf/nan
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return static_cast<unsigned short>(u >> 16);
}

template <int TA, int TB>
__global__ __launch_bounds__(256) void k_gemm_bf16(const float* __restrict__ A,
                                                   const float* __restrict__ B,
                                                   float* __restrict__ C, int M, int N, int K,
                                                   int lda, int ldb, int ldc, long long sA,
                                                   long long sB, long long sC, float alpha,
                                                   float beta, int splits, int kchunk) {
    // LDS tiles as bf16: As[m][k], Bs[n][k] (k contiguous, padded against
    // bank conflicts)
    __shared__ unsigned short As[GBM][GBK + 4];
    __shared__ unsigned short Bs[GBN][GBK + 4];
    // grid.z = batch x K-splits; a split adds its partial product atomically
    const int z = blockIdx.z / splits, split = blockIdx.z % splits;
    const int kb = split * kchunk, ke = min(K, kb + kchunk);
    A += z * sA;
    B += z * sB;
    C += z * sC;
    const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;   // M blocks on x (up to 2^31)
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;

    for (int k0 = kb; k0 < ke; k0 += GBK) {
        // stage: 64 x 32 elements of each operand, 8 per thread
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int e = r * 256 + tid;          // 0 .. 2047
            int mm, kk;
            if (TA) { kk = e / GBM; mm = e % GBM; } else { mm = e / GBK; kk = e % GBK; }
            const int gm = m0 + mm, gk = k0 + kk;
            float v = 0.f;
            if (gm < M && gk < ke) v = TA ? A[static_cast<long long>(gk) * lda + gm]
                                         : A[static_cast<long long>(gm) * lda + gk];
            As[mm][kk] = to_bf16(v);
            int nn, kb;
            if (TB) { nn = e / GBK; kb = e % GBK; } else { kb = e / GBN; nn = e % GBN; }
            const int gn = n0 + nn, gkb = k0 + kb;
            float w = 0.f;
            if (gn < N && gkb < ke) w = TB ? B[static_cast<long long>(gn) * ldb + gkb]
                                          : B[static_cast<long long>(gkb) * ldb + gn];
            Bs[nn][kb] = to_bf16(w);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GBK; ks += 8) {
            // lane l: A row wm + l%32, k = ks + 4*(l/32) .. +3; B column wn + l%32
            const int r = lane & 31, kq = ks + 4 * (lane >> 5);
            bf16x4 a, b;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = static_cast<short>(As[wm + r][kq + i]);
                b[i] = static_cast<short>(Bs[wn + r][kq + i]);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D[row][col]: row = 8*(j/4) + 4*(l/32) + j%4, col = l%32
    const int col = n0 + wn + (lane & 31);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int row = m0 + wm + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        if (row < M && col < N) {
            float* c = C + static_cast<long long>(row) * ldc + col;
            const float v = alpha * acc[j];
            if (splits > 1) atomicAdd(c, v);        // C was scaled by beta first
            else *c = (beta == 0.f) ? v : fmaf(beta, *c, v);
        }
    }
}

// C[b][m][n] *= beta (0 writes zeros), ahead of a split-K product
__global__ void k_scale(float* __restrict__ C, int M, int N, int ldc, long long sC, float beta) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i >= static_cast<long long>(M) * N) return;
    float* c = C + blockIdx.y * sC + (i / N) * ldc + (i % N);
    *c = (beta == 0.f) ? 0.f : beta * *c;
}

__device__ __forceinline__ float silu(float x) { return x / (1.f + __expf(-x)); }
__device__ __forceinline__ float silu_grad(float x) {
    const float s = 1.f / (1.f + __expf(-x));
    return s * (1.f + x * (1.f - s));
}

// Z[b][rows][cols] += bias[b][col % bias_period]; Y = act(Z) (act 0: none, 1: SiLU)
__global__ void k_bias_act(float* __restrict__ Z, float* __restrict__ Y,
                           const float* __restrict__ bias, long long rows, int cols,
                           int bias_period, long long sZ, long long sBias, int act) {
    const int b = blockIdx.y;
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i >= rows * cols) return;
    const int c = static_cast<int>(i % cols);
    float z = Z[b * sZ + i] + bias[b * sBias + (c % bias_period)];
    Z[b * sZ + i] = z;
    Y[b * sZ + i] = act ? silu(z) : z;
}

// dZ = dY * act'(Z)
__global__ void k_act_bwd(const float* __restrict__ Z, const float* __restrict__ dY,
                          float* __restrict__ dZ, long long n, int act) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    dZ[i] = act ? dY[i] * silu_grad(Z[i]) : dY[i];
}

// dbias[b][c % period] = sum over rows of dZ[b][row][c] (one block per (c, b))
__global__ void k_bias_grad(const float* __restrict__ dZ, float* __restrict__ db, long long rows,
                            int cols, int period, long long sZ, long long sB) {
    __shared__ float red[256];
    const int p = blockIdx.x, b = blockIdx.y;
    float s = 0.f;
    const int reps = cols / period;
    for (long long r = threadIdx.x; r < rows * reps; r += blockDim.x) {
        const long long row = r / reps;
        const int c = static_cast<int>(r % reps) * period + p;
        s += dZ[b * sZ + row * cols + c];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[b * sB + p] = red[0];
}

// 2x2 stride-2 transposed-convolution shuffle (channels last):
//   G[b][(n*H + h)*W + w][(dy*2 + dx)*Co + c]  <->  Y[b][n][2h+dy][2w+dx][c]
// forward (dir 0) G -> Y, backward (dir 1) Y -> G.
__global__ void k_shuffle2(float* __restrict__ G, float* __restrict__ Y, int nimg, int H, int W,
                           int Co, long long sG, int dir) {
    const int b = blockIdx.y;
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    const long long total = static_cast<long long>(nimg) * H * W * 4 * Co;
    if (i >= total) return;
    const int c = static_cast<int>(i % Co);
    long long t = i / Co;
    const int q = static_cast<int>(t % 4);
    t /= 4;
    const int w = static_cast<int>(t % W);
    t /= W;
    const int h = static_cast<int>(t % H);
    const long long n = t / H;
    const int dy = q >> 1, dx = q & 1;
    const long long y = ((n * (2 * H) + (2 * h + dy)) * (2 * W) + (2 * w + dx)) * Co + c;
    if (dir == 0) Y[b * sG + y] = G[b * sG + i];
    else G[b * sG + i] = Y[b * sG + y];
}

// Mean-squared error over each member's outputs; dP = 2 (P - T) / n.
// Weighted mean squared error of member b: loss[b] += mean_i w(i) (P_b - T)^2,
// w(i) = row_w[i / row_len] (1 without row weights).
__global__ void k_mse(const float* __restrict__ P, const float* __restrict__ T,
                      float* __restrict__ dP, float* __restrict__ loss, long long n,
                      long long sP, const float* __restrict__ row_w, long long row_len) {
    __shared__ float red[256];
    const int b = blockIdx.y;
    float s = 0.f;
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const float w = row_w ? row_w[i / row_len] : 1.f;
        const float d = P[b * sP + i] - T[i];
        dP[b * sP + i] = 2.f * w * d / static_cast<float>(n);
        s += w * d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss + b, red[0] / static_cast<float>(n));
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, float wd) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * ((mi / bc1) / (sqrtf(vi / bc2) + eps) + wd * w[i]);
}

// Ensemble statistics per frame: mean over members and, per frame, the mean
// over pixels of the members' standard deviation.
__global__ void k_ens_stats(const float* __restrict__ P, float* __restrict__ mean,
                            float* __restrict__ spread, int members, int frames, int npix,
                            long long sP) {
    __shared__ float red[256];
    const int f = blockIdx.x;
    float s = 0.f;
    for (int p = threadIdx.x; p < npix; p += blockDim.x) {
        float a = 0.f, q = 0.f;
        for (int e = 0; e < members; ++e) {
            const float x = P[e * sP + static_cast<long long>(f) * npix + p];
            a += x;
            q += x * x;
        }
        const float mu = a / members;
        mean[static_cast<long long>(f) * npix + p] = mu;
        s += fmaxf(q / members - mu * mu, 0.f);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) spread[f] = sqrtf(red[0] / npix);
}

int grid1(long long n) { return static_cast<int>((n + 255) / 256); }

}  // namespace
}  // namespace mosaic

using namespace mosaic;

extern "C" int mosaic_sg_gemm(int transA, int transB, int M, int N, int K, const float* A,
                              int lda, long long strideA, const float* B, int ldb,
                              long long strideB, float* C, int ldc, long long strideC,
                              int batch, float alpha, float beta, mosaic_stream_t stream) {
    if (M < 0 || N < 0 || K < 0 || batch < 1 || !A || !B || !C) return MOSAIC_ERR_ARGS;
    if (M == 0 || N == 0) return MOSAIC_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (K == 0) {   // an empty product: C = beta C
        const long long mn = static_cast<long long>(M) * N;
        if (mn == 0) return MOSAIC_OK;
        hipLaunchKernelGGL(k_scale, dim3(static_cast<unsigned>((mn + 255) / 256), batch), dim3(256),
                           0, st, C, M, N, ldc, strideC, beta);
        MOSAIC_HIP_CHECK(hipGetLastError());
        return MOSAIC_OK;
    }
    // Few output tiles over a long K (weight gradients: K = batch x pixels):
    // split K across workgroups so the chip fills, adding partials atomically.
    const long long tiles = static_cast<long long>((N + GBN - 1) / GBN) * ((M + GBM - 1) / GBM);
    int splits = 1;
    if (tiles * batch < 1024 && K > 4 * GBK) {
        const long long want = (2048 + tiles * batch - 1) / (tiles * batch);
        splits = static_cast<int>(std::min<long long>(want, (K + 4 * GBK - 1) / (4 * GBK)));
        splits = std::max(1, std::min(splits, 65535 / batch));
    }
    const int kchunk = ((K + splits - 1) / splits + GBK - 1) / GBK * GBK;
    splits = (K + kchunk - 1) / kchunk;
    dim3 grid((M + GBM - 1) / GBM, (N + GBN - 1) / GBN, batch * splits);
    if (grid.y > 65535 || grid.z > 65535) return MOSAIC_ERR_ARGS;
    if (splits > 1) {
        const long long mn = static_cast<long long>(M) * N;
        hipLaunchKernelGGL(k_scale, dim3(static_cast<unsigned>((mn + 255) / 256), batch), dim3(256),
                           0, st, C, M, N, ldc, strideC, beta);
    }
#define MOSAIC_GEMM(ta, tb)                                                                   \
    hipLaunchKernelGGL((k_gemm_bf16<ta, tb>), grid, dim3(256), 0, st, A, B, C, M, N, K, lda, \
                       ldb, ldc, strideA, strideB, strideC, alpha, beta, splits, kchunk)
    if (!transA && !transB) MOSAIC_GEMM(0, 0);
    else if (!transA && transB) MOSAIC_GEMM(0, 1);
    else if (transA && !transB) MOSAIC_GEMM(1, 0);
    else MOSAIC_GEMM(1, 1);
#undef MOSAIC_GEMM
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_sg_bias_act(float* Z, float* Y, const float* bias, long long rows,
                                  int cols, int bias_period, long long strideZ,
                                  long long strideBias, int batch, int act,
                                  mosaic_stream_t stream) {
    if (!Z || !Y || !bias || rows < 0 || cols < 1 || bias_period < 1 || cols % bias_period)
        return MOSAIC_ERR_ARGS;
    hipLaunchKernelGGL(k_bias_act, dim3(grid1(rows * cols), batch), dim3(256), 0,
                       static_cast<hipStream_t>(stream), Z, Y, bias, rows, cols, bias_period,
                       strideZ, strideBias, act);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_sg_act_bwd(const float* Z, const float* dY, float* dZ, long long n,
                                 int act, mosaic_stream_t stream) {
    if (!Z || !dY || !dZ || n < 0) return MOSAIC_ERR_ARGS;
    hipLaunchKernelGGL(k_act_bwd, dim3(grid1(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       Z, dY, dZ, n, act);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_sg_bias_grad(const float* dZ, float* dbias, long long rows, int cols,
                                   int period, long long strideZ, long long strideBias,
                                   int batch, mosaic_stream_t stream) {
    if (!dZ || !dbias || cols < 1 || period < 1 || cols % period) return MOSAIC_ERR_ARGS;
    hipLaunchKernelGGL(k_bias_grad, dim3(period, batch), dim3(256), 0,
                       static_cast<hipStream_t>(stream), dZ, dbias, rows, cols, period, strideZ,
                       strideBias);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_sg_shuffle2(float* G, float* Y, int nimg, int H, int W, int Co,
                                  long long stride, int batch, int backward,
                                  mosaic_stream_t stream) {
    if (!G || !Y || nimg < 0 || H < 1 || W < 1 || Co < 1) return MOSAIC_ERR_ARGS;
    const long long total = static_cast<long long>(nimg) * H * W * 4 * Co;
    hipLaunchKernelGGL(k_shuffle2, dim3(grid1(total), batch), dim3(256), 0,
                       static_cast<hipStream_t>(stream), G, Y, nimg, H, W, Co, stride,
                       backward ? 1 : 0);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_sg_mse(const float* P, const float* T, float* dP, float* loss, long long n,
                             long long strideP, int batch, const float* row_weight,
                             long long row_len, mosaic_stream_t stream) {
    if (!P || !T || !dP || !loss || n < 1) return MOSAIC_ERR_ARGS;
    if (row_weight && (row_len < 1 || n % row_len)) return MOSAIC_ERR_ARGS;
    hipLaunchKernelGGL(k_mse, dim3(std::min<long long>(1024, grid1(n)), batch), dim3(256), 0,
                       static_cast<hipStream_t>(stream), P, T, dP, loss, n, strideP, row_weight,
                       row_len);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_sg_adam(float* w, const float* g, float* m, float* v, long long n,
                              float lr, float beta1, float beta2, float eps, int step,
                              float weight_decay, mosaic_stream_t stream) {
    if (!w || !g || !m || !v || n < 0 || step < 1) return MOSAIC_ERR_ARGS;
    const float bc1 = 1.f - std::pow(beta1, static_cast<float>(step));
    const float bc2 = 1.f - std::pow(beta2, static_cast<float>(step));
    hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(256), 0, static_cast<hipStream_t>(stream), w,
                       g, m, v, n, lr, beta1, beta2, eps, bc1, bc2, weight_decay);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

extern "C" int mosaic_sg_ensemble_stats(const float* P, float* mean, float* spread, int members,
                                        int frames, int npix, long long strideP,
                                        mosaic_stream_t stream) {
    if (!P || !mean || !spread || members < 1 || frames < 1 || npix < 1) return MOSAIC_ERR_ARGS;
    hipLaunchKernelGGL(k_ens_stats, dim3(frames), dim3(256), 0, static_cast<hipStream_t>(stream),
                       P, mean, spread, members, frames, npix, strideP);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}
