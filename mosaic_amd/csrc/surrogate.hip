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
// quadrant held in one v_mfma_f32_32x32x16_bf16 accumulator (16 FP32 per
// lane); K advances 32 per LDS stage (2 MFMAs).  grid.z indexes the batch
// (ensemble member) with element strides sA, sB, sC.
//
// Inference (mosaic_sg_infer) runs the whole ensemble forward in two
// launches on weights packed once into the matrix cores' operand order
// (mosaic_sg_pack): the MLP's first two layers, then one workgroup per
// (member, 8x8-level pixel, 32 frames) for the 512 -> 4096 layer's 64
// columns of that pixel and the whole transposed-convolution tree below it.
// Each 2x2 stride-2 transposed convolution maps one input pixel to its own
// 2x2 block, so an 8x8-level pixel's 16x16 block of the 128x128 output
// depends on that pixel's 64 channels alone, and every intermediate stays in
// LDS.

#include "common.h"

namespace mosaic {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
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
    __shared__ __attribute__((aligned(16))) unsigned short As[GBM][GBK + 8];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[GBN][GBK + 8];
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
        for (int ks = 0; ks < GBK; ks += 16) {
            // lane l: A row wm + l%32, k = ks + 8*(l/32) .. +7; B column wn + l%32
            const int r = lane & 31, kq = ks + 8 * (lane >> 5);
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + r][kq]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + r][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
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

// ------------------------------------------------------------ inference
// Operand order of v_mfma_f32_32x32x16_bf16: lane l holds A row l%32 (B
// column l%32) at k = 8*(l/32) .. +7; accumulator element j of lane l is
// row 8*(j/4) + 4*(l/32) + j%4, column l%32.  A packed weight matrix W
// (K x N) is stored per member as [N/32][K/16][64 lanes][8] bf16, K and N
// zero-padded to multiples of 16 and 32, so that a wave's B operand for one
// instruction is one contiguous 1 KB load.
constexpr int SG_FB = 32;           // frames per inference block
constexpr int SG_IN = 11;           // network inputs
constexpr int SG_H1 = 256, SG_H2 = 512, SG_H3 = 4096;

__device__ __forceinline__ bf16x8 ld_frag(const unsigned short* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}
__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int acc_row(int j, int lane) {
    return 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
}

__global__ void k_sg_pack(const float* __restrict__ W, int K, int N, int Kp, int Np,
                          long long sW, unsigned short* __restrict__ out) {
    const long long per = static_cast<long long>(Kp) * Np;
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    const int e = blockIdx.y;
    if (i >= per) return;
    const int ii = static_cast<int>(i & 7), lane = static_cast<int>((i >> 3) & 63);
    const long long t = i >> 9;
    const int nks = Kp / 16;
    const int ks = static_cast<int>(t % nks), nb = static_cast<int>(t / nks);
    const int k = ks * 16 + 8 * (lane >> 5) + ii, n = nb * 32 + (lane & 31);
    const float v = (k < K && n < N) ? W[e * sW + static_cast<long long>(k) * N + n] : 0.f;
    out[e * per + i] = to_bf16(v);
}

struct SgNet {
    const unsigned short *W1, *W2, *W3, *D1, *D2, *D3, *D4;   // packed, per member
    const float *b1, *b2, *b3, *d1, *d2, *d3, *d4;            // biases (E, n)
};

// Layers 1-2 (11 -> 256 -> 512, SiLU) of members blockIdx.z for frames
// 32*blockIdx.x .. +31; blockIdx.y picks 128 of layer 2's 512 columns.  The
// output goes to h2 in A-operand order, [E][blocks][32 k-steps][64][8].
__global__ __launch_bounds__(256) void k_sg_mlp12(const float* __restrict__ X, int B, SgNet net,
                                                  unsigned short* __restrict__ h2, int nfb) {
    __shared__ __attribute__((aligned(16))) unsigned short X1s[SG_FB][SG_H1 + 8];
    const int fb = blockIdx.x, split = blockIdx.y, e = blockIdx.z;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    u16x8 xa;
    {
        const int f = fb * SG_FB + (lane & 31);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = 8 * (lane >> 5) + i;
            xa[i] = to_bf16((f < B && k < SG_IN) ? X[static_cast<long long>(f) * SG_IN + k] : 0.f);
        }
    }
    const bf16x8 xf = __builtin_bit_cast(bf16x8, xa);
    const unsigned short* W1 = net.W1 + static_cast<long long>(e) * 16 * SG_H1;
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2) {
        const int nb = wave * 2 + c2;
        f32x16 acc = {};
        acc = mfma16(xf, ld_frag(W1 + nb * 512 + lane * 8), acc);
        const int n = nb * 32 + (lane & 31);
        const float bias = net.b1[e * SG_H1 + n];
#pragma unroll
        for (int j = 0; j < 16; ++j) X1s[acc_row(j, lane)][n] = to_bf16(silu(acc[j] + bias));
    }
    __syncthreads();
    const int nb = split * 4 + wave;
    const unsigned short* W2 = net.W2 + static_cast<long long>(e) * SG_H1 * SG_H2 + nb * 16 * 512;
    f32x16 acc = {};
#pragma unroll
    for (int ks = 0; ks < SG_H1 / 16; ++ks) {
        const bf16x8 a = ld_frag(&X1s[lane & 31][ks * 16 + 8 * (lane >> 5)]);
        acc = mfma16(a, ld_frag(W2 + ks * 512 + lane * 8), acc);
    }
    const int n = nb * 32 + (lane & 31);
    const float bias = net.b2[e * SG_H2 + n];
    unsigned short* dst = h2 + static_cast<long long>(e * nfb + fb) * (SG_H2 / 16) * 512;
    const int ks2 = n >> 4, half = (n >> 3) & 1, ii = n & 7;
#pragma unroll
    for (int j = 0; j < 16; ++j)
        dst[(ks2 * 64 + acc_row(j, lane) + 32 * half) * 8 + ii] = to_bf16(silu(acc[j] + bias));
}

// Layer 3 for 8x8-level pixel q = blockIdx.x (its 64 channels) and the
// decoder tree under it, for frames 32*blockIdx.y .. +31 of member
// blockIdx.z.  Stage rows: stage 1 (f), 2 (f, s1), 3 (f, s1, s2), 4 (f, s1,
// s2, s3), with s = 2 dy + dx the sub-pixel a transposed convolution adds;
// the output pixel is y = 16 (q / 8) + 8 dy1 + 4 dy2 + 2 dy3 + dy4, x alike.
constexpr int SG_X3_LD = 40, SG_X12_LD = 72, SG_X4_LD = 24;
constexpr int SG_X3_BYTES = 512 * SG_X3_LD * 2;                       // 40960
constexpr int SG_X4_BYTES = 128 * SG_X4_LD * 2;                       // 6144 per wave
constexpr int SG_UNION_BYTES = 4 * SG_X4_BYTES;                       // >= X1s + X2s
constexpr int SG_OUT_BYTES = 2 * 256 * 4;                             // 2 frames x 16 x 16
static_assert((32 + 128) * SG_X12_LD * 2 <= SG_UNION_BYTES, "X1s + X2s fit the union");

__global__ __launch_bounds__(256) void k_sg_decode(const unsigned short* __restrict__ h2, int nfb,
                                                   int B, SgNet net, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16)))
        unsigned char smem[SG_X3_BYTES + SG_UNION_BYTES + 4 * SG_OUT_BYTES];
    const int q = blockIdx.x, fb = blockIdx.y, e = blockIdx.z;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    auto X3 = reinterpret_cast<unsigned short(*)[SG_X3_LD]>(smem);
    auto X1 = reinterpret_cast<unsigned short(*)[SG_X12_LD]>(smem + SG_X3_BYTES);
    auto X2 = reinterpret_cast<unsigned short(*)[SG_X12_LD]>(smem + SG_X3_BYTES +
                                                            32 * SG_X12_LD * 2);
    auto X4 = reinterpret_cast<unsigned short(*)[SG_X4_LD]>(smem + SG_X3_BYTES +
                                                           wave * SG_X4_BYTES);
    float* Ow = reinterpret_cast<float*>(smem + SG_X3_BYTES + SG_UNION_BYTES +
                                         wave * SG_OUT_BYTES);
    float* red = reinterpret_cast<float*>(smem);   // aliases X3 (not yet written)

    // layer 3: columns 64 q .. +63 (two blocks of 32), K = 512 split over wave pairs
    {
        const int cb = wave & 1, kh = wave >> 1;
        const unsigned short* A = h2 + static_cast<long long>(e * nfb + fb) * 32 * 512;
        const unsigned short* Bw = net.W3 + static_cast<long long>(e) * SG_H2 * SG_H3 +
                                   static_cast<long long>(2 * q + cb) * 32 * 512;
        f32x16 acc = {};
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int ks = kh * 16 + s;
            acc = mfma16(ld_frag(A + ks * 512 + lane * 8), ld_frag(Bw + ks * 512 + lane * 8), acc);
        }
        if (kh == 1)
#pragma unroll
            for (int j = 0; j < 16; ++j) red[(cb * 16 + j) * 64 + lane] = acc[j];
        __syncthreads();
        if (kh == 0) {
            const int c = cb * 32 + (lane & 31);
            const float bias = net.b3[e * SG_H3 + 64 * q + c];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                X1[acc_row(j, lane)][c] =
                    to_bf16(silu(acc[j] + red[(cb * 16 + j) * 64 + lane] + bias));
        }
        __syncthreads();
    }
    // stage 1: 64 -> 4 x 64 (rows f)
    {
        bf16x8 a[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a[ks] = ld_frag(&X1[lane & 31][ks * 16 + 8 * (lane >> 5)]);
        const unsigned short* Bw = net.D1 + static_cast<long long>(e) * 64 * 256;
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2) {
            const int nb = wave * 2 + c2;
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                acc = mfma16(a[ks], ld_frag(Bw + (nb * 4 + ks) * 512 + lane * 8), acc);
            const int n = nb * 32 + (lane & 31), s1 = n >> 6, c = n & 63;
            const float bias = net.d1[e * 64 + c];
#pragma unroll
            for (int j = 0; j < 16; ++j) X2[acc_row(j, lane) * 4 + s1][c] = to_bf16(silu(acc[j] + bias));
        }
        __syncthreads();
    }
    // stage 2: 64 -> 4 x 32 (rows (f, s1)); wave = row block
    {
        bf16x8 a[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            a[ks] = ld_frag(&X2[32 * wave + (lane & 31)][ks * 16 + 8 * (lane >> 5)]);
        const unsigned short* Bw = net.D2 + static_cast<long long>(e) * 64 * 128;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                acc = mfma16(a[ks], ld_frag(Bw + (nb * 4 + ks) * 512 + lane * 8), acc);
            const int n = nb * 32 + (lane & 31), s2 = n >> 5, c = n & 31;
            const float bias = net.d2[e * 32 + c];
#pragma unroll
            for (int j = 0; j < 16; ++j)
                X3[(32 * wave + acc_row(j, lane)) * 4 + s2][c] = to_bf16(silu(acc[j] + bias));
        }
        __syncthreads();
    }
    // stages 3 and 4, one block of 32 stage-3 rows (two frames) at a time
    const unsigned short* B3 = net.D3 + static_cast<long long>(e) * 32 * 64;
    bf16x8 w3[2][2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) w3[cb][ks] = ld_frag(B3 + (cb * 2 + ks) * 512 + lane * 8);
    const bf16x8 w4 = ld_frag(net.D4 + static_cast<long long>(e) * 16 * 32 + lane * 8);
    const float b4 = net.d4[e];
    const int Y0 = 16 * (q >> 3), X0 = 16 * (q & 7);
    for (int t = 0; t < 4; ++t) {
        const int rb = wave * 4 + t;
        bf16x8 a[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            a[ks] = ld_frag(&X3[32 * rb + (lane & 31)][ks * 16 + 8 * (lane >> 5)]);
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) acc = mfma16(a[ks], w3[cb][ks], acc);
            const int n = cb * 32 + (lane & 31), s3 = n >> 4, c = n & 15;
            const float bias = net.d3[e * 16 + c];
#pragma unroll
            for (int j = 0; j < 16; ++j) X4[acc_row(j, lane) * 4 + s3][c] = to_bf16(silu(acc[j] + bias));
        }
        __syncthreads();
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
            f32x16 acc = {};
            acc = mfma16(ld_frag(&X4[32 * r4 + (lane & 31)][8 * (lane >> 5)]), w4, acc);
            const int s4 = lane & 31;
            if (s4 < 4) {
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int row4 = 32 * r4 + acc_row(j, lane);   // (r3, s3) in this block
                    const int r3 = row4 >> 2, s3 = row4 & 3;
                    const int R3 = 32 * rb + r3;                   // (f, s1, s2)
                    const int s2 = R3 & 3, s1 = (R3 >> 2) & 3, fl = r3 >> 4;
                    const int yl = 8 * (s1 >> 1) + 4 * (s2 >> 1) + 2 * (s3 >> 1) + (s4 >> 1);
                    const int xl = 8 * (s1 & 1) + 4 * (s2 & 1) + 2 * (s3 & 1) + (s4 & 1);
                    Ow[(fl * 16 + yl) * 16 + xl] = acc[j] + b4;
                }
            }
        }
        __syncthreads();
        // two frames' 16 x 16 blocks: lane -> (frame, row, 8-float half row)
        {
            const int fl = lane >> 5, yl = (lane >> 1) & 15, xh = (lane & 1) * 8;
            const int f = fb * SG_FB + 2 * rb + fl;
            if (f < B) {
                const float4* src = reinterpret_cast<const float4*>(Ow + (fl * 16 + yl) * 16 + xh);
                float4* dst = reinterpret_cast<float4*>(
                    out + (static_cast<long long>(e) * B + f) * 16384 + (Y0 + yl) * 128 + X0 + xh);
                dst[0] = src[0];
                dst[1] = src[1];
            }
        }
        __syncthreads();
    }
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

extern "C" int mosaic_sg_pack(int members, int K, int N, const float* W, long long strideW,
                              void* packed, mosaic_stream_t stream) {
    if (members < 1 || K < 1 || N < 1 || !W || !packed) return MOSAIC_ERR_ARGS;
    const int Kp = (K + 15) / 16 * 16, Np = (N + 31) / 32 * 32;
    const long long per = static_cast<long long>(Kp) * Np;
    hipLaunchKernelGGL(k_sg_pack, dim3(static_cast<unsigned>((per + 255) / 256), members),
                       dim3(256), 0, static_cast<hipStream_t>(stream), W, K, N, Kp, Np, strideW,
                       static_cast<unsigned short*>(packed));
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

static size_t sg_infer_bytes(int members, int batch) {
    const size_t nfb = (static_cast<size_t>(batch) + SG_FB - 1) / SG_FB;
    return static_cast<size_t>(members) * nfb * SG_FB * SG_H2 * sizeof(unsigned short);
}

extern "C" int mosaic_sg_infer_workspace(int members, int batch, size_t* bytes) {
    if (members < 1 || batch < 0 || !bytes) return MOSAIC_ERR_ARGS;
    *bytes = sg_infer_bytes(members, batch);
    return MOSAIC_OK;
}

extern "C" int mosaic_sg_infer(int members, int batch, const float* X, const mosaic_sg_net* net,
                               void* workspace, size_t workspace_bytes, float* out,
                               mosaic_stream_t stream) {
    if (members < 1 || batch < 0 || !net) return MOSAIC_ERR_ARGS;
    if (batch == 0) return MOSAIC_OK;
    if (!X || !out || !workspace) return MOSAIC_ERR_ARGS;
    if (workspace_bytes < sg_infer_bytes(members, batch)) return MOSAIC_ERR_WORKSPACE;
    const void* const ptrs[] = {net->W1, net->b1, net->W2, net->b2, net->W3, net->b3, net->D1,
                                net->d1, net->D2, net->d2, net->D3, net->d3, net->D4, net->d4};
    for (const void* p : ptrs)
        if (!p) return MOSAIC_ERR_ARGS;
    SgNet n;
    n.W1 = static_cast<const unsigned short*>(net->W1);
    n.W2 = static_cast<const unsigned short*>(net->W2);
    n.W3 = static_cast<const unsigned short*>(net->W3);
    n.D1 = static_cast<const unsigned short*>(net->D1);
    n.D2 = static_cast<const unsigned short*>(net->D2);
    n.D3 = static_cast<const unsigned short*>(net->D3);
    n.D4 = static_cast<const unsigned short*>(net->D4);
    n.b1 = net->b1;
    n.b2 = net->b2;
    n.b3 = net->b3;
    n.d1 = net->d1;
    n.d2 = net->d2;
    n.d3 = net->d3;
    n.d4 = net->d4;
    const int nfb = (batch + SG_FB - 1) / SG_FB;
    if (nfb > 65535 || members > 65535) return MOSAIC_ERR_ARGS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned short* h2 = static_cast<unsigned short*>(workspace);
    hipLaunchKernelGGL(k_sg_mlp12, dim3(nfb, SG_H2 / 128, members), dim3(256), 0, st, X, batch, n,
                       h2, nfb);
    hipLaunchKernelGGL(k_sg_decode, dim3(64, nfb, members), dim3(256), 0, st, h2, nfb, batch, n,
                       out);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}
