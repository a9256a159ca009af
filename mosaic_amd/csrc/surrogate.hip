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

// dbias[b][c % period] = sum over rows of dZ[b][row][c]: block (p, b, z) sums
// its share of the (row, repeat) elements of bias p and adds it atomically
// (dbias cleared first by k_fill).
__global__ void k_bias_grad(const float* __restrict__ dZ, float* __restrict__ db, long long rows,
                            int cols, int period, long long sZ, long long sB, long long per_block) {
    __shared__ float red[256];
    const int p = blockIdx.x, b = blockIdx.y;
    const int reps = cols / period;
    const long long total = rows * reps;
    const long long r0 = blockIdx.z * per_block, r1 = min(total, r0 + per_block);
    float s = 0.f;
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
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
    if (threadIdx.x == 0) atomicAdd(db + b * sB + p, red[0]);
}

__global__ void k_fill(float* __restrict__ x, long long n, float v) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < n) x[i] = v;
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
// zero-padded to multiples of 16 and 32, so that a wave's operand for one
// instruction is one contiguous 1 KB load.
//
// The inference kernels form each layer transposed, Z^T = W^T X^T: the
// packed weights are the A operand (row n = output channel) and the
// activations, rows r with channels along k, the B operand.  A lane then
// holds, for activation row l%32, the output channels 8a + 4(l/32) + 0..3
// (a = 0..3): four runs of four consecutive channels, each stored to LDS as
// one 8-byte write of four bf16.
constexpr int SG_FB = 32;           // frames per inference block
constexpr int SG_IN = 11;           // network inputs
constexpr int SG_H1 = 256, SG_H2 = 512, SG_H3 = 4096;

__device__ __forceinline__ bf16x8 ld_frag(const unsigned short* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}
__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
// four floats -> four bf16 (round to nearest even), in memory order
__device__ __forceinline__ uint2 bf16x4_bits(float a, float b, float c, float d) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 lo = {static_cast<__bf16>(a), static_cast<__bf16>(b)};
    const bf16x2 hi = {static_cast<__bf16>(c), static_cast<__bf16>(d)};
    return make_uint2(__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi));
}
// SiLU with the hardware reciprocal (1 ulp), whose result is rounded to bf16
__device__ __forceinline__ float silu_rcp(float x) {
    return x * __builtin_amdgcn_rcpf(1.f + __expf(-x));
}
// bias + SiLU of accumulator run a (elements 4a .. 4a+3), channels c0 .. c0+3
// (bias in LDS)
__device__ __forceinline__ uint2 act_run(const f32x16& acc, int a, const float* bias) {
    const float4 b = *reinterpret_cast<const float4*>(bias);
    return bf16x4_bits(silu_rcp(acc[4 * a] + b.x), silu_rcp(acc[4 * a + 1] + b.y),
                       silu_rcp(acc[4 * a + 2] + b.z), silu_rcp(acc[4 * a + 3] + b.w));
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

// optional output map y -> c exp(a y + b) + d (the log-intensity target's inverse)
struct SgPost {
    float a, b, c, d;
    int on;
};

// Layers 1-2 (11 -> 256 -> 512, SiLU) of member blockIdx.z for frames
// 32*blockIdx.x .. +31; blockIdx.y picks 128 of layer 2's 512 columns.  The
// output goes to h2 in B-operand order, [E][blocks][32 k-steps][64][8].
__global__ __launch_bounds__(256) void k_sg_mlp12(const float* __restrict__ X, int B, SgNet net,
                                                  unsigned short* __restrict__ h2, int nfb) {
    __shared__ __attribute__((aligned(16))) unsigned short X1s[SG_FB][SG_H1 + 8];
    __shared__ __attribute__((aligned(16))) float bs1[SG_H1], bs2[128];
    const int fb = blockIdx.x, split = blockIdx.y, e = blockIdx.z;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    bs1[threadIdx.x] = net.b1[e * SG_H1 + threadIdx.x];
    if (threadIdx.x < 128) bs2[threadIdx.x] = net.b2[e * SG_H2 + split * 128 + threadIdx.x];
    __syncthreads();
    // layer 2's weights first, in flight while layer 1 runs
    const int nb2 = split * 4 + wave;
    const unsigned short* W2 = net.W2 + static_cast<long long>(e) * SG_H1 * SG_H2 + nb2 * 16 * 512;
    bf16x8 w2[SG_H1 / 16];
#pragma unroll
    for (int ks = 0; ks < SG_H1 / 16; ++ks) w2[ks] = ld_frag(W2 + ks * 512 + lane * 8);
    const unsigned short* W1 = net.W1 + static_cast<long long>(e) * 16 * SG_H1;
    bf16x8 w1[2];
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2) w1[c2] = ld_frag(W1 + (wave * 2 + c2) * 512 + lane * 8);
    u16x8 xa;
    {
        const int f = fb * SG_FB + r;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = 8 * h + i;
            xa[i] = to_bf16((f < B && k < SG_IN) ? X[static_cast<long long>(f) * SG_IN + k] : 0.f);
        }
    }
    const bf16x8 xf = __builtin_bit_cast(bf16x8, xa);
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2) {
        const int nb = wave * 2 + c2;
        f32x16 acc = {};
        acc = mfma16(w1[c2], xf, acc);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int c0 = nb * 32 + 8 * a + 4 * h;
            *reinterpret_cast<uint2*>(&X1s[r][c0]) = act_run(acc, a, bs1 + c0);
        }
    }
    __syncthreads();
    f32x16 acc = {};
#pragma unroll
    for (int ks = 0; ks < SG_H1 / 16; ++ks)
        acc = mfma16(w2[ks], ld_frag(&X1s[r][ks * 16 + 8 * h]), acc);
    unsigned short* dst = h2 + static_cast<long long>(e * nfb + fb) * (SG_H2 / 16) * 512;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int n = nb2 * 32 + 8 * a + 4 * h;
        const int ks2 = n >> 4, half = (n >> 3) & 1, i0 = n & 7;
        *reinterpret_cast<uint2*>(dst + (ks2 * 64 + r + 32 * half) * 8 + i0) =
            act_run(acc, a, bs2 + (n - split * 128));
    }
}

// Layer 3 for 8x8-level pixel q = blockIdx.x (its 64 channels) and the
// decoder tree under it, for frames 32*blockIdx.y .. +31 of member
// blockIdx.z.  s = 2 dy + dx is the sub-pixel a transposed convolution adds;
// the output pixel is y = 16 (q / 8) + 8 dy1 + 4 dy2 + 2 dy3 + dy4, x alike.
// Wave w computes stage 1's channels of s1 = w for all 32 frames, and from
// there on works alone on its s1: per s2 one 32-row block of stage 2, its
// stage 3 (two s3 at a time) and stage 4.  Stage rows are frames (row = lane
// % 32), so neighbouring lanes store to neighbouring rows (2-way LDS bank
// conflicts at most).  The per-wave buffers keep the workgroup at 46 KB of
// LDS: three workgroups per CU.
constexpr int SG_LD72 = 72, SG_LD40 = 40, SG_LD24 = 24;
constexpr int SG_X1_BYTES = 32 * SG_LD72 * 2;                                 // 4608
constexpr int SG_X2W = 32 * SG_LD72 * 2, SG_X3W = 32 * SG_LD40 * 2;           // 4608, 2560
constexpr int SG_X4W = 64 * SG_LD24 * 2;                                      // 3072
constexpr int SG_WAVE_BYTES = SG_X2W + SG_X3W + SG_X4W;                       // 10240
static_assert(2 * 16 * 64 * 4 <= 4 * SG_WAVE_BYTES, "the layer-3 reduction fits");

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_sg_decode(const unsigned short* __restrict__ h2, int nfb,
                                                   int B, SgNet net, SgPost post,
                                                   float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[SG_X1_BYTES + 4 * SG_WAVE_BYTES];
    // biases of this pixel's layer-3 channels and of the decoder stages
    __shared__ __attribute__((aligned(16))) float bs3[64], bsd1[64], bsd2[32], bsd3[16];
    const int q = blockIdx.x, fb = blockIdx.y, e = blockIdx.z;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    {
        const int t = threadIdx.x;
        if (t < 64) bs3[t] = net.b3[e * SG_H3 + 64 * q + t];
        else if (t < 128) bsd1[t - 64] = net.d1[e * 64 + t - 64];
        else if (t < 160) bsd2[t - 128] = net.d2[e * 32 + t - 128];
        else if (t < 176) bsd3[t - 160] = net.d3[e * 16 + t - 160];
    }
    auto X1 = reinterpret_cast<unsigned short(*)[SG_LD72]>(smem);
    unsigned char* wbuf = smem + SG_X1_BYTES + wave * SG_WAVE_BYTES;
    auto X2 = reinterpret_cast<unsigned short(*)[SG_LD72]>(wbuf);
    auto X3 = reinterpret_cast<unsigned short(*)[SG_LD40]>(wbuf + SG_X2W);
    auto X4 = reinterpret_cast<unsigned short(*)[SG_LD24]>(wbuf + SG_X2W + SG_X3W);
    float* red = reinterpret_cast<float*>(smem + SG_X1_BYTES);   // before stage 1 only

    // layer 3: columns 64 q .. +63 (two blocks of 32), K = 512 split over wave
    // pairs; every operand load issued before the first product
    const int cb3 = wave & 1, kh = wave >> 1;
    {
        bf16x8 la[16], lw[16];
        const unsigned short* A = h2 + static_cast<long long>(e * nfb + fb) * 32 * 512 +
                                  kh * 16 * 512 + lane * 8;
        const unsigned short* Bw = net.W3 + static_cast<long long>(e) * SG_H2 * SG_H3 +
                                   static_cast<long long>(2 * q + cb3) * 32 * 512 +
                                   kh * 16 * 512 + lane * 8;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            lw[s] = ld_frag(Bw + s * 512);
            la[s] = ld_frag(A + s * 512);
        }
        __builtin_amdgcn_sched_barrier(0);
        f32x16 acc = {};
#pragma unroll
        for (int s = 0; s < 16; ++s) acc = mfma16(lw[s], la[s], acc);
        if (kh == 1)
#pragma unroll
            for (int j = 0; j < 16; ++j) red[(cb3 * 16 + j) * 64 + lane] = acc[j];
        __syncthreads();
        if (kh == 0) {
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[j] += red[(cb3 * 16 + j) * 64 + lane];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int c0 = cb3 * 32 + 8 * a + 4 * h;
                *reinterpret_cast<uint2*>(&X1[r][c0]) = act_run(acc, a, bs3 + c0);
            }
        }
    }
    // the decoder's weights, in flight through the layer-3 epilogue
    bf16x8 w1[2][4], w2[4][4], w3[2][2];
#pragma unroll
    for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            w1[c2][ks] = ld_frag(net.D1 + static_cast<long long>(e) * 64 * 256 +
                                 ((wave * 2 + c2) * 4 + ks) * 512 + lane * 8);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            w2[nb][ks] = ld_frag(net.D2 + static_cast<long long>(e) * 64 * 128 +
                                 (nb * 4 + ks) * 512 + lane * 8);
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            w3[cb][ks] = ld_frag(net.D3 + static_cast<long long>(e) * 32 * 64 +
                                 (cb * 2 + ks) * 512 + lane * 8);
    const bf16x8 w4 = ld_frag(net.D4 + static_cast<long long>(e) * 16 * 32 + lane * 8);
    const float b4 = net.d4[e];
    __syncthreads();
    // stage 1: 64 -> 4 x 64; wave w computes channels 64 w .. +63, i.e. s1 = w
    {
        bf16x8 x[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) x[ks] = ld_frag(&X1[r][ks * 16 + 8 * h]);
#pragma unroll
        for (int c2 = 0; c2 < 2; ++c2) {
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) acc = mfma16(w1[c2][ks], x[ks], acc);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int c0 = c2 * 32 + 8 * a + 4 * h;
                *reinterpret_cast<uint2*>(&X2[r][c0]) = act_run(acc, a, bsd1 + c0);
            }
        }
    }
    __syncthreads();
    bf16x8 x2[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) x2[ks] = ld_frag(&X2[r][ks * 16 + 8 * h]);
    const int f = fb * SG_FB + r;
    float* orow = out + (static_cast<long long>(e) * B + (f < B ? f : 0)) * 16384;
    const int s1 = wave;
    for (int s2 = 0; s2 < 4; ++s2) {
        // stage 2 block s2: 64 -> 32 channels
        {
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) acc = mfma16(w2[s2][ks], x2[ks], acc);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int c0 = 8 * a + 4 * h;
                *reinterpret_cast<uint2*>(&X3[r][c0]) = act_run(acc, a, bsd2 + c0);
            }
        }
        __syncthreads();
        bf16x8 x3[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) x3[ks] = ld_frag(&X3[r][ks * 16 + 8 * h]);
        float4 o[4];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            // stage 3 block cb: s3 = 2 cb + a / 2, channel 8 (a % 2) + 4 h + 0..3
            f32x16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) acc = mfma16(w3[cb][ks], x3[ks], acc);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int c0 = 8 * (a & 1) + 4 * h;
                *reinterpret_cast<uint2*>(&X4[(a >> 1) * 32 + r][c0]) =
                    act_run(acc, a, bsd3 + c0);
            }
            __syncthreads();
            // stage 4: rows n = s4 = 0..3 are elements 0..3 of lanes 0..31
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int s3 = 2 * cb + sl;
                f32x16 acc4 = {};
                acc4 = mfma16(w4, ld_frag(&X4[sl * 32 + r][8 * h]), acc4);
                const int yi = 2 * (s3 >> 1), xi = 2 * (s3 & 1);
                reinterpret_cast<float*>(&o[yi])[xi] = acc4[0] + b4;
                reinterpret_cast<float*>(&o[yi])[xi + 1] = acc4[1] + b4;
                reinterpret_cast<float*>(&o[yi + 1])[xi] = acc4[2] + b4;
                reinterpret_cast<float*>(&o[yi + 1])[xi + 1] = acc4[3] + b4;
            }
            __syncthreads();      // X4 is rewritten by the next block
        }
        // the frame's 4 x 4 block under (s1, s2)
        if (h == 0 && f < B) {
            const int yb = 16 * (q >> 3) + 8 * (s1 >> 1) + 4 * (s2 >> 1);
            const int xb = 16 * (q & 7) + 8 * (s1 & 1) + 4 * (s2 & 1);
#pragma unroll
            for (int yi = 0; yi < 4; ++yi) {
                float4 v = o[yi];
                if (post.on) {
                    v.x = fmaf(post.c, __expf(fmaf(post.a, v.x, post.b)), post.d);
                    v.y = fmaf(post.c, __expf(fmaf(post.a, v.y, post.b)), post.d);
                    v.z = fmaf(post.c, __expf(fmaf(post.a, v.z, post.b)), post.d);
                    v.w = fmaf(post.c, __expf(fmaf(post.a, v.w, post.b)), post.d);
                }
                *reinterpret_cast<float4*>(orow + (yb + yi) * 128 + xb) = v;
            }
        }
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
    if (!dZ || !dbias || cols < 1 || period < 1 || cols % period || rows < 0 || batch < 1)
        return MOSAIC_ERR_ARGS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // dbias rows: batch blocks of strideBias, the first `period` of each used
    const long long nb = static_cast<long long>(batch - 1) * strideBias + period;
    hipLaunchKernelGGL(k_fill, dim3(grid1(nb)), dim3(256), 0, st, dbias, nb, 0.f);
    // about 1024 workgroups, each over at least 4096 elements of its bias
    const long long total = rows * (cols / period);
    if (total == 0) {
        MOSAIC_HIP_CHECK(hipGetLastError());
        return MOSAIC_OK;
    }
    long long splits = std::max(1LL, std::min<long long>((total + 4095) / 4096,
                                                        1024LL / (static_cast<long long>(period) *
                                                                  batch)));
    splits = std::min<long long>(splits, 65535);
    const long long per_block = (total + splits - 1) / splits;
    splits = std::max(1LL, (total + per_block - 1) / per_block);
    hipLaunchKernelGGL(k_bias_grad, dim3(period, batch, static_cast<unsigned>(splits)), dim3(256),
                       0, st, dZ, dbias, rows, cols, period, strideZ, strideBias, per_block);
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
                               const float* post, void* workspace, size_t workspace_bytes,
                               float* out, mosaic_stream_t stream) {
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
    SgPost pm = {0.f, 0.f, 0.f, 0.f, 0};
    if (post) pm = {post[0], post[1], post[2], post[3], 1};
    hipLaunchKernelGGL(k_sg_decode, dim3(64, nfb, members), dim3(256), 0, st, h2, nfb, batch, n,
                       pm, out);
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}
