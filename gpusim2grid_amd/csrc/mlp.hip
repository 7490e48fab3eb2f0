// Learned warm start: the MLP's training and inference kernels (gfx950).
//
// A 3-layer MLP (in -> hidden -> hidden -> out, GELU) maps a row's per-bus
// injections (and outage mask) to its voltage angles and magnitudes.  Every
// product runs on the bf16 matrix cores with fp32 accumulation
// (v_mfma_f32_32x32x16_bf16): one tiled GEMM kernel, templated on the operand
// layouts (each operand row- or column-major, which covers the forward
// X W, the backward dY W^T and the weight gradient X^T dY) and on a fused
// epilogue (bias + GELU with the pre-activation kept for the backward pass,
// bias only, GELU derivative).  Weights are kept in fp32 (Adam state) with a
// bf16 copy the GEMMs read.
//
// The gpusim2grid reference has no learned component; this follows the
// model the reference's paper benchmarks time as "PFdelta" MLPs
// (20260831_neurips_paper/scripts/README.md, ml_injection.py) only in kind.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.hpp"

namespace g2g {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128, kBK = 32;
constexpr int kPadK = kBK + 8;  // LDS row pitch (bf16): 80 bytes, spreads the 16-byte reads over banks

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// A (M x K) and B (K x N) tiles of one K step, staged k-contiguous in LDS:
// sA[m][k], sB[n][k].
template <bool TA, bool TB>
__device__ __forceinline__ void load_tiles(const GemmArgs& g, int m0, int n0, int k0, __bf16 (*sA)[kPadK],
                                           __bf16 (*sB)[kPadK])
{
    const int t = threadIdx.x;  // 256 threads
    const __bf16 zero = static_cast<__bf16>(0.0f);
    // ---- A: 128 x 32 ------------------------------------------------------
    if (!TA) {  // element (m, k) at A[m * lda + k]
        const int m = t >> 1, kh = (t & 1) * 16;
        const int gm = m0 + m, gk = k0 + kh;
        const __bf16* src = g.A + static_cast<long long>(gm) * g.lda + gk;
        if (gm < g.M && gk + 16 <= g.K && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
            const u32x4 v0 = reinterpret_cast<const u32x4*>(src)[0];
            const u32x4 v1 = reinterpret_cast<const u32x4*>(src)[1];
            *reinterpret_cast<u32x4*>(&sA[m][kh]) = v0;
            *reinterpret_cast<u32x4*>(&sA[m][kh + 8]) = v1;
        } else {
            for (int i = 0; i < 16; ++i) sA[m][kh + i] = (gm < g.M && gk + i < g.K) ? src[i] : zero;
        }
    } else {  // element (m, k) at A[k * lda + m]
        const int k = t >> 3, mh = (t & 7) * 16;
        const int gk = k0 + k, gm = m0 + mh;
        const __bf16* src = g.A + static_cast<long long>(gk) * g.lda + gm;
        for (int i = 0; i < 16; ++i) sA[mh + i][k] = (gk < g.K && gm + i < g.M) ? src[i] : zero;
    }
    // ---- B: 32 x 128 ------------------------------------------------------
    if (TB) {  // element (k, n) at B[n * ldb + k]
        const int n = t >> 1, kh = (t & 1) * 16;
        const int gn = n0 + n, gk = k0 + kh;
        const __bf16* src = g.B + static_cast<long long>(gn) * g.ldb + gk;
        if (gn < g.N && gk + 16 <= g.K && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
            const u32x4 v0 = reinterpret_cast<const u32x4*>(src)[0];
            const u32x4 v1 = reinterpret_cast<const u32x4*>(src)[1];
            *reinterpret_cast<u32x4*>(&sB[n][kh]) = v0;
            *reinterpret_cast<u32x4*>(&sB[n][kh + 8]) = v1;
        } else {
            for (int i = 0; i < 16; ++i) sB[n][kh + i] = (gn < g.N && gk + i < g.K) ? src[i] : zero;
        }
    } else {  // element (k, n) at B[k * ldb + n]
        const int k = t >> 3, nh = (t & 7) * 16;
        const int gk = k0 + k, gn = n0 + nh;
        const __bf16* src = g.B + static_cast<long long>(gk) * g.ldb + gn;
        for (int i = 0; i < 16; ++i) sB[nh + i][k] = (gk < g.K && gn + i < g.N) ? src[i] : zero;
    }
}

// C = A B, 128 x 128 output tile per workgroup of 4 waves (each 64 x 64 as
// 2 x 2 MFMA 32 x 32 tiles), K in steps of 32 through LDS.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    __shared__ __align__(16) __bf16 sA[kBM][kPadK];
    __shared__ __align__(16) __bf16 sB[kBN][kPadK];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const int lr = lane & 31, lk = (lane >> 5) * 8;
    for (int k0 = 0; k0 < g.K; k0 += kBK) {
        load_tiles<TA, TB>(g, m0, n0, k0, sA, sB);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 16) {
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[i] = *reinterpret_cast<const bf16x8*>(&sA[wm * 64 + i * 32 + lr][kk + lk]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                fb[j] = *reinterpret_cast<const bf16x8*>(&sB[wn * 64 + j * 32 + lr][kk + lk]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue: lane holds column lane % 32 and rows 8 (r / 4) + 4 (lane / 32) + r % 4
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int n = n0 + wn * 64 + j * 32 + lr;
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[i][j][r];
                const long long mc = static_cast<long long>(m);
                if (EPI == kEpiF32) {
                    g.C[mc * g.ldc + n] = v;
                } else if (EPI == kEpiBiasF32) {
                    g.C[mc * g.ldc + n] = v + g.bias[n];
                } else if (EPI == kEpiBiasGeluBf16) {
                    const float z = v + g.bias[n];
                    g.Z[mc * g.ldz + n] = z;
                    g.H[mc * g.ldh + n] = static_cast<__bf16>(gelu(z));
                } else {  // kEpiGeluBwd: dZ = (dH) * gelu'(Z), fp32 and bf16
                    const float dz = v * gelu_grad(g.Z[mc * g.ldz + n]);
                    g.C[mc * g.ldc + n] = dz;
                    g.H[mc * g.ldh + n] = static_cast<__bf16>(dz);
                }
            }
}

// Forward GEMM with both operands K-contiguous: C = A Bt^T, A: M x K (row
// m at A + m lda), Bt: N x K (row n at B + n ldb), K a multiple of 8 and rows
// 16-byte aligned.  128 x 128 output tile per workgroup of 4 waves (each
// 64 x 64 as 2 x 2 MFMA 32x32x16 tiles); K in steps of 64 through a
// double-buffered LDS stage, the next step's global loads in flight while
// the current one is multiplied.
constexpr int kFK = 64, kFPad = 8;

__device__ __forceinline__ void kc_load(const __bf16* base, long long ld, int rows, int r0, int k0, int K, int t,
                                        u32x4 (&reg)[4])
{
    const int kc = (t & 7) * 8;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + (t >> 3) + 32 * i;
        const int k = k0 + kc;
        if (r < rows && k < K) reg[i] = *reinterpret_cast<const u32x4*>(base + static_cast<long long>(r) * ld + k);
        else reg[i] = u32x4{0u, 0u, 0u, 0u};
    }
}

__device__ __forceinline__ void kc_store(__bf16 (*s)[kFK + kFPad], int t, const u32x4 (&reg)[4])
{
    const int kc = (t & 7) * 8;
#pragma unroll
    for (int i = 0; i < 4; ++i) *reinterpret_cast<u32x4*>(&s[(t >> 3) + 32 * i][kc]) = reg[i];
}

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_kc(GemmArgs g)
{
    __shared__ __align__(16) __bf16 sA[2][kBM][kFK + kFPad];
    __shared__ __align__(16) __bf16 sB[2][kBN][kFK + kFPad];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const int lr = lane & 31, lk = (lane >> 5) * 8;
    u32x4 ra[4], rb[4];
    kc_load(g.A, g.lda, g.M, m0, 0, g.K, t, ra);
    kc_load(g.B, g.ldb, g.N, n0, 0, g.K, t, rb);
    kc_store(sA[0], t, ra);
    kc_store(sB[0], t, rb);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < g.K; k0 += kFK) {
        const bool more = k0 + kFK < g.K;
        if (more) {
            kc_load(g.A, g.lda, g.M, m0, k0 + kFK, g.K, t, ra);
            kc_load(g.B, g.ldb, g.N, n0, k0 + kFK, g.K, t, rb);
        }
#pragma unroll
        for (int kk = 0; kk < kFK; kk += 16) {
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[i] = *reinterpret_cast<const bf16x8*>(&sA[cur][wm * 64 + i * 32 + lr][kk + lk]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                fb[j] = *reinterpret_cast<const bf16x8*>(&sB[cur][wn * 64 + j * 32 + lr][kk + lk]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            kc_store(sA[cur ^ 1], t, ra);
            kc_store(sB[cur ^ 1], t, rb);
        }
        __syncthreads();
        cur ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int n = n0 + wn * 64 + j * 32 + lr;
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[i][j][r];
                const long long mc = static_cast<long long>(m);
                if (EPI == kEpiF32) {
                    g.C[mc * g.ldc + n] = v;
                } else if (EPI == kEpiBiasF32) {
                    g.C[mc * g.ldc + n] = v + g.bias[n];
                } else if (EPI == kEpiBiasGeluBf16) {
                    const float z = v + g.bias[n];
                    if (g.Z) g.Z[mc * g.ldz + n] = z;
                    g.H[mc * g.ldh + n] = static_cast<__bf16>(gelu(z));
                }
            }
}

// Wt (N x K, bf16) = transpose of W (K x N, fp32), 32 x 32 tiles through LDS
__global__ void k_transpose_bf16(const float* __restrict__ W, int K, int N, __bf16* __restrict__ Wt)
{
    __shared__ float tile[32][33];
    const int k0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int k = k0 + r, n = n0 + tx;
        tile[r][tx] = (k < K && n < N) ? W[static_cast<long long>(k) * N + n] : 0.0f;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int n = n0 + r, k = k0 + tx;
        if (n < N && k < K) Wt[static_cast<long long>(n) * K + k] = static_cast<__bf16>(tile[tx][r]);
    }
}

// column sums of an fp32 M x N matrix (bias gradients), one thread per column
// Column sums of X [M][N] accumulated into out (zeroed by the caller): a
// block sums kColRows rows of 64 columns, four row groups per column reduced
// in LDS, then one atomic per column.  Grid: (ceil(N / 64), ceil(M / kColRows)).
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ X, long long ldx, int M, int N,
                                                float* __restrict__ out)
{
    __shared__ float part[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    const int m0 = blockIdx.y * kColRows;
    const int m1 = m0 + kColRows < M ? m0 + kColRows : M;
    float s = 0.0f;
    if (n < N)
        for (int m = m0 + g; m < m1; m += 4) s += X[static_cast<long long>(m) * ldx + n];
    part[g][c] = s;
    __syncthreads();
    if (g == 0 && n < N) atomicAdd(&out[n], part[0][c] + part[1][c] + part[2][c] + part[3][c]);
}

// MSE loss gradient dY = 2 (Y - T) / (M N) (fp32 and bf16) and the loss sum:
// grid-stride over the elements, one atomic per block
__global__ __launch_bounds__(256) void k_mse_grad(const float* __restrict__ Y, const float* __restrict__ T,
                                                  long long n_el, float scale, float* __restrict__ dY,
                                                  __bf16* __restrict__ dYb, float* __restrict__ loss)
{
    __shared__ float part[4];
    float e = 0.0f;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_el;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const float d = Y[i] - T[i];
        e += d * d;
        dY[i] = 2.0f * scale * d;
        dYb[i] = static_cast<__bf16>(2.0f * scale * d);
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (part[0] + part[1] + part[2] + part[3]) * scale);
}

// sum of squares of x[0..n) added into *out (the gradient's norm for clipping)
__global__ __launch_bounds__(256) void k_sumsq(const float* __restrict__ x, long long n, float* __restrict__ out)
{
    __shared__ float s[256];
    float a = 0.0f;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        a = fmaf(x[i], x[i], a);
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, s[0]);
}

// Adam on fp32 master weights, refreshing the bf16 copy (nullable).  With
// gnorm2 (nullable) the gradient is scaled by min(1, clip / ||g||) first.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, __bf16* __restrict__ wb, const float* __restrict__ gnorm2, float clip)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    if (gnorm2) {
        const float nrm = sqrtf(*gnorm2);
        if (nrm > clip) gi *= clip / nrm;
    }
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = static_cast<__bf16>(wi);
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, long long n, __bf16* __restrict__ y)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = static_cast<__bf16>(x[i]);
}

// Features and start voltages in the batch's tiled layout, [row / 64][bus][row % 64]:
//   features: per bus (P - mu) / sigma, (Q - mu) / sigma, then the optional
//             0/1 outage mask -> bf16 [M][F];
//   start V:  V = Vm exp(j Va) from the network output Y (normalised Va, Vm
//             per bus), Vm held at the set-point where it is fixed (PV, slack)
//             and Va = 0 at the slack.
// Both sides are coalesced: a block moves a patch
// of one tile's 64 rows by 32 buses through LDS (the tiled side is read or
// written 64 lanes = 1 KB at a time, the row-major side 64 features of a row
// at a time).  Grid: (ceil(n / 32), ceil(M / 64)), 256 threads.
__global__ __launch_bounds__(256) void k_features_tiles_t(const double2* __restrict__ S_tiles, int M, int n,
                                                          const float* __restrict__ mu,
                                                          const float* __restrict__ sigma, __bf16* __restrict__ X,
                                                          long long ldx)
{
    __shared__ double2 sS[kPatchBuses][64];
    const int t = threadIdx.x;
    const int b0 = blockIdx.x * kPatchBuses, tile = blockIdx.y;
    for (int i = t; i < kPatchBuses * 64; i += 256) {
        const int bl = i >> 6, lane = i & 63;
        const int b = b0 + bl;
        sS[bl][lane] = b < n ? S_tiles[(static_cast<long long>(tile) * n + b) * 64 + lane] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int rr = t >> 2, q = t & 3;  // row of the patch, quarter of its 64 features
    const int r = tile * 64 + rr;
    if (r >= M) return;
    __bf16* x = X + static_cast<long long>(r) * ldx;
    constexpr int kQ = kPatchBuses / 4;  // buses of this thread: 2 kQ contiguous features
    const int bq = b0 + q * kQ;
    if (bq + kQ <= n && (ldx & 7) == 0) {
        // the 2 kQ features as 16-byte stores (row starts and bq are 16-byte aligned)
        __bf16 o[2 * kQ];
#pragma unroll
        for (int k = 0; k < kQ; ++k) {
            const int b = bq + k;
            const double2 v = sS[q * kQ + k][rr];
            o[2 * k] = static_cast<__bf16>(static_cast<float>((v.x - mu[2 * b]) / sigma[2 * b]));
            o[2 * k + 1] = static_cast<__bf16>(static_cast<float>((v.y - mu[2 * b + 1]) / sigma[2 * b + 1]));
        }
        uint4 w[2 * kQ * sizeof(__bf16) / sizeof(uint4)];
        __builtin_memcpy(w, o, sizeof(o));
        uint4* dst = reinterpret_cast<uint4*>(x + 2 * bq);
#pragma unroll
        for (int i = 0; i < static_cast<int>(sizeof(w) / sizeof(uint4)); ++i) dst[i] = w[i];
        return;
    }
    for (int k = 0; k < kQ; ++k) {
        const int bl = q * kQ + k;
        const int b = b0 + bl;
        if (b >= n) break;
        const double2 v = sS[bl][rr];
        x[2 * b] = static_cast<__bf16>(static_cast<float>((v.x - mu[2 * b]) / sigma[2 * b]));
        x[2 * b + 1] = static_cast<__bf16>(static_cast<float>((v.y - mu[2 * b + 1]) / sigma[2 * b + 1]));
    }
}

// outage-mask features 2n .. 2n + n_mask of each row
__global__ void k_mask_features(const unsigned char* __restrict__ mask, int M, int n, int n_mask,
                                __bf16* __restrict__ X, long long ldx)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(M) * n_mask) return;
    const int r = static_cast<int>(gid / n_mask);
    const int m = static_cast<int>(gid - static_cast<long long>(r) * n_mask);
    X[static_cast<long long>(r) * ldx + 2 * n + m] = static_cast<__bf16>(static_cast<float>(mask[gid]));
}

__global__ __launch_bounds__(256) void k_warm_v_tiles_t(const float* __restrict__ Y, long long ldy, int M, int n,
                                                        const float* __restrict__ mu,
                                                        const float* __restrict__ sigma,
                                                        const double* __restrict__ vfix, const int* __restrict__ kind,
                                                        double2* __restrict__ V_tiles)
{
    __shared__ float2 sY[kPatchBuses][64];
    const int t = threadIdx.x;
    const int b0 = blockIdx.x * kPatchBuses, tile = blockIdx.y;
    {
        const int rr = t >> 2, q = t & 3;
        const int r = tile * 64 + rr;
        const float* y = Y + static_cast<long long>(r) * ldy;
        constexpr int kQ = kPatchBuses / 4;  // buses of this thread: 2 kQ contiguous outputs
        const int bq = b0 + q * kQ;
        if (r < M && bq + kQ <= n && (ldy & 3) == 0) {
            // 16-byte loads (row starts and bq are 16-byte aligned)
            const float4* src = reinterpret_cast<const float4*>(y + 2 * bq);
#pragma unroll
            for (int k = 0; k < kQ / 2; ++k) {
                const float4 f = src[k];
                sY[q * kQ + 2 * k][rr] = make_float2(f.x, f.y);
                sY[q * kQ + 2 * k + 1][rr] = make_float2(f.z, f.w);
            }
        } else {
            for (int k = 0; k < kQ; ++k) {
                const int bl = q * kQ + k;
                const int b = b0 + bl;
                sY[bl][rr] = (r < M && b < n) ? make_float2(y[2 * b], y[2 * b + 1]) : make_float2(0.f, 0.f);
            }
        }
    }
    __syncthreads();
    for (int i = t; i < kPatchBuses * 64; i += 256) {
        const int bl = i >> 6, lane = i & 63;
        const int b = b0 + bl, r = tile * 64 + lane;
        if (b >= n || r >= M) continue;
        if (kind[b] == 2) continue;  // a fixed bus keeps the row's start V
        const float2 o = sY[bl][lane];
        double va = static_cast<double>(o.x) * sigma[2 * b] + mu[2 * b];
        double vm = static_cast<double>(o.y) * sigma[2 * b + 1] + mu[2 * b + 1];
        if (kind[b] != 1) vm = vfix[b];
        double sn, cs;
        sincos(va, &sn, &cs);
        V_tiles[(static_cast<long long>(tile) * n + b) * 64 + lane] = make_double2(vm * cs, vm * sn);
    }
}

template __global__ void k_gemm_kc<kEpiF32>(GemmArgs);
template __global__ void k_gemm_kc<kEpiBiasF32>(GemmArgs);
template __global__ void k_gemm_kc<kEpiBiasGeluBf16>(GemmArgs);
template __global__ void k_gemm<false, false, kEpiF32>(GemmArgs);
template __global__ void k_gemm<false, false, kEpiBiasF32>(GemmArgs);
template __global__ void k_gemm<false, false, kEpiBiasGeluBf16>(GemmArgs);
template __global__ void k_gemm<false, true, kEpiF32>(GemmArgs);
template __global__ void k_gemm<false, true, kEpiGeluBwd>(GemmArgs);
template __global__ void k_gemm<true, false, kEpiF32>(GemmArgs);
template __global__ void k_gemm<true, true, kEpiF32>(GemmArgs);

}  // namespace g2g
