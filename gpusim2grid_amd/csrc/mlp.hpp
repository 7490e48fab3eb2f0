// MLP warm start: kernel interfaces shared by mlp.hip and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace g2g {

enum GemmEpilogue : int {
    kEpiF32 = 0,           // C = A B
    kEpiBiasF32 = 1,       // C = A B + bias
    kEpiBiasGeluBf16 = 2,  // Z = A B + bias (fp32), H = gelu(Z) (bf16)
    kEpiGeluBwd = 3,       // C = (A B) * gelu'(Z) (fp32), H = same (bf16)
};

struct GemmArgs {
    const __bf16* A;  // M x K; element (m, k) at A[m lda + k], or A[k lda + m] when transposed
    long long lda;
    const __bf16* B;  // K x N; element (k, n) at B[k ldb + n], or B[n ldb + k] when transposed
    long long ldb;
    int M, N, K;
    float* C;
    long long ldc;
    __bf16* H;
    long long ldh;
    float* Z;
    long long ldz;
    const float* bias;
};

template <bool TA, bool TB, int EPI>
__global__ void k_gemm(GemmArgs g);
template <int EPI>
__global__ void k_gemm_kc(GemmArgs g);
__global__ void k_transpose_bf16(const float* W, int K, int N, __bf16* Wt);
constexpr int kColRows = 128;  // rows a k_colsum block sums
__global__ void k_colsum(const float* X, long long ldx, int M, int N, float* out);
__global__ void k_mse_grad(const float* Y, const float* T, long long n_el, float scale, float* dY, __bf16* dYb,
                           float* loss);
__global__ void k_adam(float* w, float* m, float* v, const float* g, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2, __bf16* wb, const float* gnorm2, float clip);
__global__ void k_sumsq(const float* x, long long n, float* out);
__global__ void k_fill_int(int* p, int v, long long n);
__global__ void k_f32_to_bf16(const float* x, long long n, __bf16* y);
constexpr int kPatchBuses = 32;  // buses per block of the tiled-layout conversions
__global__ void k_features_tiles_t(const double2* S_tiles, int M, int n, const float* mu, const float* sigma,
                                   __bf16* X, long long ldx);
__global__ void k_mask_features(const unsigned char* mask, int M, int n, int n_mask, __bf16* X, long long ldx);
__global__ void k_warm_v_tiles_t(const float* Y, long long ldy, int M, int n, const float* mu, const float* sigma,
                                 const double* vfix, const int* kind, double2* V_tiles);

}  // namespace g2g
