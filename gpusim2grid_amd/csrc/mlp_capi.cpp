// C ABI of the learned warm start (see include/gpusim2grid_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpusim2grid_amd.h"
#include "capi_common.hpp"
#include "mlp.hpp"

using namespace g2g;

struct g2g_mlp_s {
    int n_in = 0, H = 0, n_out = 0;
    hipStream_t stream = nullptr;
    // fp32 parameters, contiguous: W1 [n_in x H], b1 [H], W2 [H x H], b2 [H], W3 [H x n_out], b3 [n_out]
    size_t o_W1 = 0, o_b1 = 0, o_W2 = 0, o_b2 = 0, o_W3 = 0, o_b3 = 0, n_params = 0;
    float* w = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* g = nullptr;
    __bf16* wb = nullptr;  // bf16 copy of all parameters (the GEMMs read the weight parts)
    __bf16* wt = nullptr;  // bf16 transposed weights (N x K) for the K-contiguous forward GEMM
    bool kc = false;       // every width a multiple of 8: the forward uses k_gemm_kc
    int step = 0;
    // normalisation (features: z-score per input; outputs: per output)
    float* in_mu = nullptr;
    float* in_sigma = nullptr;
    float* out_mu = nullptr;
    float* out_sigma = nullptr;
    // buses for the warm start: kind (0 pv, 1 pq, 2 fixed), |V| of pv buses
    int* kind = nullptr;
    double* vfix = nullptr;
    int n_bus = 0;
    // workspace for M rows
    int ws_M = 0;
    __bf16* X = nullptr;
    float* Z1 = nullptr;
    float* Z2 = nullptr;
    __bf16* H1 = nullptr;
    __bf16* H2 = nullptr;
    float* Y = nullptr;
    float* T = nullptr;
    float* dY = nullptr;
    __bf16* dYb = nullptr;
    float* dZ1 = nullptr;
    float* dZ2 = nullptr;
    __bf16* dZ1b = nullptr;
    __bf16* dZ2b = nullptr;
    float* loss = nullptr;
    float clip = 0.0f;          // gradient-norm clip of the training steps (0: off)
    float* gnorm2 = nullptr;    // its squared-norm accumulator
    DeviceScratch ws;   // workspace buffers (freed on growth / destruction)
    DeviceScratch own;  // parameters and state

    __bf16* W1b() const { return wb + o_W1; }
    __bf16* W2b() const { return wb + o_W2; }
    __bf16* W3b() const { return wb + o_W3; }
    __bf16* W1t() const { return wt + o_W1; }
    __bf16* W2t() const { return wt + o_W2; }
    __bf16* W3t() const { return wt + o_W3; }

    // refresh the transposed bf16 weights from the fp32 masters
    void transpose_weights(hipStream_t st)
    {
        if (!kc) return;
        auto tr = [&](size_t off, int K, int N) {
            const dim3 grid((N + 31) / 32, (K + 31) / 32);
            k_transpose_bf16<<<grid, 256, 0, st>>>(w + off, K, N, wt + off);
        };
        tr(o_W1, n_in, H);
        tr(o_W2, H, H);
        tr(o_W3, H, n_out);
        HIPCHK(hipGetLastError());
    }

    void workspace(int M)
    {
        if (M <= ws_M) return;
        HIPCHK(hipStreamSynchronize(stream));
        ws.release();
        const size_t m = static_cast<size_t>(M);
        X = ws.get<__bf16>(m * n_in);
        Z1 = ws.get<float>(m * H);
        Z2 = ws.get<float>(m * H);
        H1 = ws.get<__bf16>(m * H);
        H2 = ws.get<__bf16>(m * H);
        Y = ws.get<float>(m * n_out);
        T = ws.get<float>(m * n_out);
        dY = ws.get<float>(m * n_out);
        dYb = ws.get<__bf16>(m * n_out);
        dZ1 = ws.get<float>(m * H);
        dZ2 = ws.get<float>(m * H);
        dZ1b = ws.get<__bf16>(m * H);
        dZ2b = ws.get<__bf16>(m * H);
        loss = ws.get<float>(1);
        ws_M = M;
    }
};

namespace {

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t st, const __bf16* A, long long lda, const __bf16* B, long long ldb, int M, int N, int K,
          float* C, long long ldc, __bf16* Hh = nullptr, long long ldh = 0, float* Z = nullptr, long long ldz = 0,
          const float* bias = nullptr)
{
    GemmArgs a{A, lda, B, ldb, M, N, K, C, ldc, Hh, ldh, Z, ldz, bias};
    const dim3 grid((N + 127) / 128, (M + 127) / 128);
    k_gemm<TA, TB, EPI><<<grid, 256, 0, st>>>(a);
    HIPCHK(hipGetLastError());
}

// X (fp32, M x n_in, host or device) -> the workspace's bf16 X
void stage_x(g2g_mlp_s* p, const float* X, int M, int on_device)
{
    const size_t n = static_cast<size_t>(M) * p->n_in;
    const float* src = X;
    DeviceScratch tmp;
    if (!on_device) {
        float* d = tmp.get<float>(n);
        HIPCHK(hipMemcpyAsync(d, X, n * sizeof(float), hipMemcpyHostToDevice, p->stream));
        src = d;
    }
    k_f32_to_bf16<<<blocks_for(static_cast<long long>(n), 256), 256, 0, p->stream>>>(src, static_cast<long long>(n),
                                                                                     p->X);
    HIPCHK(hipGetLastError());
    if (!on_device) HIPCHK(hipStreamSynchronize(p->stream));
}

template <int EPI>
void gemm_kc(hipStream_t st, const __bf16* A, long long lda, const __bf16* Bt, long long ldb, int M, int N, int K,
             float* C, long long ldc, __bf16* Hh, long long ldh, float* Z, long long ldz, const float* bias)
{
    GemmArgs a{A, lda, Bt, ldb, M, N, K, C, ldc, Hh, ldh, Z, ldz, bias};
    const dim3 grid((N + 127) / 128, (M + 127) / 128);
    k_gemm_kc<EPI><<<grid, 256, 0, st>>>(a);
    HIPCHK(hipGetLastError());
}

// forward of the staged X into p->Y; keep_z: the pre-activations the backward
// pass needs (inference skips them)
void forward(g2g_mlp_s* p, int M, bool keep_z = true)
{
    if (p->kc) {
        gemm_kc<kEpiBiasGeluBf16>(p->stream, p->X, p->n_in, p->W1t(), p->n_in, M, p->H, p->n_in, nullptr, 0, p->H1,
                                  p->H, keep_z ? p->Z1 : nullptr, p->H, p->w + p->o_b1);
        gemm_kc<kEpiBiasGeluBf16>(p->stream, p->H1, p->H, p->W2t(), p->H, M, p->H, p->H, nullptr, 0, p->H2, p->H,
                                  keep_z ? p->Z2 : nullptr, p->H, p->w + p->o_b2);
        gemm_kc<kEpiBiasF32>(p->stream, p->H2, p->H, p->W3t(), p->H, M, p->n_out, p->H, p->Y, p->n_out, nullptr, 0,
                             nullptr, 0, p->w + p->o_b3);
        return;
    }
    gemm<false, false, kEpiBiasGeluBf16>(p->stream, p->X, p->n_in, p->W1b(), p->H, M, p->H, p->n_in, nullptr, 0,
                                         p->H1, p->H, p->Z1, p->H, p->w + p->o_b1);
    gemm<false, false, kEpiBiasGeluBf16>(p->stream, p->H1, p->H, p->W2b(), p->H, M, p->H, p->H, nullptr, 0, p->H2,
                                         p->H, p->Z2, p->H, p->w + p->o_b2);
    gemm<false, false, kEpiBiasF32>(p->stream, p->H2, p->H, p->W3b(), p->n_out, M, p->n_out, p->H, p->Y, p->n_out,
                                    nullptr, 0, nullptr, 0, p->w + p->o_b3);
}

// out[n] = sum over the M rows of X[m][n]
void colsum(hipStream_t st, const float* X, long long ldx, int M, int N, float* out)
{
    HIPCHK(hipMemsetAsync(out, 0, static_cast<size_t>(N) * sizeof(float), st));
    const dim3 grid((N + 63) / 64, (M + kColRows - 1) / kColRows);
    k_colsum<<<grid, 256, 0, st>>>(X, ldx, M, N, out);
    HIPCHK(hipGetLastError());
}

// loss and gradients (into p->g) of the staged X against targets p->T
void backward(g2g_mlp_s* p, int M)
{
    hipStream_t st = p->stream;
    const long long ne = static_cast<long long>(M) * p->n_out;
    HIPCHK(hipMemsetAsync(p->loss, 0, sizeof(float), st));
    k_mse_grad<<<std::min<unsigned>(blocks_for(ne, 256), 2048), 256, 0, st>>>(p->Y, p->T, ne,
                                                                              1.0f / static_cast<float>(ne), p->dY,
                                                                              p->dYb,
                                                    p->loss);
    HIPCHK(hipGetLastError());
    // layer 3: dW3 = H2^T dY, db3 = sum dY, dZ2 = (dY W3^T) gelu'(Z2)
    gemm<true, false, kEpiF32>(st, p->H2, p->H, p->dYb, p->n_out, p->H, p->n_out, M, p->g + p->o_W3, p->n_out);
    colsum(st, p->dY, p->n_out, M, p->n_out, p->g + p->o_b3);
    gemm<false, true, kEpiGeluBwd>(st, p->dYb, p->n_out, p->W3b(), p->n_out, M, p->H, p->n_out, p->dZ2, p->H,
                                   p->dZ2b, p->H, p->Z2, p->H);
    // layer 2
    gemm<true, false, kEpiF32>(st, p->H1, p->H, p->dZ2b, p->H, p->H, p->H, M, p->g + p->o_W2, p->H);
    colsum(st, p->dZ2, p->H, M, p->H, p->g + p->o_b2);
    gemm<false, true, kEpiGeluBwd>(st, p->dZ2b, p->H, p->W2b(), p->H, M, p->H, p->H, p->dZ1, p->H, p->dZ1b, p->H,
                                   p->Z1, p->H);
    // layer 1
    gemm<true, false, kEpiF32>(st, p->X, p->n_in, p->dZ1b, p->H, p->n_in, p->H, M, p->g + p->o_W1, p->H);
    colsum(st, p->dZ1, p->H, M, p->H, p->g + p->o_b1);
    HIPCHK(hipGetLastError());
}

void stage_t(g2g_mlp_s* p, const float* T, int M, int on_device)
{
    const size_t n = static_cast<size_t>(M) * p->n_out;
    HIPCHK(hipMemcpyAsync(p->T, T, n * sizeof(float), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          p->stream));
    if (!on_device) HIPCHK(hipStreamSynchronize(p->stream));
}

}  // namespace

namespace g2g {

// Warm-start voltages of rows [0, M) of a batch from its tiled Sbus (used by capi.cpp)
void mlp_warm_start_tiles(g2g_mlp_s* p, const double2* S_tiles, int M, int n, const unsigned char* d_mask, int n_mask,
                          double2* V_tiles, hipStream_t st)
{
    if (!p->kind) throw std::invalid_argument("the MLP has no bus types: call g2g_mlp_set_buses");
    if (p->n_bus != n) throw std::invalid_argument("the MLP's bus count differs from the batch's");
    if (p->n_in != 2 * n + n_mask) throw std::invalid_argument("the MLP's input width is not 2 n + mask");
    if (p->n_out != 2 * n) throw std::invalid_argument("the MLP's output width is not 2 n");
    if (!p->in_mu) throw std::invalid_argument("the MLP has no normalisation: call g2g_mlp_set_normalization");
    hipStream_t keep = p->stream;
    p->stream = st;
    p->workspace(M);
    const dim3 patches((n + kPatchBuses - 1) / kPatchBuses, (M + 63) / 64);
    k_features_tiles_t<<<patches, 256, 0, st>>>(S_tiles, M, n, p->in_mu, p->in_sigma, p->X, p->n_in);
    if (n_mask > 0)
        k_mask_features<<<blocks_for(static_cast<long long>(M) * n_mask, 256), 256, 0, st>>>(d_mask, M, n, n_mask,
                                                                                            p->X, p->n_in);
    HIPCHK(hipGetLastError());
    forward(p, M, false);
    k_warm_v_tiles_t<<<patches, 256, 0, st>>>(p->Y, p->n_out, M, n, p->out_mu, p->out_sigma, p->vfix, p->kind,
                                              V_tiles);
    HIPCHK(hipGetLastError());
    p->stream = keep;
}

}  // namespace g2g

extern "C" {

int g2g_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint64_t seed, void* stream, g2g_mlp_t* out)
{
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        if (n_in <= 0 || n_hidden <= 0 || n_out <= 0) throw std::invalid_argument("layer widths must be positive");
        auto p = std::make_unique<g2g_mlp_s>();
        p->n_in = n_in;
        p->H = n_hidden;
        p->n_out = n_out;
        p->stream = static_cast<hipStream_t>(stream);
        size_t o = 0;
        auto take = [&](size_t n) { const size_t at = o; o += (n + 63) / 64 * 64; return at; };
        p->o_W1 = take(static_cast<size_t>(n_in) * n_hidden);
        p->o_b1 = take(n_hidden);
        p->o_W2 = take(static_cast<size_t>(n_hidden) * n_hidden);
        p->o_b2 = take(n_hidden);
        p->o_W3 = take(static_cast<size_t>(n_hidden) * n_out);
        p->o_b3 = take(n_out);
        p->n_params = o;
        p->w = p->own.get<float>(o);
        p->m = p->own.get<float>(o);
        p->v = p->own.get<float>(o);
        p->g = p->own.get<float>(o);
        p->wb = p->own.get<__bf16>(o);
        p->kc = n_in % 8 == 0 && n_hidden % 8 == 0 && n_out % 8 == 0;
        if (p->kc) p->wt = p->own.get<__bf16>(o);
        // Glorot-uniform weights, zero biases (seeded)
        std::vector<float> h(o, 0.0f);
        std::mt19937_64 rng(seed);
        auto fill = [&](size_t at, int fan_in, int fan_out) {
            const float a = std::sqrt(6.0f / static_cast<float>(fan_in + fan_out));
            std::uniform_real_distribution<float> u(-a, a);
            for (size_t i = 0; i < static_cast<size_t>(fan_in) * fan_out; ++i) h[at + i] = u(rng);
        };
        fill(p->o_W1, n_in, n_hidden);
        fill(p->o_W2, n_hidden, n_hidden);
        fill(p->o_W3, n_hidden, n_out);
        HIPCHK(hipMemcpy(p->w, h.data(), o * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(p->m, 0, o * sizeof(float)));
        HIPCHK(hipMemset(p->v, 0, o * sizeof(float)));
        // the layers' gradients never write the padding between them, which
        // the gradient norm and the Adam update read: zero it once
        HIPCHK(hipMemset(p->g, 0, o * sizeof(float)));
        k_f32_to_bf16<<<blocks_for(static_cast<long long>(o), 256), 256>>>(p->w, static_cast<long long>(o), p->wb);
        HIPCHK(hipGetLastError());
        p->transpose_weights(nullptr);
        HIPCHK(hipDeviceSynchronize());
        *out = p.release();
    });
}

void g2g_mlp_destroy(g2g_mlp_t p)
{
    if (!p) return;
    (void)hipStreamSynchronize(p->stream);
    DeviceScratch::free_one(p->in_mu);
    DeviceScratch::free_one(p->in_sigma);
    DeviceScratch::free_one(p->out_mu);
    DeviceScratch::free_one(p->out_sigma);
    DeviceScratch::free_one(p->kind);
    DeviceScratch::free_one(p->vfix);
    delete p;
}

int64_t g2g_mlp_num_params(g2g_mlp_t p) { return p ? static_cast<int64_t>(p->n_params) : -1; }

int g2g_mlp_param_offsets(g2g_mlp_t p, int64_t* offsets)
{
    return guarded([&] {
        if (!p || !offsets) throw std::invalid_argument("null argument");
        const size_t o[6] = {p->o_W1, p->o_b1, p->o_W2, p->o_b2, p->o_W3, p->o_b3};
        for (int i = 0; i < 6; ++i) offsets[i] = static_cast<int64_t>(o[i]);
    });
}

int g2g_mlp_get_params(g2g_mlp_t p, float* out, int64_t capacity)
{
    return guarded([&] {
        if (!p || !out) throw std::invalid_argument("null argument");
        if (capacity < static_cast<int64_t>(p->n_params)) throw std::invalid_argument("output buffer too small");
        HIPCHK(hipStreamSynchronize(p->stream));
        HIPCHK(hipMemcpy(out, p->w, p->n_params * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int g2g_mlp_set_params(g2g_mlp_t p, const float* in, int64_t n)
{
    return guarded([&] {
        if (!p || !in) throw std::invalid_argument("null argument");
        if (n != static_cast<int64_t>(p->n_params)) throw std::invalid_argument("parameter count mismatch");
        HIPCHK(hipStreamSynchronize(p->stream));
        HIPCHK(hipMemcpy(p->w, in, p->n_params * sizeof(float), hipMemcpyHostToDevice));
        k_f32_to_bf16<<<blocks_for(static_cast<long long>(n), 256), 256, 0, p->stream>>>(p->w, n, p->wb);
        HIPCHK(hipGetLastError());
        p->transpose_weights(p->stream);
        HIPCHK(hipStreamSynchronize(p->stream));
    });
}

int g2g_mlp_set_normalization(g2g_mlp_t p, const float* in_mu, const float* in_sigma, const float* out_mu,
                              const float* out_sigma)
{
    return guarded([&] {
        if (!p || !in_mu || !in_sigma || !out_mu || !out_sigma) throw std::invalid_argument("null argument");
        auto up = [&](float*& dst, const float* src, int n) {
            if (!dst) dst = dalloc<float>(static_cast<size_t>(n));
            HIPCHK(hipMemcpy(dst, src, static_cast<size_t>(n) * sizeof(float), hipMemcpyHostToDevice));
        };
        up(p->in_mu, in_mu, p->n_in);
        up(p->in_sigma, in_sigma, p->n_in);
        up(p->out_mu, out_mu, p->n_out);
        up(p->out_sigma, out_sigma, p->n_out);
    });
}

int g2g_mlp_set_buses(g2g_mlp_t p, const int32_t* kind, const double* vfix, int32_t n)
{
    return guarded([&] {
        if (!p || !kind || !vfix) throw std::invalid_argument("null argument");
        if (n <= 0) throw std::invalid_argument("n must be positive");
        DeviceScratch::free_one(p->kind);
        DeviceScratch::free_one(p->vfix);
        p->kind = dupload(std::vector<int>(kind, kind + n));
        p->vfix = dupload(std::vector<double>(vfix, vfix + n));
        p->n_bus = n;
    });
}

int g2g_mlp_forward(g2g_mlp_t p, const float* X, int32_t M, float* Y, int32_t on_device)
{
    return guarded([&] {
        if (!p || !X || !Y) throw std::invalid_argument("null argument");
        if (M <= 0) throw std::invalid_argument("M must be positive");
        p->workspace(M);
        stage_x(p, X, M, on_device);
        forward(p, M, false);
        HIPCHK(hipMemcpyAsync(Y, p->Y, static_cast<size_t>(M) * p->n_out * sizeof(float),
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
    });
}

int g2g_mlp_loss_grad(g2g_mlp_t p, const float* X, const float* T, int32_t M, float* loss, float* grads,
                      int32_t on_device)
{
    return guarded([&] {
        if (!p || !X || !T) throw std::invalid_argument("null argument");
        if (M <= 0) throw std::invalid_argument("M must be positive");
        p->workspace(M);
        stage_x(p, X, M, on_device);
        stage_t(p, T, M, on_device);
        forward(p, M);
        backward(p, M);
        if (loss) HIPCHK(hipMemcpyAsync(loss, p->loss, sizeof(float), hipMemcpyDeviceToHost, p->stream));
        if (grads)
            HIPCHK(hipMemcpyAsync(grads, p->g, p->n_params * sizeof(float), hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
    });
}

int g2g_mlp_train_step(g2g_mlp_t p, const float* X, const float* T, int32_t M, float lr, float* loss,
                       int32_t on_device)
{
    return guarded([&] {
        if (!p || !X || !T) throw std::invalid_argument("null argument");
        if (M <= 0) throw std::invalid_argument("M must be positive");
        p->workspace(M);
        stage_x(p, X, M, on_device);
        stage_t(p, T, M, on_device);
        forward(p, M);
        backward(p, M);
        ++p->step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f / (1.0f - std::pow(b1, static_cast<float>(p->step)));
        const float c2 = 1.0f / (1.0f - std::pow(b2, static_cast<float>(p->step)));
        const long long n = static_cast<long long>(p->n_params);
        if (p->clip > 0.0f) {
            if (!p->gnorm2) p->gnorm2 = p->own.get<float>(1);
            k_fill_int<<<1, 64, 0, p->stream>>>(reinterpret_cast<int*>(p->gnorm2), 0, 1);  // 0.0f
            k_sumsq<<<std::min<long long>(1024, blocks_for(n, 256)), 256, 0, p->stream>>>(p->g, n, p->gnorm2);
        }
        k_adam<<<blocks_for(n, 256), 256, 0, p->stream>>>(p->w, p->m, p->v, p->g, n, lr, b1, b2, 1e-8f, c1, c2, p->wb,
                                                           p->clip > 0.0f ? p->gnorm2 : nullptr, p->clip);
        HIPCHK(hipGetLastError());
        p->transpose_weights(p->stream);
        if (loss) {
            HIPCHK(hipMemcpyAsync(loss, p->loss, sizeof(float), hipMemcpyDeviceToHost, p->stream));
            HIPCHK(hipStreamSynchronize(p->stream));
        }
    });
}

int g2g_mlp_set_grad_clip(g2g_mlp_t p, float max_norm)
{
    return guarded([&] {
        if (!p) throw std::invalid_argument("null argument");
        if (!(max_norm >= 0.0f)) throw std::invalid_argument("max_norm must be >= 0 (0: off)");
        p->clip = max_norm;
    });
}

int g2g_gemm_bf16(int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, const uint16_t* A, int64_t lda,
                  const uint16_t* B, int64_t ldb, float* C, int64_t ldc)
{
    return guarded([&] {
        if (!A || !B || !C) throw std::invalid_argument("null argument");
        if (M <= 0 || N <= 0 || K <= 0) throw std::invalid_argument("sizes must be positive");
        const size_t na = static_cast<size_t>(trans_a ? K : M) * static_cast<size_t>(lda);
        const size_t nb = static_cast<size_t>(trans_b ? N : K) * static_cast<size_t>(ldb);
        DeviceScratch s;
        __bf16* dA = s.get<__bf16>(na);
        __bf16* dB = s.get<__bf16>(nb);
        float* dC = s.get<float>(static_cast<size_t>(M) * ldc);
        HIPCHK(hipMemcpy(dA, A, na * 2, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dB, B, nb * 2, hipMemcpyHostToDevice));
        if (!trans_a && !trans_b) gemm<false, false, kEpiF32>(nullptr, dA, lda, dB, ldb, M, N, K, dC, ldc);
        else if (!trans_a && trans_b && K % 8 == 0 && lda % 8 == 0 && ldb % 8 == 0)
            gemm_kc<kEpiF32>(nullptr, dA, lda, dB, ldb, M, N, K, dC, ldc, nullptr, 0, nullptr, 0, nullptr);
        else if (!trans_a && trans_b) gemm<false, true, kEpiF32>(nullptr, dA, lda, dB, ldb, M, N, K, dC, ldc);
        else if (trans_a && !trans_b) gemm<true, false, kEpiF32>(nullptr, dA, lda, dB, ldb, M, N, K, dC, ldc);
        else gemm<true, true, kEpiF32>(nullptr, dA, lda, dB, ldb, M, N, K, dC, ldc);
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(C, dC, static_cast<size_t>(M) * ldc * sizeof(float), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
