// mlp.hip -- the learned GR15 step proposer: a 12 -> 64 -> 64 -> 1 ReLU MLP
// in FP32, with hand-written forward, backward and Adam kernels.
//
// Weight layout (ABIE_MLP_NPARAM floats): W1[12][64], b1[64], W2[64][64],
// b2[64], w3[64], b3[1]; W[i][j] connects input i to unit j.
//
// Training is deterministic: the per-sample pass writes activations and
// gradients to a workspace, the weight gradients are reduced over fixed
// chunks of samples, then the chunk partials are summed in chunk order.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <vector>

#include "abie_hip.h"

namespace abie {

constexpr int IN = ABIE_MLP_IN, H = ABIE_MLP_HIDDEN;
constexpr int OFF_W1 = 0, OFF_B1 = IN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H,
              OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;
static_assert(NPARAM == ABIE_MLP_NPARAM, "layout");

constexpr int TB = 128;  // threads per block, one sample per thread

__device__ __forceinline__ void load_weights(const float *__restrict__ w, float *sw) {
    for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) sw[i] = w[i];
    __syncthreads();
}

// Forward for one sample; returns the output and (optionally) activations.
template <bool KEEP>
__device__ __forceinline__ float forward_one(const float *sw, const float *x, float *h1_out, float *h2_out) {
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float a = sw[OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < IN; ++i) a = fmaf(x[i], sw[OFF_W1 + i * H + j], a);
        h1[j] = fmaxf(a, 0.0f);
        if (KEEP) h1_out[j] = h1[j];
    }
    float y = sw[OFF_B3];
    for (int j = 0; j < H; ++j) {
        float a = sw[OFF_B2 + j];
#pragma unroll
        for (int i = 0; i < H; ++i) a = fmaf(h1[i], sw[OFF_W2 + i * H + j], a);
        a = fmaxf(a, 0.0f);
        if (KEEP) h2_out[j] = a;
        y = fmaf(a, sw[OFF_W3 + j], y);
    }
    return y;
}

__global__ __launch_bounds__(TB) void mlp_forward_kernel(const float *__restrict__ w, const float *__restrict__ x,
                                                         float *__restrict__ y, int n) {
    __shared__ float sw[NPARAM];
    load_weights(w, sw);
    const int s = blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    float xs[IN];
#pragma unroll
    for (int i = 0; i < IN; ++i) xs[i] = x[(size_t)s * IN + i];
    y[s] = forward_one<false>(sw, xs, nullptr, nullptr);
}

// Per sample: forward, then backward of L = sum_s sw_s (y_s - t_s)^2 / sum(sw).
// Writes H1, H2 (post-ReLU), dH1, dH2 (pre-activation gradients), dy, loss_s.
__global__ __launch_bounds__(TB) void mlp_sample_kernel(const float *__restrict__ w, const float *__restrict__ x,
                                                        const float *__restrict__ t, const float *__restrict__ swt,
                                                        const float *__restrict__ inv_wsum, int n, float *H1, float *H2,
                                                        float *dH1, float *dH2, float *dY, float *L) {
    __shared__ float sw[NPARAM];
    load_weights(w, sw);
    const int s = blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    float xs[IN];
#pragma unroll
    for (int i = 0; i < IN; ++i) xs[i] = x[(size_t)s * IN + i];
    float *h1 = H1 + (size_t)s * H, *h2 = H2 + (size_t)s * H;
    const float y = forward_one<true>(sw, xs, h1, h2);
    const float wt = (swt ? swt[s] : 1.0f) * inv_wsum[0];
    const float r = y - t[s];
    L[s] = wt * r * r;
    const float dy = 2.0f * wt * r;
    dY[s] = dy;
    float d2[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        d2[j] = h2[j] > 0.0f ? dy * sw[OFF_W3 + j] : 0.0f;
        dH2[(size_t)s * H + j] = d2[j];
    }
    for (int i = 0; i < H; ++i) {
        float a = 0.0f;
#pragma unroll
        for (int j = 0; j < H; ++j) a = fmaf(d2[j], sw[OFF_W2 + i * H + j], a);
        dH1[(size_t)s * H + i] = h1[i] > 0.0f ? a : 0.0f;
    }
}

// Partial gradient of parameter p over samples [c*CH, (c+1)*CH).
constexpr int CH = 1024;
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float *__restrict__ x, const float *__restrict__ H1,
                                                         const float *__restrict__ H2, const float *__restrict__ dH1,
                                                         const float *__restrict__ dH2, const float *__restrict__ dY,
                                                         int n, float *__restrict__ part) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (p >= NPARAM) return;
    const int s0 = c * CH, s1 = min(n, s0 + CH);
    float acc = 0.0f;
    if (p < OFF_B1) {
        const int i = p / H, j = p % H;
        for (int s = s0; s < s1; ++s) acc = fmaf(x[(size_t)s * IN + i], dH1[(size_t)s * H + j], acc);
    } else if (p < OFF_W2) {
        const int j = p - OFF_B1;
        for (int s = s0; s < s1; ++s) acc += dH1[(size_t)s * H + j];
    } else if (p < OFF_B2) {
        const int q = p - OFF_W2, i = q / H, j = q % H;
        for (int s = s0; s < s1; ++s) acc = fmaf(H1[(size_t)s * H + i], dH2[(size_t)s * H + j], acc);
    } else if (p < OFF_W3) {
        const int j = p - OFF_B2;
        for (int s = s0; s < s1; ++s) acc += dH2[(size_t)s * H + j];
    } else if (p < OFF_B3) {
        const int j = p - OFF_W3;
        for (int s = s0; s < s1; ++s) acc = fmaf(H2[(size_t)s * H + j], dY[s], acc);
    } else {
        for (int s = s0; s < s1; ++s) acc += dY[s];
    }
    part[(size_t)c * NPARAM + p] = acc;
}

__global__ __launch_bounds__(256) void mlp_sum_kernel(const float *__restrict__ part, int chunks,
                                                      float *__restrict__ grad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NPARAM) return;
    float a = 0.0f;
    for (int c = 0; c < chunks; ++c) a += part[(size_t)c * NPARAM + p];
    grad[p] = a;
}

// sum of per-sample weights (or n) and of the loss terms, in sample order
__global__ void mlp_wsum_kernel(const float *__restrict__ swt, int n, float *inv_wsum) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0;
    for (int s = 0; s < n; ++s) a += swt ? swt[s] : 1.0f;
    inv_wsum[0] = a > 0 ? (float)(1.0 / a) : 0.0f;
}

__global__ __launch_bounds__(256) void mlp_loss_kernel(const float *__restrict__ L, int n, float *loss) {
    __shared__ double red[256];
    double a = 0.0;
    for (int s = threadIdx.x; s < n; s += 256) a += L[s];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = (float)red[0];
}

__global__ __launch_bounds__(256) void adam_kernel(float *w, const float *g, float *m, float *v, int n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace abie

using namespace abie;

extern "C" {

int abie_mlp_forward(const float *w, const float *x, float *y, int n, void *stream) {
    if (!w || !x || !y || n < 0) return ABIE_HIP_EINVAL;
    if (n == 0) return 0;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + TB - 1) / TB), dim3(TB), 0, (hipStream_t)stream, w, x, y, n);
    return hipGetLastError() == hipSuccess ? 0 : ABIE_HIP_EHIP;
}

int abie_mlp_loss_grad(const float *w, const float *x, const float *target, const float *sw, int n, float *grad,
                       float *loss, void *stream) {
    if (!w || !x || !target || !grad || !loss || n <= 0) return ABIE_HIP_EINVAL;
    const int chunks = (n + CH - 1) / CH;
    const size_t need = (size_t)4 * n * H + 2 * (size_t)n + (size_t)chunks * NPARAM + 1;
    hipStream_t st = (hipStream_t)stream;
    // stream-ordered workspace: concurrent calls on different streams each
    // get their own, and the pool reuses it once the stream passes the free
    float *ws = nullptr;
    if (hipMallocAsync((void **)&ws, need * sizeof(float), st) != hipSuccess) return ABIE_HIP_ENOMEM;
    float *H1 = ws, *H2 = H1 + (size_t)n * H, *dH1 = H2 + (size_t)n * H, *dH2 = dH1 + (size_t)n * H;
    float *dY = dH2 + (size_t)n * H, *L = dY + n, *part = L + n, *inv = part + (size_t)chunks * NPARAM;
    hipLaunchKernelGGL(mlp_wsum_kernel, dim3(1), dim3(64), 0, st, sw, n, inv);
    hipLaunchKernelGGL(mlp_sample_kernel, dim3((n + TB - 1) / TB), dim3(TB), 0, st, w, x, target, sw, inv, n, H1, H2,
                       dH1, dH2, dY, L);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((NPARAM + 255) / 256, chunks), dim3(256), 0, st, x, H1, H2, dH1, dH2,
                       dY, n, part);
    hipLaunchKernelGGL(mlp_sum_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, part, chunks, grad);
    hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, st, L, n, loss);
    const hipError_t le = hipGetLastError();
    const hipError_t fe = hipFreeAsync(ws, st);
    return (le == hipSuccess && fe == hipSuccess) ? 0 : ABIE_HIP_EHIP;
}

int abie_mlp_adam(float *w, const float *g, float *m, float *v, int n_param, int step, float lr, float beta1,
                  float beta2, float eps, void *stream) {
    if (!w || !g || !m || !v || n_param <= 0 || step <= 0) return ABIE_HIP_EINVAL;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((n_param + 255) / 256), dim3(256), 0, (hipStream_t)stream, w, g, m, v,
                       n_param, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? 0 : ABIE_HIP_EHIP;
}

}  // extern "C"
