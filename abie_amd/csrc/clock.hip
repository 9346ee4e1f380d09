// clock.hip -- abie_hip_issue_rate: the shader clock and the SIMD cycles per
// wave instruction of FP64 VALU work, measured in a kernel.
//
// The FP64 vector peak (78.6 TFLOP/s on MI355X) assumes the 2.4 GHz peak
// engine clock; under a load that keeps every SIMD issuing FP64 instructions
// the clock settles lower, and a transcendental such as v_rsq_f64 holds the
// SIMD longer than an FMA.  Pricing a VALU-bound kernel against its
// instruction stream therefore needs both, measured.  The kernel here keeps
// 8 waves per SIMD issuing independent FP64 chains (FMA only, rsq only, or
// one rsq per 16 FMAs) and reads, per wave, the shader cycle counter
// (s_memtime) and the constant-rate wall clock (s_memrealtime) around the
// loop: their ratio is the shader clock; the kernel's event-timed duration
// at that clock, over the instructions each SIMD issued, gives the cycles
// per wave instruction.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "abie_hip.h"

namespace abie {

int report_error(int code, const char *msg);  // runtime.hip: sets abie_hip_last_error

constexpr int kClockChains = 8;     // independent chains per lane
constexpr int kClockUnroll = 8;     // chain rounds per loop iteration

__device__ __forceinline__ void record_ticks(long long *__restrict__ ticks, long long w0, long long k0) {
    const long long w1 = wall_clock64(), k1 = clock64();
    if ((threadIdx.x & 63) == 0) {
        const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
        ticks[4 * w + 0] = w0;
        ticks[4 * w + 1] = w1;
        ticks[4 * w + 2] = k0;
        ticks[4 * w + 3] = k1;
    }
}

// KIND 0: v_fma_f64 only; 1: v_rsq_f64 only; 2: per chain round, one
// v_rsq_f64 and 16 v_fma_f64 on independent registers (the force
// interaction's mix).  Every chain is independent of the others.
template <int KIND>
__global__ __launch_bounds__(256) void issue_rate_kernel(int iters, double *__restrict__ sink,
                                                         long long *__restrict__ ticks) {
    double a[kClockChains], f[16];
#pragma unroll
    for (int c = 0; c < kClockChains; ++c) a[c] = 1.5 + threadIdx.x + c;
#pragma unroll
    for (int q = 0; q < 16; ++q) f[q] = q;
    const long long w0 = wall_clock64(), k0 = clock64();
    for (int i = 0; i < iters; ++i) {
        if (KIND == 2) {
#pragma unroll
            for (int c = 0; c < kClockChains; ++c) {
                a[c] = __builtin_amdgcn_rsq(a[c]);
#pragma unroll
                for (int q = 0; q < 16; ++q) f[q] = fma(f[q], 0.999999, 1e-7);
            }
        } else {
#pragma unroll
            for (int u = 0; u < kClockUnroll; ++u)
#pragma unroll
                for (int c = 0; c < kClockChains; ++c)
                    a[c] = KIND == 0 ? fma(a[c], 0.999999, 1e-7) : __builtin_amdgcn_rsq(a[c]);
        }
    }
    record_ticks(ticks, w0, k0);
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < kClockChains; ++c) s += a[c];
#pragma unroll
    for (int q = 0; q < 16; ++q) s += f[q];
    sink[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace abie

using namespace abie;

#define CLK_TRY(x)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            snprintf(msg, sizeof msg, "issue_rate: %s: %s", #x, hipGetErrorString(e_));   \
            rc = report_error(ABIE_HIP_EHIP, msg);                                        \
            goto done;                                                                    \
        }                                                                                 \
    } while (0)

extern "C" int abie_hip_issue_rate(int kind, double milliseconds, double *shader_ghz,
                                   double *cycles_per_instruction) {
    char msg[256];
    int rc = 0;
    int dev = 0, n_cu = 0, khz = 0;
    double *sink = nullptr;
    long long *ticks = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!shader_ghz || !cycles_per_instruction || kind < 0 || kind > 2 || !(milliseconds > 0.0) ||
        milliseconds > 10000.0) {
        snprintf(msg, sizeof msg, "issue_rate: bad arguments (kind %d, milliseconds %g)", kind, milliseconds);
        return report_error(ABIE_HIP_EINVAL, msg);
    }
    {
        CLK_TRY(hipGetDevice(&dev));
        CLK_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
        CLK_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        // 8 waves per SIMD: 8 blocks of 4 waves per CU, all resident at once
        const int blocks = 8 * n_cu, waves = 4 * blocks;
        CLK_TRY(hipMalloc((void **)&sink, sizeof(double) * 256 * (size_t)blocks));
        CLK_TRY(hipMalloc((void **)&ticks, sizeof(long long) * 4 * (size_t)waves));
        CLK_TRY(hipEventCreate(&e0));
        CLK_TRY(hipEventCreate(&e1));
        // wave instructions per loop iteration and their nominal cycles
        const double inst = kind == 2 ? kClockChains * 17.0 : (double)kClockChains * kClockUnroll;
        const double nominal = kind == 0 ? 4.0 * inst : kind == 1 ? 16.0 * inst : kClockChains * (16.0 * 4 + 16);
        // sized for about `milliseconds` at 8 waves per SIMD and ~2 GHz
        const int iters = std::max(10, (int)(milliseconds * 1e-3 * 2.0e9 / (8.0 * nominal)));
        auto launch = [&](int it) {
            switch (kind) {
                case 0: hipLaunchKernelGGL(issue_rate_kernel<0>, dim3(blocks), dim3(256), 0, 0, it, sink, ticks); break;
                case 1: hipLaunchKernelGGL(issue_rate_kernel<1>, dim3(blocks), dim3(256), 0, 0, it, sink, ticks); break;
                default: hipLaunchKernelGGL(issue_rate_kernel<2>, dim3(blocks), dim3(256), 0, 0, it, sink, ticks); break;
            }
        };
        launch(iters / 4 + 1);
        CLK_TRY(hipGetLastError());
        CLK_TRY(hipEventRecord(e0, 0));
        launch(iters);
        CLK_TRY(hipGetLastError());
        CLK_TRY(hipEventRecord(e1, 0));
        CLK_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        CLK_TRY(hipEventElapsedTime(&ms, e0, e1));
        std::vector<long long> h(4 * (size_t)waves);
        CLK_TRY(hipMemcpy(h.data(), ticks, sizeof(long long) * h.size(), hipMemcpyDeviceToHost));
        std::vector<double> ghz;
        ghz.reserve(waves);
        for (int w = 0; w < waves; ++w) {
            const double wall_s = (double)(h[4 * w + 1] - h[4 * w]) / (khz * 1e3);
            if (wall_s > 0) ghz.push_back((double)(h[4 * w + 3] - h[4 * w + 2]) / wall_s / 1e9);
        }
        if (ghz.empty() || !(ms > 0.0f)) {
            snprintf(msg, sizeof msg, "issue_rate: no timing recorded");
            rc = report_error(ABIE_HIP_EHIP, msg);
            goto done;
        }
        std::sort(ghz.begin(), ghz.end());
        *shader_ghz = ghz[ghz.size() / 2];
        // SIMD cycles in the kernel over the wave instructions each SIMD issued
        // (8 waves x iters x inst)
        *cycles_per_instruction = (ms * 1e-3) * (*shader_ghz * 1e9) / (8.0 * (double)iters * inst);
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (ticks) (void)hipFree(ticks);
    return rc;
}
