// force.hip -- all-pairs Newtonian accelerations on gfx950.
//
// Replaces src/common.c:39-77 (CPU) and src/gpuforce.cu:60-84 (CUDA).
// Layout: sources and sinks in SoA double arrays in HBM.  Sources are the
// massive bodies only (m > 0), compacted in ascending body order, with
// gm = G*m precomputed; sinks [0, n_self) are those same bodies, sinks beyond
// are test particles.  Two kernel forms, both summing each sink's sources in
// ascending order (identical bits):
//  * accel_scalar_kernel (up to 512 sources): one sink per thread; every lane
//    of a wave needs the same source at the same time, so sources come in
//    through scalar loads into SGPRs and each VALU op takes one as its scalar
//    operand -- no LDS, no barrier, 8 waves per SIMD.
//  * accel_kernel (more sources): a 256-thread block owns BLOCK*SPT sinks
//    (SPT in registers, strided by BLOCK so loads and stores coalesce) and
//    streams the sources through LDS in tiles of BLOCK, read as broadcasts.
// Either can split the source range over blockIdx.y when the sinks alone do
// not fill the chip; the partial sums are then added in part order.
//
// Interaction (ABIE_FORCE_FP64), 16 VALU ops + 1 v_rsq_f64:
//   d = p_i - s_k (3)   r2 = |d|^2 (1 mul + 2 fma)
//   y = rsq(r2); e = 1 - r2 y^2; s = gm y^3 (1 + 3/2 e + 15/8 e^2)   (7)
//   a -= s d (3 fma)
// The rsq seed error e enters only at third order (35/16 e^3), so the term is
// accurate to a few ulp as long as the seed is good to ~2^-20.
#include <cstdio>
#include <cstdlib>

#include "devmath.hpp"
#include "force.hpp"

namespace abie {

// Correction coefficients held in registers: gfx950 VOP3 takes no literal,
// and a literal forces a v_fmac with two v_mov per interaction.
struct Coef {
    double c1, c2;  // 3/2, 15/8
};

__device__ __forceinline__ Coef make_coef() {
    Coef k{1.5, 1.875};
    asm volatile("" : "+v"(k.c1), "+v"(k.c2));
    return k;
}

template <int MODE>
struct Interact;

// FP64: rsqrt seed + second-order correction, FMA accumulation.
template <>
struct Interact<ABIE_FORCE_FP64> {
    static __device__ __forceinline__ void run(double px, double py, double pz, double sx, double sy,
                                               double sz, double gm, double &ax, double &ay, double &az, Coef k) {
        const double dx = px - sx, dy = py - sy, dz = pz - sz;
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double y = __builtin_amdgcn_rsq(r2);
        const double y2 = y * y;
        const double e = fma(-r2, y2, 1.0);
        const double g3 = (gm * y) * y2;
        const double p = fma(e, k.c2, k.c1);
        const double s = fma(g3 * e, p, g3);
        ax = fma(-s, dx, ax);
        ay = fma(-s, dy, ay);
        az = fma(-s, dz, az);
    }
};

// Reference operation order, src/common.c:62-70: sqrt, r*r2, (gm*d)/r3.
template <>
struct Interact<ABIE_FORCE_FP64_EXACT> {
    static __device__ __forceinline__ void run(double px, double py, double pz, double sx, double sy,
                                               double sz, double gm, double &ax, double &ay, double &az, Coef) {
        const double dx = px - sx, dy = py - sy, dz = pz - sz;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r3 = sqrt(r2) * r2;
        ax = ax - gm * dx / r3;
        ay = ay - gm * dy / r3;
        az = az - gm * dz / r3;
    }
};

// Test particle from a massive body at single-precision accuracy: the
// hardware reciprocal square root (relative error below 2^-23, measured
// 2^-24.2) used without its correction terms, so r^-3 carries a relative
// error of about 3 * 2^-24.  FP32 arithmetic would not be cheaper: without
// packed FP32 an FP32 op issues at the FP64 rate, and converting the
// coordinates and results costs more ops than the correction saved here
// (12 VALU ops + 1 v_rsq_f64 per pair against 16 + 1 for FP64 mode).
template <>
struct Interact<ABIE_FORCE_MIXED> {
    static __device__ __forceinline__ void run(double px, double py, double pz, double sx, double sy,
                                               double sz, double gm, double &ax, double &ay, double &az, Coef) {
        const double dx = px - sx, dy = py - sy, dz = pz - sz;
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double y = __builtin_amdgcn_rsq(r2);
        const double s = (gm * y) * (y * y);
        ax = fma(-s, dx, ax);
        ay = fma(-s, dy, ay);
        az = fma(-s, dz, az);
    }
};

// Test particle from a massive body in FP32 arithmetic (ABIE_FORCE_FP32):
// FP64 relative coordinates rounded to FP32, v_rsq_f32 + one Newton step,
// FP32 products, the term added to an FP64 sum (more accurate than a
// float-float pair, and one add per component).  Kept to measure against the
// mixed mode: without packed FP32 each FP32 op issues at the FP64 rate, and
// the six conversions make it the slower of the two.
template <>
struct Interact<ABIE_FORCE_FP32> {
    static __device__ __forceinline__ void run(double px, double py, double pz, double sx, double sy,
                                               double sz, double gm, double &ax, double &ay, double &az, Coef) {
        const float dx = (float)(px - sx), dy = (float)(py - sy), dz = (float)(pz - sz);
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        float y = __builtin_amdgcn_rsqf(r2);
        const float hy = 0.5f * y;
        const float e = fmaf(-r2 * y, y, 1.0f);
        y = fmaf(hy, e, y);
        const float t = ((float)gm * y) * (y * y);
        ax = ax - (double)(t * dx);
        ay = ay - (double)(t * dy);
        az = az - (double)(t * dz);
    }
};

constexpr int BLOCK = 256;

// modes whose test-particle terms are reduced precision; their massive sinks
// are evaluated in FP64 mode
constexpr bool tp_reduced(int mode) { return mode == ABIE_FORCE_MIXED || mode == ABIE_FORCE_FP32; }

// Final accelerations are written once and not read again by this launch:
// streaming (non-temporal) stores keep them from piling up as dirty lines
// in L2, which would otherwise be written back at the end of the kernel
// (2-4 % of the launch on 64 + 2^20 sinks).
__device__ __forceinline__ void store_acc(const AccelArgs &a, int i, double x, double y, double z) {
    __builtin_nontemporal_store(x, &a.ax[i]);
    __builtin_nontemporal_store(y, &a.ay[i]);
    __builtin_nontemporal_store(z, &a.az[i]);
}

// Sum over sources [k0, k1) of one LDS tile for SPT sinks.  CHECK: compare
// the source index with each sink's own index (diagonal tiles only).
template <int SPT, int MODE, bool CHECK>
__device__ __forceinline__ void tile_sum(const double *__restrict__ tx, const double *__restrict__ ty,
                                         const double *__restrict__ tz, const double *__restrict__ tg,
                                         int cnt, int kbase, const int (&idx)[SPT],
                                         const double (&px)[SPT], const double (&py)[SPT],
                                         const double (&pz)[SPT], double (&ax)[SPT], double (&ay)[SPT],
                                         double (&az)[SPT], Coef kc) {
    if (!CHECK) {
        // groups of 4 sources, fully unrolled with the SPT sinks: 4*SPT
        // independent chains in flight; then the remainder
        int k = 0;
        for (; k + 4 <= cnt; k += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double sx = tx[k + u], sy = ty[k + u], sz = tz[k + u], g = tg[k + u];
#pragma unroll
                for (int s = 0; s < SPT; ++s)
                    Interact<MODE>::run(px[s], py[s], pz[s], sx, sy, sz, g, ax[s], ay[s], az[s], kc);
            }
        }
        for (; k < cnt; ++k) {
            const double sx = tx[k], sy = ty[k], sz = tz[k], g = tg[k];
#pragma unroll
            for (int s = 0; s < SPT; ++s) Interact<MODE>::run(px[s], py[s], pz[s], sx, sy, sz, g, ax[s], ay[s], az[s], kc);
        }
        return;
    }
#pragma unroll 2
    for (int k = 0; k < cnt; ++k) {
        const double sx = tx[k], sy = ty[k], sz = tz[k], g = tg[k];
#pragma unroll
        for (int s = 0; s < SPT; ++s) {
            if (!CHECK || idx[s] != kbase + k)
                Interact<MODE>::run(px[s], py[s], pz[s], sx, sy, sz, g, ax[s], ay[s], az[s], kc);
        }
    }
}

// One block: BLOCK*SPT sinks against sources [j0, j1).  OUT_PARTIAL writes
// to part[(p*3 + c)*n_sink + i] for the j-split reduction instead of a.
template <int SPT, int MODE, bool OUT_PARTIAL>
__global__ __launch_bounds__(BLOCK) void accel_kernel(AccelArgs a, int j_chunk, double *__restrict__ part) {
    __shared__ double tx[BLOCK], ty[BLOCK], tz[BLOCK], tg[BLOCK];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * BLOCK * SPT;
    const int p = blockIdx.y;
    const int j0 = p * j_chunk;
    const int j1 = min(a.n_src, j0 + j_chunk);

    int idx[SPT];   // own index for sinks that are also sources, else -1
    double px[SPT], py[SPT], pz[SPT], ax[SPT], ay[SPT], az[SPT];
#pragma unroll
    for (int s = 0; s < SPT; ++s) {
        const int i = i0 + s * BLOCK + tid;
        idx[s] = i < a.n_self ? i : -1;
        const int ic = i < a.n_sink ? i : a.n_sink - 1;
        px[s] = a.px[ic]; py[s] = a.py[ic]; pz[s] = a.pz[ic];
        ax[s] = 0.0; ay[s] = 0.0; az[s] = 0.0;
    }
    // MIXED / FP32: blocks holding any massive sink evaluate them all in FP64.
    const bool block_fp64 = tp_reduced(MODE) && (i0 < a.n_self);
    // sinks of this block that are also sources lie in [i0, i_self_end)
    const int i_self_end = min(i0 + BLOCK * SPT, a.n_self);
    const Coef kc = make_coef();

    for (int k0 = j0; k0 < j1; k0 += BLOCK) {
        const int cnt = min(BLOCK, j1 - k0);
        __syncthreads();
        if (tid < cnt) {
            tx[tid] = a.sx[k0 + tid]; ty[tid] = a.sy[k0 + tid];
            tz[tid] = a.sz[k0 + tid]; tg[tid] = a.gm[k0 + tid];
        }
        __syncthreads();
        const bool diag = (i0 < a.n_self) && (k0 < i_self_end) && (k0 + cnt > i0);
        if (tp_reduced(MODE)) {
            if (block_fp64) {
                if (diag) tile_sum<SPT, ABIE_FORCE_FP64, true>(tx, ty, tz, tg, cnt, k0, idx, px, py, pz, ax, ay, az, kc);
                else tile_sum<SPT, ABIE_FORCE_FP64, false>(tx, ty, tz, tg, cnt, k0, idx, px, py, pz, ax, ay, az, kc);
            } else {
                tile_sum<SPT, MODE, false>(tx, ty, tz, tg, cnt, k0, idx, px, py, pz, ax, ay, az, kc);
            }
        } else {
            if (diag) tile_sum<SPT, MODE, true>(tx, ty, tz, tg, cnt, k0, idx, px, py, pz, ax, ay, az, kc);
            else tile_sum<SPT, MODE, false>(tx, ty, tz, tg, cnt, k0, idx, px, py, pz, ax, ay, az, kc);
        }
    }
#pragma unroll
    for (int s = 0; s < SPT; ++s) {
        const int i = i0 + s * BLOCK + tid;
        if (i < a.n_sink) {
            if (OUT_PARTIAL) {
                const size_t base = (size_t)p * 3 * a.n_sink;
                part[base + i] = ax[s];
                part[base + a.n_sink + i] = ay[s];
                part[base + 2 * (size_t)a.n_sink + i] = az[s];
            } else {
                store_acc(a, i, ax[s], ay[s], az[s]);
            }
        }
    }
}

constexpr int kScalarSrcMax = 512;

// Scalar-source form: every lane of a wave needs the same source at the same
// time, so the sources are read with uniform loads straight from memory (the
// compiler turns them into scalar loads through the scalar cache into SGPRs),
// and each VALU op takes the source value as its one scalar operand.  No LDS,
// no barrier; the VGPRs hold only the SPT sinks and their sums.  Same
// per-sink source order as accel_kernel, so identical bits.
// Loads through the constant address space: with a uniform index they are
// selected as scalar loads (s_load) into SGPRs.  The source arrays are not
// written while the kernel runs, which that address space requires.
typedef const __attribute__((address_space(4))) double cdouble;
__device__ __forceinline__ double ldc(const double *p, int k) { return ((cdouble *)p)[k]; }

template <int SPT, int MODE, bool CHECK>
__device__ __forceinline__ void scalar_sum(const double *__restrict__ sx, const double *__restrict__ sy,
                                           const double *__restrict__ sz, const double *__restrict__ sg,
                                           int k0, int k1, const int (&idx)[SPT], const double (&px)[SPT],
                                           const double (&py)[SPT], const double (&pz)[SPT], double (&ax)[SPT],
                                           double (&ay)[SPT], double (&az)[SPT], Coef kc) {
    int k = k0;
    if (!CHECK) {
        for (; k + 4 <= k1; k += 4) {
            double qx[4], qy[4], qz[4], qg[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                qx[u] = ldc(sx, k + u); qy[u] = ldc(sy, k + u); qz[u] = ldc(sz, k + u); qg[u] = ldc(sg, k + u);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int s = 0; s < SPT; ++s)
                    Interact<MODE>::run(px[s], py[s], pz[s], qx[u], qy[u], qz[u], qg[u], ax[s], ay[s], az[s], kc);
        }
    }
    for (; k < k1; ++k) {
        const double qx = ldc(sx, k), qy = ldc(sy, k), qz = ldc(sz, k), qg = ldc(sg, k);
#pragma unroll
        for (int s = 0; s < SPT; ++s)
            if (!CHECK || idx[s] != k)
                Interact<MODE>::run(px[s], py[s], pz[s], qx, qy, qz, qg, ax[s], ay[s], az[s], kc);
    }
}

// Grid (sink blocks, source parts): part p sums sources [p*j_chunk, ...);
// OUT_PARTIAL writes part[(p*3 + c)*n_sink + i] for reduce_partials.
template <int SPT, int MODE, bool OUT_PARTIAL>
__device__ __forceinline__ void accel_scalar_body(const AccelArgs &a, int j_chunk, double *__restrict__ part) {
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * BLOCK * SPT;
    const int p = blockIdx.y;
    const int j0 = p * j_chunk;
    const int j1 = min(a.n_src, j0 + j_chunk);
    const Coef kc = make_coef();
    int idx[SPT];
    double px[SPT], py[SPT], pz[SPT], ax[SPT], ay[SPT], az[SPT];
#pragma unroll
    for (int s = 0; s < SPT; ++s) {
        const int i = i0 + s * BLOCK + tid;
        const int ic = i < a.n_sink ? i : a.n_sink - 1;
        idx[s] = i < a.n_self ? i : -1;
        px[s] = a.px[ic]; py[s] = a.py[ic]; pz[s] = a.pz[ic];
        ax[s] = 0.0; ay[s] = 0.0; az[s] = 0.0;
    }
    // sinks of this block that are sources [i0, i_self_end) meet sources
    // [j0, j1) only when the ranges overlap: then skip k == i
    const int i_self_end = min(i0 + BLOCK * SPT, a.n_self);
    if (i0 < a.n_self && j0 < i_self_end && j1 > i0) {
        constexpr int M = tp_reduced(MODE) ? ABIE_FORCE_FP64 : MODE;
        scalar_sum<SPT, M, true>(a.sx, a.sy, a.sz, a.gm, j0, j1, idx, px, py, pz, ax, ay, az, kc);
    } else if (tp_reduced(MODE) && i0 < a.n_self) {
        scalar_sum<SPT, ABIE_FORCE_FP64, false>(a.sx, a.sy, a.sz, a.gm, j0, j1, idx, px, py, pz, ax, ay, az, kc);
    } else {
        scalar_sum<SPT, MODE, false>(a.sx, a.sy, a.sz, a.gm, j0, j1, idx, px, py, pz, ax, ay, az, kc);
    }
#pragma unroll
    for (int s = 0; s < SPT; ++s) {
        const int i = i0 + s * BLOCK + tid;
        if (i < a.n_sink) {
            if (OUT_PARTIAL) {
                const size_t base = (size_t)p * 3 * a.n_sink;
                part[base + i] = ax[s];
                part[base + a.n_sink + i] = ay[s];
                part[base + 2 * (size_t)a.n_sink + i] = az[s];
            } else {
                store_acc(a, i, ax[s], ay[s], az[s]);
            }
        }
    }
}

template <int SPT, int MODE, bool OUT_PARTIAL>
__global__ __launch_bounds__(BLOCK) void accel_scalar_kernel(AccelArgs a, int j_chunk, double *__restrict__ part) {
    accel_scalar_body<SPT, MODE, OUT_PARTIAL>(a, j_chunk, part);
}

// The one-sink-per-thread scalar-source kernel with each wave's shader cycle
// counter (s_memtime) and constant-rate wall clock (s_memrealtime) read at
// its start and end: ticks[4 w ..] = wall0, wall1, cycles0, cycles1.  The
// clock a VALU-bound kernel runs at depends on its own power draw, so the
// shader clock for pricing this kernel's instruction stream is read inside it.
template <int MODE>
__global__ __launch_bounds__(BLOCK) void accel_scalar_clocked(AccelArgs a, long long *__restrict__ ticks) {
    const long long w0 = wall_clock64(), k0 = clock64();
    accel_scalar_body<1, MODE, false>(a, a.n_src, nullptr);
    const long long w1 = wall_clock64(), k1 = clock64();
    if ((threadIdx.x & 63) == 0) {
        const size_t w = (size_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
        ticks[4 * w + 0] = w0;
        ticks[4 * w + 1] = w1;
        ticks[4 * w + 2] = k0;
        ticks[4 * w + 3] = k1;
    }
}

int clocked_waves(const AccelArgs &a, const AccelPlan &pl) {
    if (!pl.scalar_src || pl.parts != 1 || pl.spt != 1 || a.n_sink <= 0) return 0;
    return (a.n_sink + BLOCK - 1) / BLOCK * (BLOCK / 64);
}

hipError_t launch_accel_clocked(const AccelArgs &a, const AccelPlan &pl, long long *ticks, hipStream_t st) {
    if (clocked_waves(a, pl) == 0) return hipErrorNotSupported;
    const int nblk = (a.n_sink + BLOCK - 1) / BLOCK;
    switch (a.mode) {
        case ABIE_FORCE_FP64:
            hipLaunchKernelGGL(accel_scalar_clocked<ABIE_FORCE_FP64>, dim3(nblk), dim3(BLOCK), 0, st, a, ticks);
            break;
        case ABIE_FORCE_FP64_EXACT:
            hipLaunchKernelGGL(accel_scalar_clocked<ABIE_FORCE_FP64_EXACT>, dim3(nblk), dim3(BLOCK), 0, st, a, ticks);
            break;
        case ABIE_FORCE_MIXED:
            hipLaunchKernelGGL(accel_scalar_clocked<ABIE_FORCE_MIXED>, dim3(nblk), dim3(BLOCK), 0, st, a, ticks);
            break;
        case ABIE_FORCE_FP32:
            hipLaunchKernelGGL(accel_scalar_clocked<ABIE_FORCE_FP32>, dim3(nblk), dim3(BLOCK), 0, st, a, ticks);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Sum the j-split partials in part order 0..P-1 (deterministic).
__global__ __launch_bounds__(BLOCK) void reduce_partials(const double *__restrict__ part, int P, int n,
                                                         double *__restrict__ ax, double *__restrict__ ay,
                                                         double *__restrict__ az) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double sx = part[i], sy = part[n + i], sz = part[2 * (size_t)n + i];
    for (int p = 1; p < P; ++p) {
        const size_t b = (size_t)p * 3 * n;
        sx += part[b + i]; sy += part[b + n + i]; sz += part[b + 2 * (size_t)n + i];
    }
    ax[i] = sx; ay[i] = sy; az[i] = sz;
}

template <int SPT, int MODE>
static hipError_t launch_mode(const AccelArgs &a, const AccelPlan &pl, double *part, hipStream_t st) {
    const int nblk = (a.n_sink + BLOCK * SPT - 1) / (BLOCK * SPT);
    if (pl.scalar_src) {
        if (pl.parts <= 1) {
            hipLaunchKernelGGL((accel_scalar_kernel<SPT, MODE, false>), dim3(nblk, 1), dim3(BLOCK), 0, st, a,
                               a.n_src, nullptr);
        } else {
            const int chunk = (a.n_src + pl.parts - 1) / pl.parts;
            hipLaunchKernelGGL((accel_scalar_kernel<SPT, MODE, true>), dim3(nblk, pl.parts), dim3(BLOCK), 0, st, a,
                               chunk, part);
            hipLaunchKernelGGL(reduce_partials, dim3((a.n_sink + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, part,
                               pl.parts, a.n_sink, a.ax, a.ay, a.az);
        }
    } else if (pl.parts <= 1) {
        hipLaunchKernelGGL((accel_kernel<SPT, MODE, false>), dim3(nblk, 1), dim3(BLOCK), 0, st, a, a.n_src,
                           nullptr);
    } else {
        const int chunk = (a.n_src + pl.parts - 1) / pl.parts;
        hipLaunchKernelGGL((accel_kernel<SPT, MODE, true>), dim3(nblk, pl.parts), dim3(BLOCK), 0, st, a, chunk,
                           part);
        hipLaunchKernelGGL(reduce_partials, dim3((a.n_sink + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, part,
                           pl.parts, a.n_sink, a.ax, a.ay, a.az);
    }
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_spt(const AccelArgs &a, const AccelPlan &pl, double *part, hipStream_t st) {
    switch (pl.spt) {
        case 1: return launch_mode<1, MODE>(a, pl, part, st);
        case 2: return launch_mode<2, MODE>(a, pl, part, st);
        default: return launch_mode<4, MODE>(a, pl, part, st);
    }
}

AccelPlan plan_accel(int n_src, int n_sink, int mode, int n_cu) {
    AccelPlan pl;
    const int want = 4 * n_cu;  // blocks for ~4 resident blocks per CU
    pl.spt = 4;
    while (pl.spt > 1 && (n_sink + BLOCK * pl.spt - 1) / (BLOCK * pl.spt) < want) pl.spt /= 2;
    const int nblk = (n_sink + BLOCK * pl.spt - 1) / (BLOCK * pl.spt);
    pl.parts = 1;
    pl.scalar_src = 0;
    // split the source range when the sinks alone cannot fill the chip
    if (mode != ABIE_FORCE_FP64_EXACT && nblk < want && n_src >= 8 * BLOCK) {
        pl.parts = (want + nblk - 1) / nblk;
        pl.parts = min(pl.parts, n_src / (4 * BLOCK));
        if (pl.parts < 2) pl.parts = 1;
    }
    // Up to kScalarSrcMax sources (32 B each: 16 KB, the scalar cache) the
    // scalar-source kernel with one sink per thread is fastest: measured on
    // 64 + 2^20 test particles 42-45 us against 51-58 us for the LDS-tiled
    // kernel at 4 sinks per thread, and 1.70e12 against 1.56e12
    // interactions/s with 256 sources.  With 1,024 sources and more the
    // LDS-tiled kernel is ahead (197.6 against 204.8 us; 2.6 against 4.0 ms
    // for the 65,536-body Plummer sphere).
    if (pl.parts == 1 && n_src <= kScalarSrcMax) {
        pl.spt = 1;
        pl.scalar_src = 1;
    }
    // ABIE_HIP_ACCEL_PLAN="spt,scalar" overrides the shape (tuning sweeps)
    if (const char *e = getenv("ABIE_HIP_ACCEL_PLAN")) {
        int spt = 0, sc = 0;
        if (sscanf(e, "%d,%d", &spt, &sc) == 2 && (spt == 1 || spt == 2 || spt == 4)) {
            pl.spt = spt;
            pl.scalar_src = sc != 0;
        }
    }
    return pl;
}

hipError_t launch_accel(const AccelArgs &a, const AccelPlan &pl, double *part, hipStream_t st) {
    if (a.n_sink <= 0) return hipSuccess;
    switch (a.mode) {
        case ABIE_FORCE_FP64: return launch_spt<ABIE_FORCE_FP64>(a, pl, part, st);
        case ABIE_FORCE_FP64_EXACT: return launch_spt<ABIE_FORCE_FP64_EXACT>(a, pl, part, st);
        case ABIE_FORCE_MIXED: return launch_spt<ABIE_FORCE_MIXED>(a, pl, part, st);
        case ABIE_FORCE_FP32: return launch_spt<ABIE_FORCE_FP32>(a, pl, part, st);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Host-AoS path (calculate_accelerations): gather bodies into SoA with the
// massive ones first, then scatter accelerations back to AoS.
__global__ void pack_bodies(const double *__restrict__ pos, const double *__restrict__ m,
                            const int *__restrict__ perm, int n, int n_src, double G, double *__restrict__ x,
                            double *__restrict__ y, double *__restrict__ z, double *__restrict__ gm) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = perm[p];
    x[p] = pos[3 * i]; y[p] = pos[3 * i + 1]; z[p] = pos[3 * i + 2];
    if (p < n_src) gm[p] = G * m[i];
}

__global__ void unpack_acc(const int *__restrict__ perm, const double *__restrict__ m, int n,
                           const double *__restrict__ ax, const double *__restrict__ ay,
                           const double *__restrict__ az, const double *__restrict__ ext,
                           double *__restrict__ acc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = perm[p];
    if (m[i] < 0.0) {
        // deleted body: gravity leaves the caller's value (common.c:49-52),
        // the external acceleration is still added to it (common.c:107-111)
        if (ext) {
            acc[3 * i] += ext[3 * i]; acc[3 * i + 1] += ext[3 * i + 1]; acc[3 * i + 2] += ext[3 * i + 2];
        }
        return;
    }
    double vx = ax[p], vy = ay[p], vz = az[p];
    if (ext) {
        vx += ext[3 * i]; vy += ext[3 * i + 1]; vz += ext[3 * i + 2];
    }
    acc[3 * i] = vx; acc[3 * i + 1] = vy; acc[3 * i + 2] = vz;
}

hipError_t launch_pack(const double *pos, const double *m, const int *perm, int n, int n_src, double G,
                       double *x, double *y, double *z, double *gm, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_bodies, dim3((n + 255) / 256), dim3(256), 0, st, pos, m, perm, n, n_src, G, x, y, z, gm);
    return hipGetLastError();
}

hipError_t launch_unpack(const int *perm, const double *m, int n, const double *ax, const double *ay,
                         const double *az, const double *ext, double *acc, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(unpack_acc, dim3((n + 255) / 256), dim3(256), 0, st, perm, m, n, ax, ay, az, ext, acc);
    return hipGetLastError();
}

}  // namespace abie
