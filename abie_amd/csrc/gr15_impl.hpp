// gr15_impl.hpp -- batched, device-resident Gauss-Radau15 (Everhart) integrator.
//
// Restates src/integrator_gauss_radau15.c:9-380 for MI355X: one workgroup
// integrates one system end to end inside the kernel, so a GR15 attempt costs
// no host round trip.  Each thread owns BPT bodies (body b = k*NT + tid) and
// keeps their whole integrator state in registers: y0, dy0, a0 (=ddy0),
// b[7], g[7] and the previous b[6] (the only column of bs0 ever read), 54
// doubles per body.  Massive bodies come first in engine order, so they sit
// in slot 0 of the first waves; at every force evaluation they write their
// predicted positions to an LDS source list (double-buffered, one barrier per
// evaluation) and every thread sums the sources for its bodies from LDS.
//
// Bit-exactness with the reference C (gcc -O3 -std=c99, no contraction):
//  * the library is compiled with -ffp-contract=off and every expression of
//    approx_pos/approx_vel/compute_gs/compute_bs_from_gs/refine_bs keeps the
//    reference's operation order;
//  * divisions by the predictor constants use div_const (correctly rounded);
//  * node 0 reuses a0 instead of a new evaluation: the reference evaluates
//    the force at y(s=0) = y0 again and gets a0 bit for bit;
//  * compute_gs rebuilds all g[0..ih-1] at node ih from unchanged inputs, so
//    only g[ih-1] can differ from the previous node's values: only it is
//    recomputed, by the same recurrence;
//  * vector_max_abs ignores NaNs and starts at -DBL_MAX: fmax reductions do
//    the same; the exponents 1/7 and 1/16 go through pow_pos (double-double).
//  With ABIE_FORCE_FP64_EXACT forces the whole trajectory matches the
//  reference bit for bit (tests/test_gpu_gr15.py).
#pragma once


#include "devmath.hpp"
#include "gr15.hpp"
#include "radau_consts.h"

namespace abie {

constexpr double kH[8] = ABIE_RADAU_H;
constexpr double kR[8][8] = ABIE_RADAU_RS;
constexpr double kC[8][8] = ABIE_RADAU_CS;
// integrator_gauss_radau15.h:17,58-63
constexpr double kEpsb = 1.e-4, kFac = 0.25, kTolpc = 1.e-16, kHmin = 1.e-13;
// The learned proposer's output (log10 of the proposal / dtreq) is clamped to
// [kPropLo, kPropHi].  kPropLo: the logged targets have a spread of ~0.09 dex,
// and a wider range lets an out-of-distribution output hold h well below
// dtreq step after step, which shrinks h without bound (err stays
// ~10^kPropLo, dtreq ~ 10 h).  kPropHi = 0: a proposal never exceeds the
// plain controller's dtreq, so each step is at most as long as the plain
// rule's and the proposer only removes rejected attempts.
constexpr double kPropLo = -0.3, kPropHi = 0.0;

// Every FP64 constant of the step is read from this table through a pointer
// the compiler cannot see through (launder()), so each use is a scalar load
// into SGPRs instead of a value hoisted into VGPRs for the whole kernel
// (gfx950 VOP3 takes no literal; ~80 hoisted constants would cost ~160 VGPRs).
struct KTab {
    double R[8][8], C[8][8];
    double pc[8], prc[8];   // approx_pos divisors 3,6,10,15,21,28,36 and 1/them
    double vc[8], vrc[8];   // approx_vel divisors 2..8 and 1/them
    double h[8];
    double ref[8];          // 7, 6, 5, 4, 3, 2, 21, 15
    double ref2[8];         // 10, 35, 20, 1.5, 1.875, 0.5, 0.25, 0.125
};

constexpr KTab make_ktab() {
    KTab k{};
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
            k.R[i][j] = kR[i][j];
            k.C[i][j] = kC[i][j];
        }
    const double pc[7] = {3, 6, 10, 15, 21, 28, 36}, vc[7] = {2, 3, 4, 5, 6, 7, 8};
    for (int i = 0; i < 7; ++i) {
        k.pc[i] = pc[i];
        k.prc[i] = 1.0 / pc[i];
        k.vc[i] = vc[i];
        k.vrc[i] = 1.0 / vc[i];
    }
    for (int i = 0; i < 8; ++i) k.h[i] = kH[i];
    const double r1[8] = {7, 6, 5, 4, 3, 2, 21, 15}, r2[8] = {10, 35, 20, 1.5, 1.875, 0.5, 0.25, 0.125};
    for (int i = 0; i < 8; ++i) {
        k.ref[i] = r1[i];
        k.ref2[i] = r2[i];
    }
    return k;
}

static __constant__ KTab gK = make_ktab();

// The table through a constant-address-space pointer: a load through it at a
// uniform address is selected as a scalar load (s_load into SGPRs); through a
// generic pointer it would be a vector load whose results occupy VGPRs.
typedef const __attribute__((address_space(4))) KTab CKTab;

__device__ __forceinline__ CKTab *launder() {
    CKTab *p = (CKTab *)&gK;
    asm volatile("" : "+s"(p));
    return p;
}

// The step-size error err = (estim / epsb)^(1/7) (integrator_gauss_radau15.c:316):
// the reference's pow, correctly rounded, under the plain controller (its
// trajectories are the reference's bit for bit); the cheap seventh root
// (root7, an ulp or two) under the massive-only controllers, which are not
// bound to the reference's bits and run the same code in every engine form.
template <int CT>
__device__ __forceinline__ double step_err(double estim) {
    return CT == ABIE_CTRL_PLAIN ? pow_pos(estim / kEpsb, 1. / 7) : root7(estim / kEpsb);
}

// |v_ij|^2 / r_ij^2 of one massive pair, for the proposer's encounter rate
// (the maximum over the pairs): the hardware reciprocal (v_rcp_f64) instead
// of a division, since the feature needs a few digits.  Every engine form
// evaluates a pair with these same operations, so the maximum does not
// depend on which lane or in which order the pairs are visited.
__device__ __forceinline__ double pair_rate2(double dx, double dy, double dz, double vx, double vy, double vz) {
    const double r2 = dx * dx + dy * dy + dz * dz, v2 = vx * vx + vy * vy + vz * vz;
    return v2 * __builtin_amdgcn_rcp(r2);
}

// -------------------------------------------------------------- predictors
// approx_pos (integrator_gauss_radau15.c:56-61) at node s, T = dt*s.
// approx_pos (integrator_gauss_radau15.c:56-61).  FAST (the massive-only
// controllers, which do not reproduce the reference bit for bit): products
// with the reciprocals and fused multiply-adds instead of the correctly
// rounded divisions and separate roundings -- 17 instead of 38 operations
// per component, within an ulp or two of the reference form.
template <bool FAST>
__device__ __forceinline__ double pred_pos(CKTab *K, double y0, double dy0, double a0, const double (&b)[7],
                                           double s, double T) {
    if (FAST) {
        double p = fma(s * b[6], K->prc[6], b[5] * K->prc[5]);
        p = fma(s, p, b[4] * K->prc[4]);
        p = fma(s, p, b[3] * K->prc[3]);
        p = fma(s, p, b[2] * K->prc[2]);
        p = fma(s, p, b[1] * K->prc[1]);
        p = fma(s, p, b[0] * K->prc[0]);
        p = fma(s, p, a0);
        return fma(T, fma(T * p, K->ref2[5], dy0), y0);
    }
    double p = div_const(b[5], K->pc[5], K->prc[5]) + div_const(s * b[6], K->pc[6], K->prc[6]);
    p = div_const(b[4], K->pc[4], K->prc[4]) + s * p;
    p = div_const(b[3], K->pc[3], K->prc[3]) + s * p;
    p = div_const(b[2], K->pc[2], K->prc[2]) + s * p;
    p = div_const(b[1], K->pc[1], K->prc[1]) + s * p;
    p = div_const(b[0], K->pc[0], K->prc[0]) + s * p;
    p = a0 + s * p;
    return y0 + T * (dy0 + (T * p) * K->ref2[5]);
}

// approx_vel (integrator_gauss_radau15.c:63-68) at node s = 1.
template <bool FAST>
__device__ __forceinline__ double pred_vel1(CKTab *K, double dy0, double a0, const double (&b)[7], double dt) {
    if (FAST) {
        double p = fma(b[6], K->ref2[7], b[5] * K->vrc[5]);
        p = fma(b[4], K->vrc[4], p);
        p = fma(b[3], K->vrc[3], p);
        p = fma(b[2], K->ref2[6], p);
        p = fma(b[1], K->vrc[1], p);
        p = fma(b[0], K->ref2[5], p);
        return fma(dt, a0 + p, dy0);
    }
    double p = div_const(b[5], K->vc[5], K->vrc[5]) + b[6] * K->ref2[7];
    p = div_const(b[4], K->vc[4], K->vrc[4]) + p;
    p = div_const(b[3], K->vc[3], K->vrc[3]) + p;
    p = b[2] * K->ref2[6] + p;
    p = div_const(b[1], K->vc[1], K->vrc[1]) + p;
    p = b[0] * K->ref2[5] + p;
    p = a0 + p;
    return dy0 + dt * p;
}

// compute_gs + compute_bs_from_gs at node IH (1..7) for one component; FAST
// fuses the b sums.
template <int IH, bool FAST>
__device__ __forceinline__ void update_gb(CKTab *K, double F, double a0, double (&g)[7], double (&b)[7]) {
    constexpr int k = IH - 1;
    double v = (F - a0) * K->R[k + 1][0];
#pragma unroll
    for (int l = 1; l <= k; ++l) v = (v - g[l - 1]) * K->R[k + 1][l];
    g[k] = v;
#pragma unroll
    for (int kk = 0; kk < IH; ++kk) {
        double s = g[kk];  // cs[kk][kk] == 1 exactly
#pragma unroll
        for (int j = kk + 1; j < 7; ++j) s = FAST ? fma(K->C[j][kk], g[j], s) : s + K->C[j][kk] * g[j];
        b[kk] = s;
    }
}

// refine_bs (integrator_gauss_radau15.c:178-218) for one component.
__device__ __forceinline__ void refine(CKTab *K, double (&b)[7], double q) {
    const double q2 = q * q, q3 = q2 * q, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q2 * q5;
    const double c7 = K->ref[0], c6 = K->ref[1], c5 = K->ref[2], c4 = K->ref[3], c3 = K->ref[4], c2 = K->ref[5];
    const double c21 = K->ref[6], c15 = K->ref[7], c10 = K->ref2[0], c35 = K->ref2[1], c20 = K->ref2[2];
    double e[7];
    e[0] = q * (b[6] * c7 + b[5] * c6 + b[4] * c5 + b[3] * c4 + b[2] * c3 + b[1] * c2 + b[0]);
    e[1] = q2 * (b[6] * c21 + b[5] * c15 + b[4] * c10 + b[3] * c6 + b[2] * c3 + b[1]);
    e[2] = q3 * (b[6] * c35 + b[5] * c20 + b[4] * c10 + b[3] * c4 + b[2]);
    e[3] = q4 * (b[6] * c35 + b[5] * c15 + b[4] * c5 + b[3]);
    e[4] = q5 * (b[6] * c21 + b[5] * c6 + b[4]);
    e[5] = q6 * (b[6] * c7 + b[5]);
    e[6] = q7 * b[6];
#pragma unroll
    for (int j = 0; j < 7; ++j) b[j] = e[j] + 0.0;
}

// -------------------------------------------------------------- pair terms
struct Coef2 {
    double c1, c2;
};

template <int FM>
__device__ __forceinline__ void pair(double px, double py, double pz, const double *s, double &ax, double &ay,
                                     double &az, Coef2 k) {
    const double dx = px - s[0], dy = py - s[1], dz = pz - s[2];
    if (FM == ABIE_FORCE_FP64_EXACT) {
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r3 = sqrt(r2) * r2;
        ax = ax - s[3] * dx / r3;
        ay = ay - s[3] * dy / r3;
        az = az - s[3] * dz / r3;
    } else if (FM == ABIE_FORCE_FP64) {
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double y = __builtin_amdgcn_rsq(r2);
        const double y2 = y * y;
        const double e = fma(-r2, y2, 1.0);
        const double g3 = (s[3] * y) * y2;
        const double t = fma(g3 * e, fma(e, k.c2, k.c1), g3);
        ax = fma(-t, dx, ax);
        ay = fma(-t, dy, ay);
        az = fma(-t, dz, az);
    } else {  // MIXED: uncorrected hardware rsq, single-precision accuracy
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double y = __builtin_amdgcn_rsq(r2);
        const double t = (s[3] * y) * (y * y);
        ax = fma(-t, dx, ax);
        ay = fma(-t, dy, ay);
        az = fma(-t, dz, az);
    }
}

// The same pair with the self-pair (source index == own) removed: 1.0 is
// added to r^2 of the self-pair only (d = 0 there, so its term is an exact
// zero and acc is unchanged; for every other pair r^2 + 0.0 is r^2 bit for
// bit), which costs a compare and a select per pair and keeps rsq(0) = inf
// out of the arithmetic.
template <int FM>
__device__ __forceinline__ void pair_self(double px, double py, double pz, const double *s, bool self, double &ax,
                                          double &ay, double &az, Coef2 k) {
    const double dx = px - s[0], dy = py - s[1], dz = pz - s[2];
    const double one = self ? 1.0 : 0.0;
    if (FM == ABIE_FORCE_FP64_EXACT) {
        const double r2 = (dx * dx + dy * dy + dz * dz) + one;
        const double r3 = sqrt(r2) * r2;
        ax = ax - s[3] * dx / r3;
        ay = ay - s[3] * dy / r3;
        az = az - s[3] * dz / r3;
    } else {  // massive sinks always use FP64
        const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, one)));
        const double y = __builtin_amdgcn_rsq(r2);
        const double y2 = y * y;
        const double e = fma(-r2, y2, 1.0);
        const double g3 = (s[3] * y) * y2;
        const double t = fma(g3 * e, fma(e, k.c2, k.c1), g3);
        ax = fma(-t, dx, ax);
        ay = fma(-t, dy, ay);
        az = fma(-t, dz, az);
    }
}

// The pair term for a wave that holds massive bodies, without divergence:
// self removes the self-pair (1.0 added to r^2, where d = 0, so the term is
// an exact zero), corr applies the FP64 correction of the rsqrt seed.  With
// self false the bits equal pair<ABIE_FORCE_FP64> (fma(dx, dx, 0) rounds as
// dx * dx); with corr false they equal pair<ABIE_FORCE_MIXED>.  So a test
// particle gets the same term whether or not its wave holds massive bodies.
template <int FM>
__device__ __forceinline__ void pair_sel(double px, double py, double pz, const double *s, bool self, bool corr,
                                         double &ax, double &ay, double &az, Coef2 k) {
    const double dx = px - s[0], dy = py - s[1], dz = pz - s[2];
    const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, self ? 1.0 : 0.0)));
    const double y = __builtin_amdgcn_rsq(r2);
    const double y2 = y * y;
    const double g3 = (s[3] * y) * y2;
    const double e = fma(-r2, y2, 1.0);
    const double tc = fma(g3 * e, fma(e, k.c2, k.c1), g3);
    const double t = (FM == ABIE_FORCE_FP64 || corr) ? tc : g3;
    ax = fma(-t, dx, ax);
    ay = fma(-t, dy, ay);
    az = fma(-t, dz, az);
}

// The force sum of the massive-only controllers' units of up to 256 bodies,
// in both engine forms (one lane per body, gr15_kernel; four lanes per body,
// gr15q_kernel): the sources are split by j mod 4 into four partial sums
// P_c over j = c, c + 4, ... in ascending order, combined as
// (P0 + P1) + (P2 + P3).  Both forms therefore give the same bits, and a
// unit may move from one to the other between launches.
//
// sum_stride4 adds P_c at position p to (ax, ay, az), 2 P sources per
// iteration in two register sets of P that alternate: the next P sources are
// read from LDS while the current P are summed (LDS reads return in order,
// so each wait covers only the older reads).  The order of the terms, and so
// the bits, are those of a plain loop whatever P is.  P = 2 for the quad
// form's groups of up to 512 threads (a lone wave per SIMD: the pairs in
// flight hide FP64 latency; 4 was measured 10 % slower), P = 1 for the
// one-lane form and the 1,024-thread quad group (three or four waves per
// SIMD hide it, and the freed registers spill less: 1-2 % more ensemble
// throughput, 5 % lower latency of 256 such units at once).  MW: the wave holds massive bodies (pair_sel with self =
// (j == me), corr = massive).  The reads ahead are clamped to the last
// source, so none leaves the list.
template <int FM, bool MW, int P = 2>
__device__ __forceinline__ void sum_stride4(const double *S, int c, int nm, int me, bool massive, double px,
                                            double py, double pz, double &ax, double &ay, double &az, Coef2 kc) {
    const int cnt = nm > c ? (nm - c + 3) >> 2 : 0;
    const double *Sp = S + 4 * c;  // source c + 4 k at Sp + 16 k
    auto term = [&](const double (&q)[4], int k) __attribute__((always_inline)) {
        if (MW) pair_sel<FM>(px, py, pz, q, c + 4 * k == me, massive, ax, ay, az, kc);
        else pair<FM>(px, py, pz, q, ax, ay, az, kc);
    };
    auto load = [&](double (&q)[4], int k) __attribute__((always_inline)) {
        const int kk = k < cnt ? k : (cnt > 0 ? cnt - 1 : 0);
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = Sp[16 * kk + u];
    };
    double A[P][4], B[P][4];
#pragma unroll
    for (int i = 0; i < P; ++i) load(A[i], i);
    int k = 0;
    for (; k + 2 * P <= cnt; k += 2 * P) {
#pragma unroll
        for (int i = 0; i < P; ++i) load(B[i], k + P + i);
#pragma unroll
        for (int i = 0; i < P; ++i) term(A[i], k + i);
#pragma unroll
        for (int i = 0; i < P; ++i) load(A[i], k + 2 * P + i);
#pragma unroll
        for (int i = 0; i < P; ++i) term(B[i], k + P + i);
    }
    // the remaining 0 .. 2P - 1 sources: A holds k .. k + P - 1
#pragma unroll
    for (int i = 0; i < P; ++i)
        if (k + i < cnt) term(A[i], k + i);
#pragma unroll
    for (int i = 0; i < P - 1; ++i)
        if (k + P + i < cnt) {
            load(B[i], k + P + i);
            term(B[i], k + P + i);
        }
}

// -------------------------------------------------------------- reductions
// v from another lane of the wave by a DPP pattern (both 32-bit halves
// through v_mov_dpp).  Every lane of the wave must be active: the patterns
// used (quad_perm, row_ror) then read a valid lane everywhere, so no "old"
// value is needed and the compiler copies nothing into the destination
// first (2 moves per 64-bit step fewer than with update_dpp).
template <int CTRL>
__device__ __forceinline__ double dpp64(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = (int)(unsigned)(u & 0xffffffffll), hi = (int)(u >> 32);
    const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
    const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)rhi << 32) | (long long)(unsigned)rlo);
}

// v of lane l (a wave-uniform value)
__device__ __forceinline__ double readlane64(double v, int l) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Max over the 64 lanes of a wave: within each row of 16 lanes by DPP
// (quad_perm [1,0,3,2] and [2,3,0,1], then row_ror 4 and 8), then over the
// rows' lanes 0, 16, 32, 48 by v_readlane -- no LDS round trips.  A max is
// exact, so any order gives the same value; NaNs lose, as with fmax.
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp64<0xB1>(v));
    v = fmax(v, dpp64<0x4E>(v));
    v = fmax(v, dpp64<0x124>(v));
    v = fmax(v, dpp64<0x128>(v));
    return fmax(fmax(readlane64(v, 0), readlane64(v, 16)), fmax(readlane64(v, 32), readlane64(v, 48)));
}

// Workgroup max of two values over the group's first nw waves (all of them
// unless waves without bodies have ended); NaNs lose (fmax), as in
// vector_max_abs.  Groups of up to 512 threads alternate between two halves
// of red[0, 32) (par flips each call, the same in every thread): a wave
// that writes this call's half cannot overwrite values another wave still
// reads from the previous call's, and the half it reuses was last read
// before the previous call's barrier -- one barrier per call instead of two
// (lone units in the quad form, split: 1.31-1.74 -> 1.29-1.73 us per
// evaluation; the one-lane throughput unchanged within noise).
template <int NT>
__device__ __forceinline__ void wg_max2(double &a, double &b, double *red, int &par, int nw = NT / 64) {
    a = wave_max(a);
    b = wave_max(b);
    const int w = threadIdx.x >> 6;
    if (NT > 512) {
        __syncthreads();
    } else {
        red += 16 * par;
        par ^= 1;
    }
    if ((threadIdx.x & 63) == 0) {
        red[2 * w] = a;
        red[2 * w + 1] = b;
    }
    __syncthreads();
    a = red[0];
    b = red[1];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) {
        if (i < nw) {
            a = fmax(a, red[2 * i]);
            b = fmax(b, red[2 * i + 1]);
        }
    }
}

// wg_max2, then (linked units) the maxima over every slice of the system.
// Thread 0 publishes the slice's pair in the parity-(epoch & 1) half of
// link_part, releases it, arrives at the system's counter and waits until all
// link_n slices have arrived for this epoch (the counter only grows:
// (epoch + 1) * link_n arrivals).  Two parities suffice: a slice can only
// write the next epoch's pair after every slice arrived at this epoch, i.e.
// after it read this one.  The wait is bounded by A.spin_ticks of the
// device's constant-rate clock; false on a timeout (the caller stops the
// unit with ST_ERROR).  Every slice makes the same sequence of calls.
// LDS: red[48] carries the wait's outcome, red[49..50] the maxima.
template <int NT>
__device__ __forceinline__ bool red_max2(const GR15Args &A, const SysDesc &d, long long &epoch, double &a, double &b, double *red,
                                         int &par) {
    wg_max2<NT>(a, b, red, par);
    if (d.link_n <= 1) return true;
    const int tid = threadIdx.x;
    double *P = A.link_part + ((long long)(epoch & 1) * A.n_units + d.link_base) * 2;
    unsigned long long *cnt = A.link_count + d.link_base;
    if (tid == 0) {
        P[2 * d.link_rank] = a;
        P[2 * d.link_rank + 1] = b;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(cnt, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long target = (unsigned long long)(epoch + 1) * (unsigned long long)d.link_n;
        const long long t0 = wall_clock64();
        double ok = 1.0;
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (wall_clock64() - t0 > A.spin_ticks) {
                ok = 0.0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        red[48] = ok;
    }
    __syncthreads();
    // every wave: the other slices' writes (pairs, event-scan positions) are
    // visible to the loads that follow
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const bool ok = red[48] != 0.0;
    if (tid < 64) {
        double x = -DBL_MAX, y = -DBL_MAX;
        for (int u = tid; u < d.link_n; u += 64) {
            x = fmax(x, P[2 * u]);
            y = fmax(y, P[2 * u + 1]);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            x = fmax(x, __shfl_xor(x, off));
            y = fmax(y, __shfl_xor(y, off));
        }
        if (tid == 0) {
            red[49] = x;
            red[50] = y;
        }
    }
    __syncthreads();
    a = red[49];
    b = red[50];
    epoch += 1;
    return ok;
}

// -------------------------------------------------------------- the kernel
template <int NT, int BPT>
struct Bodies {
    double y0[BPT][3], dy0[BPT][3], a0[BPT][3], bb[BPT][3][7], gg[BPT][3][7], b6p[BPT][3];
};

template <int NT, int BPT, int FM, int CT>
struct Engine {
    const GR15Args &A;
    const int sys, body0, n, nm;
    double *src;  // LDS [2][NT][4]
    double *red;  // LDS [32]
    int buf = 0;
    Coef2 kc;
    int slot_valid[BPT], slot_massive[BPT], slot_del[BPT], slot_orig[BPT];
    double gm0;  // G m of this thread's slot-0 body when massive (massive bodies sit in slot 0)
    bool norm_all;
    bool wave_massive;

    __device__ Engine(const GR15Args &a, int s, const SysDesc &d, double *lds)
        : A(a), sys(s), body0(d.body0), n(d.n), nm(d.nm), src(lds), red(lds + 2 * src_cap(NT) * 4) {
        kc = Coef2{1.5, 1.875};
        asm volatile("" : "+v"(kc.c1), "+v"(kc.c2));
        wave_massive = ((int)(threadIdx.x & ~63) < nm);
#pragma unroll
        for (int k = 0; k < BPT; ++k) {
            const int b = k * NT + threadIdx.x;
            slot_valid[k] = b < n;
            slot_massive[k] = b < nm;
            slot_del[k] = (b < n) ? (int)A.deleted[body0 + b] : 0;
            slot_orig[k] = (b < n) ? A.orig[body0 + b] : 0;
        }
        gm0 = ((int)threadIdx.x < nm) ? A.gm[body0 + threadIdx.x] : 0.0;
        norm_all = d.norm_all != 0;
    }

    // The source sum for units with many massive bodies: sources in pairs,
    // the next pair read from LDS while the current one is used (LDS reads
    // return in order, so the wait covers only the older pair).  Pair order
    // and arithmetic are those of the plain loop, so the sums are the same
    // bits.  With two pairs per iteration a lone 256-body unit evaluates
    // 11-21 % faster (12-58 massive bodies) -- what sets an ensemble's
    // makespan, the few close-encounter systems with the most steps -- and
    // 2,048 such units 2-8 % faster (massive-only controllers).  The plain
    // controller keeps the plain loop: compiled into its kernels this loop
    // made the configuration-0 drop-in (5 massive bodies, which never use
    // it) 7 % slower, and the whole-system (512, 2..5) shapes of the plain
    // ensemble 30 % slower, for 16-21 % on lone one-body-per-thread units
    // with 8+ massive bodies.
    __device__ __forceinline__ void sum_prefetched(const double *S, const double (&p)[BPT][3],
                                                   double (&acc)[BPT][3]) {
        const int me = wave_massive ? (int)threadIdx.x : -1;
        int j = 0;
        double c0[4], c1[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            c0[q] = S[q];
            c1[q] = S[4 + q];
        }
        // two pairs of sources per iteration: the scheduler overlaps their
        // dependency chains (a lone 58 + 198 unit 7.3 -> 6.8 us per
        // evaluation, 2,048 of them +8 % evaluations per second; same bits)
#pragma unroll 2
        for (; j + 2 <= nm; j += 2) {
            const int jn = (j + 4 <= nm) ? j + 2 : j;
            double n0[4], n1[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                n0[q] = S[4 * jn + q];
                n1[q] = S[4 * jn + 4 + q];
            }
            if (wave_massive) {
                pair_self<FM>(p[0][0], p[0][1], p[0][2], c0, j == me, acc[0][0], acc[0][1], acc[0][2], kc);
#pragma unroll
                for (int k = 1; k < BPT; ++k) pair<FM>(p[k][0], p[k][1], p[k][2], c0, acc[k][0], acc[k][1], acc[k][2], kc);
                pair_self<FM>(p[0][0], p[0][1], p[0][2], c1, j + 1 == me, acc[0][0], acc[0][1], acc[0][2], kc);
#pragma unroll
                for (int k = 1; k < BPT; ++k) pair<FM>(p[k][0], p[k][1], p[k][2], c1, acc[k][0], acc[k][1], acc[k][2], kc);
            } else {
#pragma unroll
                for (int k = 0; k < BPT; ++k) pair<FM>(p[k][0], p[k][1], p[k][2], c0, acc[k][0], acc[k][1], acc[k][2], kc);
#pragma unroll
                for (int k = 0; k < BPT; ++k) pair<FM>(p[k][0], p[k][1], p[k][2], c1, acc[k][0], acc[k][1], acc[k][2], kc);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                c0[q] = n0[q];
                c1[q] = n1[q];
            }
        }
        if (j < nm) {
            const double *s = S + 4 * j;
            if (wave_massive) {
                pair_self<FM>(p[0][0], p[0][1], p[0][2], s, j == me, acc[0][0], acc[0][1], acc[0][2], kc);
#pragma unroll
                for (int k = 1; k < BPT; ++k) pair<FM>(p[k][0], p[k][1], p[k][2], s, acc[k][0], acc[k][1], acc[k][2], kc);
            } else {
#pragma unroll
                for (int k = 0; k < BPT; ++k) pair<FM>(p[k][0], p[k][1], p[k][2], s, acc[k][0], acc[k][1], acc[k][2], kc);
            }
        }
    }

    // accelerations at positions p for every owned body; one barrier
    __device__ void forces(const double (&p)[BPT][3], double (&acc)[BPT][3]) {
        double *S = src + buf * src_cap(NT) * 4;
        buf ^= 1;
#pragma unroll
        for (int k = 0; k < BPT; ++k) {
            const int b = k * NT + threadIdx.x;
            if (b < nm) {
                S[4 * b] = p[k][0];
                S[4 * b + 1] = p[k][1];
                S[4 * b + 2] = p[k][2];
                S[4 * b + 3] = gm0;  // b < nm <= NT: slot 0
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < BPT; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
        if (CT != ABIE_CTRL_PLAIN && BPT == 1) {
            // the four-way summation tree shared with the quad form
            // (P0 + P1) + (P2 + P3) without selects on c: r starts at -0.0
            // (-0.0 + x is x, bit for bit), an even c keeps its partial sum
            // in h, an odd one adds (h + P_c) to r (selects on c cost ~150
            // register moves per evaluation: lone one-lane units of 32-52
            // massive bodies 4.0-5.9 -> 3.3-5.2 us per evaluation, 4,096
            // systems 0.2 yr 0.222 -> 0.208 s)
            const int me = threadIdx.x;
            const bool mass = me < nm;
            double r[3] = {-0.0, -0.0, -0.0}, h[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
            for (int c = 0; c < 4; ++c) {
                double a[3] = {0.0, 0.0, 0.0};
                if (wave_massive) sum_stride4<FM, true, 1>(S, c, nm, me, mass, p[0][0], p[0][1], p[0][2], a[0], a[1], a[2], kc);
                else sum_stride4<FM, false, 1>(S, c, nm, me, false, p[0][0], p[0][1], p[0][2], a[0], a[1], a[2], kc);
                if (c & 1) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) r[q] = r[q] + (h[q] + a[q]);
                } else {
#pragma unroll
                    for (int q = 0; q < 3; ++q) h[q] = a[q];
                }
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[0][q] = r[q];
        } else if (CT != ABIE_CTRL_PLAIN && A.prefetch > 0 && nm >= A.prefetch) {
            sum_prefetched(S, p, acc);
        } else if (wave_massive) {
            // slot 0 holds massive bodies (nm <= NT): FP64 with the self-pair removed
            const int me = threadIdx.x;
            for (int j = 0; j < nm; ++j) {
                const double *s = S + 4 * j;
                pair_self<FM>(p[0][0], p[0][1], p[0][2], s, j == me, acc[0][0], acc[0][1], acc[0][2], kc);
#pragma unroll
                for (int k = 1; k < BPT; ++k) pair<FM>(p[k][0], p[k][1], p[k][2], s, acc[k][0], acc[k][1], acc[k][2], kc);
            }
        } else {
#pragma unroll 2
            for (int j = 0; j < nm; ++j) {
                const double *s = S + 4 * j;
#pragma unroll
                for (int k = 0; k < BPT; ++k) pair<FM>(p[k][0], p[k][1], p[k][2], s, acc[k][0], acc[k][1], acc[k][2], kc);
            }
        }
#pragma unroll
        for (int k = 0; k < BPT; ++k) {
            const int b = k * NT + threadIdx.x;
            if (slot_del[k]) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
            if (A.ext != nullptr && b < n) {
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[k][c] += A.ext[c * A.total + body0 + b];
            }
        }
    }

    // which components enter the controller's norms
    __device__ __forceinline__ bool in_norm(int k) const {
        return (CT == ABIE_CTRL_PLAIN || norm_all) ? (bool)slot_valid[k] : (bool)slot_massive[k];
    }
};

// LDS scratch (doubles after the source lists, gr15_lds_bytes): [0, 32)
// reductions, [40, 48) flags broadcast by thread 0, [48, 51) linked-slice
// maxima, [64, 96) the proposer's hidden layer (64 floats), [96, 109) the
// proposer's inputs and output kept for the step log.
constexpr int kLdsFeat = 96;

__device__ __forceinline__ float mlp_eval(const float *__restrict__ w, const float *x, float *lds_h) {
    // 12 -> 64 (ReLU) -> 64 (ReLU) -> 1; one wave, lane = hidden unit; w is
    // the weights' copy in LDS
    const int lane = threadIdx.x & 63;
    const float *W1 = w, *b1 = W1 + ABIE_MLP_IN * ABIE_MLP_HIDDEN;
    const float *W2 = b1 + ABIE_MLP_HIDDEN, *b2 = W2 + ABIE_MLP_HIDDEN * ABIE_MLP_HIDDEN;
    const float *w3 = b2 + ABIE_MLP_HIDDEN, *b3 = w3 + ABIE_MLP_HIDDEN;
    // short unrolls: the GR15 state of the caller is live across this call
    // and a fully unrolled 64-term loop would spill it
    float h = b1[lane];
#pragma unroll 4
    for (int i = 0; i < ABIE_MLP_IN; ++i) h = fmaf(x[i], W1[i * ABIE_MLP_HIDDEN + lane], h);
    h = fmaxf(h, 0.0f);
    lds_h[lane] = h;
    __builtin_amdgcn_wave_barrier();
    float h2 = b2[lane];
#pragma unroll 8
    for (int i = 0; i < ABIE_MLP_HIDDEN; ++i) h2 = fmaf(lds_h[i], W2[i * ABIE_MLP_HIDDEN + lane], h2);
    h2 = fmaxf(h2, 0.0f);
    float o = h2 * w3[lane];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) o += __shfl_xor(o, off);
    return o + b3[0];
}

// What the learned proposer's inputs are formed from, after an accepted
// attempt: this attempt's h, error and PC count, the rejections before it,
// the previous accepted step's h, error, dtreq and PC count, the encounter
// rate, the unit's massive bodies, its system's test particles and its mass
// ratio.
struct PropIn {
    double h_att, err, dtreq, prev_h, prev_err, prev_dtreq, rate2, mratio;
    int pc_used, prev_pc, rej_before, nm;
    long long nt_sys;
};

// The proposer's output for an accepted attempt, by the calling (first)
// wave: lane i < 12 forms input i -- one log10 per lane, the same operations
// on the same values as the inputs of the training code -- the inputs are
// broadcast to the wave by v_readlane, and mlp_eval runs the network.
// Returns the output in every lane and lane i's input in *xi (0 beyond 12).
// Inputs (all invariant under a change of time unit): log10 h/dtreq_prev,
// log10 err, log10 err_prev, log10 h/h_prev, pc/12, log10 h sqrt(rate2),
// log10 N_m, log10(N_t + 1) (clamped to the training range 0..2,048: test
// particles never enter the massive-only error estimate, so a larger count
// says nothing more about the step), log10(1 + rejections), log10
// dtreq/dtreq_prev, pc_prev/12, log10 of the mass ratio.
__device__ __forceinline__ float proposer_eval(const float *__restrict__ w, const PropIn &f, float *lds_h,
                                               float *xi) {
    const int i = threadIdx.x & 63;
    const double eps = 1e-300;
    const double pd = f.prev_dtreq;
    double num, den = 1.0;
    switch (i) {
        case 0: num = pd > 0 ? f.h_att : 1.0; den = pd > 0 ? pd : 1.0; break;
        case 1: num = f.err + eps; break;
        case 2: num = f.prev_err + eps; break;
        case 3: num = f.h_att; den = f.prev_h > 0 ? f.prev_h : f.h_att; break;
        case 5: num = f.h_att * sqrt(f.rate2) + 1e-30; break;
        case 6: num = (double)f.nm; break;
        case 7: num = (double)f.nt_sys + 1.0; break;
        case 8: num = 1.0 + (double)f.rej_before; break;
        case 9: num = pd > 0 ? f.dtreq : 1.0; den = pd > 0 ? pd : 1.0; break;
        case 11: num = f.mratio + 1e-12; break;
        default: num = 1.0; break;
    }
    const double l = log10(num / den);  // x / 1.0 is x: the unscaled inputs keep their bits
    float x = (float)l;
    if (i == 4) x = (float)f.pc_used / 12.0f;
    if (i == 7) x = (float)fmin(l, log10(2049.0));
    if (i == 10) x = (float)f.prev_pc / 12.0f;
    if (i >= ABIE_MLP_IN) x = 0.0f;
    *xi = x;
    float xs[ABIE_MLP_IN];
#pragma unroll
    for (int k = 0; k < ABIE_MLP_IN; ++k) xs[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
    return mlp_eval(w, xs, lds_h);
}

// Close-encounter / collision filter of one body against LDS tile entries
// [q0, cnt): true if some pair passes r^2 <= max(cd^2, (r_j + r_q)^2) with a
// 2^-46 relative slack, a superset of the exact test's d <= cd or d <= r_j +
// r_q: d = sqrt(r^2) rounds, and r^2 with FMAs and the squared thresholds
// differ from the exact pass's values by a few ulp (2^-50 would be tight).  RAD = false:
// no body of the unit has a radius, the threshold is cd^2 for every pair.
// self >= 0: the diagonal tile, only entries q > self count.
template <bool RAD>
__device__ __forceinline__ bool scan_tile(const double *tile, int q0, int cnt, int self, const double (&y)[3],
                                          double rj, double thr0, double cd2) {
    bool any = false;
    for (int q = q0; q < cnt; ++q) {
        const double dx = y[0] - tile[4 * q], dy = y[1] - tile[4 * q + 1], dz = y[2] - tile[4 * q + 2];
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        double thr = thr0;
        if (RAD) {
            const double rr = rj + tile[4 * q + 3];
            thr = (rr > 0 && thr0 >= 0) ? fmax(cd2, rr * rr * (1.0 + 0x1p-46)) : thr0;
        }
        any |= (r2 <= thr) && q > self;
    }
    return any;
}

// scan_tile's test again, for a wave that found a hit in the tile: every
// pair that passes flags both its bodies (slot sj, and qbase + q for tile
// entry q) in sf, so that the exact pass visits flagged bodies only (an
// event pair passes this superset test).  Several lanes may store the same
// 1.0 to one flag.
template <bool RAD>
__device__ __forceinline__ void mark_tile(const double *tile, int q0, int cnt, int self, const double (&y)[3],
                                          double rj, double thr0, double cd2, double *sf, long long sj,
                                          long long qbase) {
    for (int q = q0; q < cnt; ++q) {
        const double dx = y[0] - tile[4 * q], dy = y[1] - tile[4 * q + 1], dz = y[2] - tile[4 * q + 2];
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        double thr = thr0;
        if (RAD) {
            const double rr = rj + tile[4 * q + 3];
            thr = (rr > 0 && thr0 >= 0) ? fmax(cd2, rr * rr * (1.0 + 0x1p-46)) : thr0;
        }
        if ((r2 <= thr) && q > self) {
            sf[sj] = 1.0;
            sf[qbase + q] = 1.0;
        }
    }
}

// x-slab prefilter of scan_tile<false>, two VALU per pair: lanes of the wave
// with |x_j - x_q| <= cx at some q in [q0, cnt) (q > self on the diagonal
// tile).  r^2 = fma(dz, dz, fma(dy, dy, dx dx)) >= round(dx dx), so a pair
// with r^2 <= cd2 has round(dx dx) <= cd2 and |dx| <= cx when cx =
// sqrt(cd2) (1 + 2^-48) + 2^-500 (the last term covers a dx dx that rounds
// to zero or to a subnormal): a zero mask means scan_tile<false> over the
// same tile finds nothing.  cx < 0 on idle lanes, NaN never passes.
template <bool DIAG>
__device__ __forceinline__ unsigned long long slab_tile(const double *tile, int q0, int cnt, int self, double x,
                                                        double cx) {
    unsigned long long m = 0;
    for (int q = q0; q < cnt; ++q) {
        const bool c = fabs(x - tile[4 * q]) <= cx;
        m |= __ballot(DIAG ? (c && q > self) : c);
    }
    return m;
}

// WPE: waves per SIMD the register allocation targets (512 / WPE registers
// per lane, AGPRs included): 1 lets a 256-thread group keep 4 bodies per
// thread in VGPRs + AGPRs; 3 for the one-body-per-thread shapes of 128 to
// 512 threads (measured +10-15 % evaluations per second on 2,048 identical
// 256-body units against 2; 4 spills too much and is slower).
template <int NT, int BPT, int FM, int CT, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void gr15_kernel(GR15Args A) {
    extern __shared__ double lds[];
    const long long t_enter = wall_clock64();
    const int sys = A.order[blockIdx.x];
    const SysDesc d = A.sys[sys];
    SysCtl *ctl = A.ctl + sys;
    if (ctl->status != ST_RUNNING) return;

    Engine<NT, BPT, FM, CT> E(A, sys, d, lds);
    // the plain controller keeps the reference's operation order (bit-exact
    // with exact forces); the massive-only controllers use the fused forms
    constexpr bool kFast = CT != ABIE_CTRL_PLAIN;
    const int tid = threadIdx.x;
    const long long T = A.total;
    const int b0 = d.body0, n = d.n, nm = d.nm;
    double *flag = E.red + 40;  // LDS scalars broadcast by thread 0
    // the proposer's weights to LDS (first read after the first force
    // evaluation's barrier)
    const float *mlp_w = nullptr;
    if (CT == ABIE_CTRL_LEARNED && A.mlp != nullptr) {
        float *wl = (float *)(E.red + 128);
        for (int i = tid; i < ABIE_MLP_NPARAM; i += NT) wl[i] = A.mlp[i];
        mlp_w = wl;
    }

    double y0[BPT][3], dy0[BPT][3], a0[BPT][3], bb[BPT][3][7], gg[BPT][3][7], b6p[BPT][3];
    double yp[BPT][3], acc[BPT][3];

    // ---- load state
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        const int b = k * NT + tid;
        const long long gi = b0 + (b < n ? b : 0);
        y0[k][0] = A.x[gi]; y0[k][1] = A.y[gi]; y0[k][2] = A.z[gi];
        dy0[k][0] = A.vx[gi]; dy0[k][1] = A.vy[gi]; dy0[k][2] = A.vz[gi];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                bb[k][c][j] = A.state[(long long)(c * 7 + j) * T + gi];
                gg[k][c][j] = A.state[(long long)(21 + c * 7 + j) * T + gi];
            }
            a0[k][c] = A.state[(long long)(42 + c) * T + gi];
            b6p[k][c] = A.state[(long long)(45 + c) * T + gi];
        }
    }
    // linked slices (plain controller only): maxima over the whole system
    const bool linked = CT == ABIE_CTRL_PLAIN && d.link_n > 1;
    const int nall = linked ? d.sys_n : n;  // the system's body count
    long long epoch = CT == ABIE_CTRL_PLAIN ? ctl->sync_epoch : 0;
    bool lk_ok = true;
    int red_par = 0;  // wg_max2's half of the reduction scratch
    auto rmax2 = [&](double &a, double &b) __attribute__((always_inline)) {
        if (CT == ABIE_CTRL_PLAIN) {
            if (!red_max2<NT>(A, d, epoch, a, b, E.red, red_par)) lk_ok = false;
        } else {
            wg_max2<NT>(a, b, E.red, red_par);
        }
    };
    // events: does any body of the system have a radius (collisions possible)?
    bool any_rad = false;
    if (A.events) {
        double rmax = 0.0, rdummy = 0.0;
#pragma unroll
        for (int k = 0; k < BPT; ++k) {
            const int b = k * NT + tid;
            if (b < n) rmax = fmax(rmax, A.rad[b0 + b]);
        }
        rmax2(rmax, rdummy);
        any_rad = rmax > 0.0;
    }
    double t = ctl->t, h = ctl->h;
    const double t_end = ctl->t_end;
    int rejects = ctl->rejects, warned = ctl->warned;
    long long n_evals = 0, n_att = 0, n_acc = 0, n_pc = 0, n_pu = 0, n_pr = 0;
    int status = ST_RUNNING;
    int prop_pending = ctl->pad;  // the current h came from the proposer

    // ---- fresh start: a0 = a(y0), initial_time_step (integrator_gauss_radau15.c:9-54,270-273)
    if (ctl->phase == 0) {
        E.forces(y0, a0);
        n_evals += 1;
        // d0, d1 over the first n of the 3n components (vector_max_abs(y0, nbodies))
        double d0 = -DBL_MAX, d1 = -DBL_MAX;
#pragma unroll
        for (int k = 0; k < BPT; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (CT == ABIE_CTRL_PLAIN ? (E.slot_valid[k] && 3 * E.slot_orig[k] + c < nall)
                                          : (E.norm_all ? (bool)E.slot_valid[k] : (bool)E.slot_massive[k])) {
                    d0 = fmax(d0, fabs(y0[k][c]));
                    d1 = fmax(d1, fabs(a0[k][c]));
                }
        rmax2(d0, d1);
        const double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
#pragma unroll
        for (int k = 0; k < BPT; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) yp[k][c] = y0[k][c] + dt0 * dy0[k][c];
        E.forces(yp, acc);
        n_evals += 1;
        double d2 = -DBL_MAX, dummy = -DBL_MAX;
#pragma unroll
        for (int k = 0; k < BPT; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (E.in_norm(k)) d2 = fmax(d2, fabs(acc[k][c] - a0[k][c]));
        rmax2(d2, dummy);
        d2 = d2 / dt0;
        double dt1;
        if (fmax(d1, d2) <= 1e-15) dt1 = fmax(1e-6, dt0 * 1e-3);
        else dt1 = pow_pos(0.01 / fmax(d1, d2), 1.0 / 16);
        h = fmin(100 * dt0, dt1);
#pragma unroll
        for (int k = 0; k < BPT; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                b6p[k][c] = 0.0;
#pragma unroll
                for (int j = 0; j < 7; ++j) bb[k][c][j] = gg[k][c][j] = 0.0;
            }
        rejects = 0;
        warned = 0;
    }

    // ---- the step loop (integrator_gauss_radau15.c:279-368)
    double dy[BPT][3];
    while (lk_ok && n_att < A.max_attempts) {
        // end the launch once its wall-clock budget is spent (the host then
        // relaunches; the state is stored below either way), so that no
        // unit waits on a fixed attempt count of the slowest one
        // (every 4 attempts; every 16 for linked slices, whose stop is a grid barrier)
        if (n_att > 0 && (n_att & (linked ? 15 : 3)) == 0 &&
            (A.budget_ticks > 0 || A.deadline_tick > 0 || A.tail_units > 0)) {
            if (tid == 0) {
                const long long now = wall_clock64();
                bool stop = (A.budget_ticks > 0 && now - t_enter > A.budget_ticks) ||
                            (A.deadline_tick > 0 && now > A.deadline_tick);
                if (!stop && A.tail_units > 0)
                    stop = __hip_atomic_load(A.running, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= A.tail_units;
                flag[2] = stop ? 1.0 : 0.0;
            }
            __syncthreads();
            double stop = flag[2], unused = 0.0;
            if (linked) rmax2(stop, unused);  // every slice stops at the same attempt
            if (stop != 0.0 || !lk_ok) break;
        }
        int pc_used = 0;
        double pc_last = DBL_MAX;
        for (int ipc = 0; ipc < 12; ++ipc) {
            pc_used = ipc + 1;
            double mdb = -DBL_MAX, mf = -DBL_MAX;
#define ABIE_NODE(IH)                                                                          \
    {                                                                                          \
        CKTab *K = launder();                                                             \
        const double s = K->h[IH];                                                             \
        const double Ts = h * s;                                                               \
        _Pragma("unroll") for (int k = 0; k < BPT; ++k) _Pragma("unroll") for (int c = 0; c < 3; ++c) \
            yp[k][c] = pred_pos<kFast>(K, y0[k][c], dy0[k][c], a0[k][c], bb[k][c], s, Ts);   \
        E.forces(yp, acc);                                                                     \
        CKTab *K2 = launder();                                                            \
        _Pragma("unroll") for (int k = 0; k < BPT; ++k) _Pragma("unroll") for (int c = 0; c < 3; ++c) \
            update_gb<IH, kFast>(K2, acc[k][c], a0[k][c], gg[k][c], bb[k][c]);                \
    }
            ABIE_NODE(1) ABIE_NODE(2) ABIE_NODE(3) ABIE_NODE(4) ABIE_NODE(5) ABIE_NODE(6) ABIE_NODE(7)
#undef ABIE_NODE
            n_evals += 7;
#pragma unroll
            for (int k = 0; k < BPT; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (E.in_norm(k)) {
                        mdb = fmax(mdb, fabs(bb[k][c][6] - b6p[k][c]));
                        mf = fmax(mf, fabs(acc[k][c]));
                    }
            rmax2(mdb, mf);
            if (!lk_ok) break;
            const double pc_err = mdb / mf;
            if (pc_err < kTolpc) break;
            // massive-only controllers: also stop once the correction no
            // longer shrinks (its round-off floor), as IAS15 does (Rein &
            // Spiegel 2015, MNRAS 446, 1424); the plain controller iterates
            // to 1e-16 or 12 times like the reference
            if (kFast && ipc >= 2 && pc_err >= pc_last) break;
            pc_last = pc_err;
#pragma unroll
            for (int k = 0; k < BPT; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c) b6p[k][c] = bb[k][c][6];
        }
        if (!lk_ok) break;
        n_pc += pc_used;
        n_att += 1;

        // advance: predictor at s = 1, final evaluation
        CKTab *KA = launder();
#pragma unroll
        for (int k = 0; k < BPT; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                yp[k][c] = pred_pos<kFast>(KA, y0[k][c], dy0[k][c], a0[k][c], bb[k][c], 1.0, h * 1.0);
                dy[k][c] = pred_vel1<kFast>(KA, dy0[k][c], a0[k][c], bb[k][c], h * 1.0);
            }
        E.forces(yp, acc);
        n_evals += 1;
        double mb = -DBL_MAX, ma = -DBL_MAX;
#pragma unroll
        for (int k = 0; k < BPT; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (E.in_norm(k)) {
                    mb = fmax(mb, fabs(bb[k][c][6]));
                    ma = fmax(ma, fabs(acc[k][c]));
                }
        rmax2(mb, ma);
        if (!lk_ok) break;
        // a system without any acceleration (one body alone, no ext-acc)
        // would give 0/0 as in the reference; it moves freely: estim 0
        const double estim = (ma > 0.0) ? mb / ma : 0.0;
        const double err = step_err<CT>(estim);
        const double dtreq = h / err;
        const double h_att = h, t_att = t;
        const int rej_before = rejects;
        bool accepted = false;
        bool finished = false;
        double h_next_plain;

        if (err <= 1 || rejects > 100) {
            accepted = true;
            t += h;
            rejects = 0;
            if (t > t_end) finished = true;
#pragma unroll
            for (int k = 0; k < BPT; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    y0[k][c] = yp[k][c];
                    dy0[k][c] = dy[k][c];
                    a0[k][c] = acc[k][c];
                    b6p[k][c] = bb[k][c][6];
                }
            n_acc += 1;
        } else {
            rejects += 1;
        }
        // Shortest encounter timescale among the massive bodies of the new
        // state, as max |v_ij|^2 / r_ij^2 (a feature of the learned proposer,
        // and logged): massive bodies put (x, v) in the LDS source area,
        // free between evaluations, and each sums its row of pairs.
        double rate2 = 0.0;
        if (accepted && ((CT == ABIE_CTRL_LEARNED && A.mlp != nullptr) || A.log != nullptr)) {
            double *L = E.src;
#pragma unroll
            for (int k = 0; k < BPT; ++k) {
                const int b = k * NT + tid;
                if (b < nm) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        L[6 * b + c] = y0[k][c];
                        L[6 * b + 3 + c] = dy0[k][c];
                    }
                }
            }
            __syncthreads();
            double best = 0.0, unused = 0.0;
#pragma unroll
            for (int k = 0; k < BPT; ++k) {
                const int b = k * NT + tid;
                if (b >= nm) continue;
#pragma unroll 4
                for (int j = 0; j < nm; ++j) {
                    if (j == b) continue;
                    const double dx = y0[k][0] - L[6 * j], dyy = y0[k][1] - L[6 * j + 1], dz = y0[k][2] - L[6 * j + 2];
                    const double vx = dy0[k][0] - L[6 * j + 3], vy = dy0[k][1] - L[6 * j + 4],
                                 vz = dy0[k][2] - L[6 * j + 5];
                    best = fmax(best, pair_rate2(dx, dyy, dz, vx, vy, vz));
                }
            }
            wg_max2<NT>(best, unused, E.red, red_par);
            rate2 = best;
        }

        // step size (integrator_gauss_radau15.c:349-367)
        if (dtreq / h > 1.0 / kFac) h_next_plain = h / kFac;
        else if (dtreq < 1.e-12) h_next_plain = h * kFac;
        else h_next_plain = dtreq;
        double h_next = h_next_plain;
        double q = dtreq / h_att;  // refine_bs ratio of the reference
        if (q > DBL_MAX) q = 1.0 / kFac;  // only when estim = 0 (free motion, all b = 0)

        // the learned proposer (after accepted steps only)
        if (CT == ABIE_CTRL_LEARNED && accepted && A.mlp != nullptr) {
            if (tid < 64) {
                const PropIn f{h_att, err, dtreq, ctl->prev_h, ctl->prev_err, ctl->prev_dtreq, rate2, A.mratio[sys],
                               pc_used, ctl->prev_pc, rej_before, nm, d.nt_sys};
                float xi;
                const float yl = proposer_eval(mlp_w, f, (float *)(E.red + 64), &xi);
                if (tid == 0) flag[0] = (double)yl;
                if (A.log != nullptr) {  // the network's inputs and raw output, for the step log
                    if (tid < ABIE_MLP_IN) E.red[kLdsFeat + tid] = (double)xi;
                    if (tid == 0) E.red[kLdsFeat + ABIE_MLP_IN] = (double)yl;
                }
            }
            __syncthreads();
            double ratio = exp10(fmin(fmax(flag[0], kPropLo), kPropHi));
            double hp = dtreq * ratio;
            // keep the reference's growth / shrink bounds
            hp = fmin(hp, h_att / kFac);
            hp = fmax(hp, h_att * kFac);
            // the refine_bs ratio follows the proposal: the reference's q
            // scaled by the factor the proposer applied to its step
            q = q * (hp / h_next_plain);
            h_next = hp;
            n_pu += 1;
        }
        if (prop_pending && !accepted) n_pr += 1;
        prop_pending = (CT == ABIE_CTRL_LEARNED && accepted && A.mlp != nullptr);

        // step log
        if (A.log != nullptr && tid == 0 && d.primary) {
            const unsigned long long slot = atomicAdd(A.log_len, 1ull);
            if ((long long)slot < A.log_cap) {
                double *r = A.log + slot * ABIE_STEPLOG_FIELDS;
                r[0] = d.sys_id; r[1] = t_att; r[2] = h_att; r[3] = estim; r[4] = err; r[5] = dtreq;
                r[6] = accepted; r[7] = pc_used; r[8] = rej_before; r[9] = ctl->prev_h; r[10] = ctl->prev_err;
                r[11] = ctl->prev_dtreq; r[12] = ctl->prev_pc; r[13] = nm; r[14] = d.nt_sys; r[15] = h_next;
                r[16] = rate2; r[17] = A.mratio[sys];
                // proposer inputs and raw output (NaN when no proposal was made)
                const bool proposed = CT == ABIE_CTRL_LEARNED && accepted && A.mlp != nullptr;
#pragma unroll
                for (int i = 0; i <= ABIE_MLP_IN; ++i) r[18 + i] = proposed ? E.red[kLdsFeat + i] : __builtin_nan("");
            }
        }
        if (accepted) {
            if (tid == 0) {
                ctl->prev_h = h_att;
                ctl->prev_err = err;
                ctl->prev_dtreq = dtreq;
                ctl->prev_pc = pc_used;
            }
            {
                // also on the last step: the state stays ready to continue
                CKTab *KR = launder();
#pragma unroll
                for (int k = 0; k < BPT; ++k)
#pragma unroll
                    for (int c = 0; c < 3; ++c) refine(KR, bb[k][c], q);
            }
            // close encounters / collisions (common.c:116-158) on the new positions
            if (A.events) {
                double *sx = A.scan, *sy = A.scan + T, *sz = A.scan + 2 * T, *sf = A.scan + 3 * T;
#pragma unroll
                for (int k = 0; k < BPT; ++k) {
                    const int b = k * NT + tid;
                    if (b < n) {
                        sx[b0 + b] = y0[k][0]; sy[b0 + b] = y0[k][1]; sz[b0 + b] = y0[k][2];
                        sf[b0 + b] = 0.0;
                    }
                }
                if (linked) {  // every slice's positions are written before any is read
                    double u0 = 0.0, u1 = 0.0;
                    rmax2(u0, u1);
                    if (!lk_ok) break;
                }
                // hit detection: a superset test over each unordered pair
                // once, bodies staged in LDS tiles of NT (x, y, z, radius);
                // a hit runs the exact pass below.  The test is symmetric in
                // the pair, so checking i > j only finds the same hits.  A
                // linked slice other than the primary leaves its copies of
                // the massive bodies out (NaN coordinates never pass), and
                // also tests its bodies against the slices link_rank + 1 ..
                // link_rank + link_n / 2 (mod link_n): every pair of slices
                // is covered once (twice for the opposite slice when link_n
                // is even, which only repeats the same hits).
                const double cd = A.ce_dist;
                const double cd2 = cd * cd * (1.0 + 0x1p-46);
                const double cx = sqrt(cd2) * (1.0 + 0x1p-48) + 0x1p-500;
                const bool skipm = linked && !d.primary;
                double *tile = E.src;  // the force source list is not live here
                const int lane0 = tid & ~63;
                bool any = false;
                double hit = 0.0, dummy2 = 0.0;
                // the tile [cnt) against the bodies of slot k (q from q0; on
                // the diagonal tile only entries above the lane's own body)
                // qbase: the body slot of tile entry 0
                auto test_tile = [&](int k, int q0, int cnt, bool diag, long long qbase)
                                     __attribute__((always_inline)) {
                    const int j = k * NT + tid;
                    const bool valid = j < n && !(skipm && j < nm);
                    const double rj = valid ? A.rad[b0 + j] : 0.0;
                    const double thr0 = valid ? cd2 : -1.0;  // -1: no pair of an idle lane passes
                    const int self = diag ? tid : -1;
                    bool hit = false;
                    if (any_rad) {
                        hit = scan_tile<true>(tile, q0, cnt, self, y0[k], rj, thr0, cd2);
                        if (__ballot(hit)) mark_tile<true>(tile, q0, cnt, self, y0[k], rj, thr0, cd2, sf, b0 + j, qbase);
                    } else {
                        // the full test only for a wave whose x-slab test
                        // passes somewhere in the tile (wave-uniform branch)
                        const double cxl = valid ? cx : -1.0;
                        const unsigned long long m = diag ? slab_tile<true>(tile, q0, cnt, tid, y0[k][0], cxl)
                                                          : slab_tile<false>(tile, q0, cnt, -1, y0[k][0], cxl);
                        if (m) hit = scan_tile<false>(tile, q0, cnt, self, y0[k], rj, thr0, cd2);
                        if (__ballot(hit)) mark_tile<false>(tile, q0, cnt, self, y0[k], rj, thr0, cd2, sf, b0 + j, qbase);
                    }
                    any |= hit;
                };
                const double qnan = __builtin_nan("");
                for (int kt = 0; kt < BPT && kt * NT < n; ++kt) {
                    const int base = kt * NT;
                    const int cnt = min(NT, n - base);
                    __syncthreads();
                    if (tid < cnt) {
                        const bool out = skipm && base + tid < nm;
                        tile[4 * tid] = out ? qnan : y0[kt][0];
                        tile[4 * tid + 1] = out ? qnan : y0[kt][1];
                        tile[4 * tid + 2] = out ? qnan : y0[kt][2];
                        tile[4 * tid + 3] = A.rad[b0 + base + tid];
                    }
                    __syncthreads();
#pragma unroll
                    for (int k = 0; k < BPT; ++k) {
                        if (k > kt || k * NT + lane0 >= n) continue;  // tile wholly below this wave's bodies
                        // on the diagonal tile only i = base + q > j: the wave's
                        // first lane bounds the loop, the rest is a lane mask
                        const bool diag = k == kt;
                        test_tile(k, diag ? lane0 + 1 : 0, cnt, diag, b0 + base);
                    }
                }
                if (linked) {
                    for (int off = 1; off <= d.link_n / 2; ++off) {
                        const int v = (d.link_rank + off) % d.link_n;
                        const SysDesc dv = A.sys[d.link_base + v];
                        const long long s0 = dv.body0 + (v == 0 ? 0 : nm);  // the partner without massive copies
                        const int sc = dv.n - (v == 0 ? 0 : nm);
                        for (int base = 0; base < sc; base += NT) {
                            const int cnt = min(NT, sc - base);
                            __syncthreads();
                            if (tid < cnt) {
                                const long long gq = s0 + base + tid;
                                tile[4 * tid] = sx[gq];
                                tile[4 * tid + 1] = sy[gq];
                                tile[4 * tid + 2] = sz[gq];
                                tile[4 * tid + 3] = A.rad[gq];
                            }
                            __syncthreads();
#pragma unroll
                            for (int k = 0; k < BPT; ++k) {
                                if (k * NT + lane0 >= n) continue;
                                test_tile(k, 0, cnt, false, s0 + base);
                            }
                        }
                    }
                }
                hit = any ? 1.0 : 0.0;
                rmax2(hit, dummy2);
                if (!lk_ok) break;
                if (hit > 0.0) {
                    // exact pass in the reference's (j, k) order by the first
                    // wave (of the primary slice): j in original order, the
                    // k > j in chunks of 64 lanes; the ring slots follow from
                    // ballots, so the rings and counters end as the serial
                    // loop leaves them.  slot(q): the body slot of the q-th
                    // body of the system by original index.
                    if (tid < 64 && (!linked || d.primary)) {
                        auto slot = [&](int q) __attribute__((always_inline)) -> long long {
                            return linked ? (long long)A.sys_inv[d.sys_inv0 + q] : b0 + A.inv[b0 + q];
                        };
                        const int lane = tid;
                        const unsigned long long below = (1ull << lane) - 1ull;
                        long long nce = ctl->n_ce, nco = ctl->n_coll;
                        double *rce = (A.ring_ce && A.max_ce > 0) ? A.ring_ce + (long long)sys * 4 * A.max_ce : nullptr;
                        double *rco = (A.ring_coll && A.max_coll > 0) ? A.ring_coll + (long long)sys * 4 * A.max_coll
                                                                       : nullptr;
                        for (int q1 = 0; q1 < nall; ++q1) {
                            const long long sj = slot(q1);
                            if (sf[sj] == 0.0) continue;  // no candidate pair (the same for every lane)
                            const double jo = A.orig[sj];
                            const double xj = sx[sj], yj = sy[sj], zj = sz[sj];
                            const double rj = A.rad[sj];
                            for (int c = q1 + 1; c < nall; c += 64) {
                                const int q2 = c + lane;
                                bool ce = false, co = false;
                                double dd = 0.0, ko = 0.0;
                                const long long sk = q2 < nall ? slot(q2) : 0;
                                if (q2 < nall && sf[sk] != 0.0) {
                                    ko = A.orig[sk];
                                    const double dx = xj - sx[sk], dyy = yj - sy[sk], dz = zj - sz[sk];
                                    const double r2 = dx * dx + dyy * dyy + dz * dz;
                                    dd = sqrt(r2);
                                    const double rr = rj + A.rad[sk];
                                    ce = dd <= cd;
                                    co = rr > 0 && dd <= rr;
                                }
                                const unsigned long long mce = __ballot(ce), mco = __ballot(co);
                                if (mce | mco) {
                                    const long long ece = nce + __popcll(mce), eco = nco + __popcll(mco);
                                    // an event is written unless a later one of
                                    // this chunk takes the same ring slot
                                    if (ce && rce) {
                                        const long long idx = nce + __popcll(mce & below);
                                        if (idx >= ece - A.max_ce) {
                                            const long long w = 4LL * A.max_ce, o = 4 * idx;
                                            rce[o % w] = t; rce[(o + 1) % w] = jo; rce[(o + 2) % w] = ko;
                                            rce[(o + 3) % w] = dd;
                                        }
                                    }
                                    if (co && rco) {
                                        const long long idx = nco + __popcll(mco & below);
                                        if (idx >= eco - A.max_coll) {
                                            const long long w = 4LL * A.max_coll, o = 4 * idx;
                                            rco[o % w] = t; rco[(o + 1) % w] = jo; rco[(o + 2) % w] = ko;
                                            rco[(o + 3) % w] = dd;
                                        }
                                    }
                                    nce = ece;
                                    nco = eco;
                                    // the next chunk's writes may reuse these slots
                                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                                }
                            }
                        }
                        if (lane == 0) {
                            ctl->n_ce = nce;
                            ctl->n_coll = nco;
                            int st = ST_RUNNING;
                            if (A.max_ce > 0 && nce >= A.max_ce) st = ST_CE;
                            else if (A.max_coll > 0 && nco >= A.max_coll) st = ST_COLL;
                            flag[1] = (double)st;
                        }
                    }
                    __syncthreads();
                    double stv = (!linked || d.primary) ? flag[1] : -DBL_MAX, unused = 0.0;
                    if (linked) {  // the primary's outcome reaches every slice
                        rmax2(stv, unused);
                        if (!lk_ok) break;
                    }
                    const int st = (int)stv;
                    if (st != ST_RUNNING) {
                        status = st;
                        h = h_next;
                        break;
                    }
                }
            }
        }
        h = h_next;
        if (h < kHmin) {
            if (!warned && tid == 0 && d.primary)
                printf("Warning! Timestep %g being too small! Imposing a minimum timestep of %g, t = %g\n", h,
                       kHmin, t);
            warned = 1;
            h = kHmin;
        }
        if (finished) {
            status = ST_DONE;
            break;
        }
    }

    // ---- store state (y0 / dy as the reference writes pos / vel).  The
    // pointers are laundered so that the load addresses are not kept live
    // across the whole kernel for reuse here.
#pragma unroll
    for (int k = 0; k < BPT; ++k) {
        int b = k * NT + tid;
        asm volatile("" : "+v"(b));
        if (b >= n) continue;
        const long long gi = b0 + b;
        double *px = A.x, *py = A.y, *pz = A.z, *pvx = A.vx, *pvy = A.vy, *pvz = A.vz, *ps = A.state;
        asm volatile("" : "+s"(px), "+s"(py), "+s"(pz), "+s"(pvx), "+s"(pvy), "+s"(pvz), "+s"(ps));
        px[gi] = y0[k][0]; py[gi] = y0[k][1]; pz[gi] = y0[k][2];
        pvx[gi] = dy0[k][0]; pvy[gi] = dy0[k][1]; pvz[gi] = dy0[k][2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                ps[(long long)(c * 7 + j) * T + gi] = bb[k][c][j];
                ps[(long long)(21 + c * 7 + j) * T + gi] = gg[k][c][j];
            }
            ps[(long long)(42 + c) * T + gi] = a0[k][c];
            ps[(long long)(45 + c) * T + gi] = b6p[k][c];
        }
    }
    if (!lk_ok) status = ST_ERROR;  // a grid barrier timed out
    if (tid == 0) {
        if (CT == ABIE_CTRL_PLAIN) ctl->sync_epoch = epoch;
        ctl->t = t;
        ctl->h = h;
        ctl->rejects = rejects;
        ctl->warned = warned;
        ctl->phase = 1;
        ctl->status = status;
        ctl->pad = prop_pending;
        ctl->n_evals += n_evals;
        ctl->n_attempts += n_att;
        ctl->n_accepted += n_acc;
        ctl->n_pc += n_pc;
        ctl->n_prop_used += n_pu;
        ctl->n_prop_rej += n_pr;
        if (ctl->start_tick == 0) ctl->start_tick = t_enter;
        ctl->busy_ticks += wall_clock64() - t_enter;
        if (status != ST_RUNNING) {
            ctl->done_tick = wall_clock64();
            if (A.running != nullptr) atomicSub(A.running, 1);
        }
    }
}


template <int NT, int BPT, int FM, int CT, int WPE>
static hipError_t launch_one(const GR15Args &a, int nb, size_t lds, hipStream_t st, bool coop) {
    if (coop) {
        GR15Args arg = a;
        void *params[] = {&arg};
        return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&gr15_kernel<NT, BPT, FM, CT, WPE>),
                                          dim3(nb), dim3(NT), params, (unsigned)lds, st);
    }
    hipLaunchKernelGGL((gr15_kernel<NT, BPT, FM, CT, WPE>), dim3(nb), dim3(NT), lds, st, a);
    return hipGetLastError();
}

// Instantiates every configuration of kCfgs for one (force mode, controller).
template <int FM, int CT>
hipError_t launch_gr15_fm_ct(int cfg, const GR15Args &a, int nb, size_t lds, hipStream_t st, bool coop) {
    switch (cfg) {
        case 0: return launch_one<64, 1, FM, CT, 2>(a, nb, lds, st, coop);
        case 1: return launch_one<128, 1, FM, CT, 3>(a, nb, lds, st, coop);
        case 2: return launch_one<256, 1, FM, CT, 3>(a, nb, lds, st, coop);
        case 3: return launch_one<256, 2, FM, CT, 1>(a, nb, lds, st, coop);
        case 4: return launch_one<512, 1, FM, CT, 3>(a, nb, lds, st, coop);
        case 5: return launch_one<256, 4, FM, CT, 1>(a, nb, lds, st, coop);
        case 6: return launch_one<512, 2, FM, CT, 2>(a, nb, lds, st, coop);
        case 7: return launch_one<1024, 1, FM, CT, 4>(a, nb, lds, st, coop);
        case 8: return launch_one<512, 3, FM, CT, 2>(a, nb, lds, st, coop);
        case 9: return launch_one<512, 4, FM, CT, 2>(a, nb, lds, st, coop);
        case 10: return launch_one<512, 5, FM, CT, 2>(a, nb, lds, st, coop);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace abie
