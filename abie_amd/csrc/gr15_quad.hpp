// gr15_quad.hpp -- the GR15 engine form for the massive-only controllers
// (ABIE_CTRL_MASSIVE, ABIE_CTRL_LEARNED): four lanes per body.
//
// Same step as gr15_kernel (integrator_gauss_radau15.c:279-368 with the
// massive-only norms, the fused predictor and the IAS15-style PC stop), laid
// out for latency: the four lanes of a quad (lanes 4b .. 4b+3 of the group
// hold body b) split the work of one body.
//  * Lane c < 3 owns component c of the body's integrator state -- y0, dy0,
//    a0, b[7], g[7] and the previous b[6], 18 doubles instead of 54 per
//    thread -- and runs the predictor and the g/b update of that component.
//    Lane 3 owns no component.
//  * The force sum is split over the quad: lane l sums the sources
//    j = l, l + 4, l + 8, ... in ascending order, and the quad combines the
//    partial sums as (P0 + P1) + (P2 + P3) by two DPP butterfly steps, which
//    leaves identical bits in the four lanes (floating-point addition is
//    commutative).  The summation tree depends only on the source count, so
//    a body's trajectory does not depend on which group, slice or launch
//    integrates it.
// A group of NT threads holds NT / 4 bodies (a system of up to 256 bodies,
// or a test-particle slice of one).  Each body's share of the source loop is
// a quarter of the sources, and the 54-double state spread over three lanes
// leaves the register file room for four waves per SIMD without spills.
#pragma once

#include "gr15_impl.hpp"

namespace abie {

// v from lane CTRL's choice within each quad (a DPP quad_perm): both 32-bit
// halves through v_mov_dpp.  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double quad_dpp(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = (int)(unsigned)(u & 0xffffffffll), hi = (int)(u >> 32);
    const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
    const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)rhi << 32) | (long long)(unsigned)rlo);
}
// quad_perm encodings: lane k of a quad reads lane sel_k (2 bits each)
constexpr int kQuadSwap1 = 0xB1;   // [1, 0, 3, 2]
constexpr int kQuadSwap2 = 0x4E;   // [2, 3, 0, 1]
constexpr int kQuadBcast0 = 0x00;  // [0, 0, 0, 0]
constexpr int kQuadBcast1 = 0x55;  // [1, 1, 1, 1]
constexpr int kQuadBcast2 = 0xAA;  // [2, 2, 2, 2]

// (P0 + P1) + (P2 + P3) over the quad, the same bits in every lane
__device__ __forceinline__ double quad_sum(double v) {
    v = v + quad_dpp<kQuadSwap1>(v);
    return v + quad_dpp<kQuadSwap2>(v);
}

// Registers: a 1,024-thread group needs four waves per SIMD (at most 128
// VGPRs) to launch at all; smaller groups serve the tail, one group per CU,
// where two waves per SIMD (256 VGPRs) leave the state, the force loop's
// pairs in flight and the step control without spills.
constexpr int quad_wpe(int nt) { return nt >= 1024 ? 4 : 2; }
template <int NT, int FM, int CT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(quad_wpe(NT), quad_wpe(NT)))) void gr15q_kernel(
    GR15Args A) {
    static_assert(CT != ABIE_CTRL_PLAIN, "the quad form serves the massive-only controllers");
    constexpr int NB = NT / 4;  // bodies per group
    extern __shared__ double lds[];
    double *const src = lds;                // [2][NB][4] source lists (x, y, z, G m)
    double *const red = lds + 2 * NB * 4;   // scratch, as gr15_kernel's (kLdsFeat)
    double *const flag = red + 40;
    int red_par = 0;  // wg_max2's half of the reduction scratch
    const long long t_enter = wall_clock64();
    const int sys = A.order[blockIdx.x];
    const SysDesc d = A.sys[sys];
    SysCtl *ctl = A.ctl + sys;
    if (ctl->status != ST_RUNNING) return;
    // Waves that hold none of the unit's bodies end here: the group's
    // barriers (s_barrier) then wait for the surviving waves only, and a
    // unit smaller than its group runs at the latency of its own waves.
    const int nw = max(1, min(NT / 64, (d.n + 15) / 16));  // waves with bodies
    if ((int)(threadIdx.x >> 6) >= nw) return;
    // the proposer's weights to LDS (first read after the first force
    // evaluation's barrier)
    const float *mlp_w = nullptr;
    if (CT == ABIE_CTRL_LEARNED && A.mlp != nullptr) {
        float *wl = (float *)(red + 128);
        for (int i = threadIdx.x; i < ABIE_MLP_NPARAM; i += 64 * nw) wl[i] = A.mlp[i];
        mlp_w = wl;
    }

    const int tid = threadIdx.x;
    const int bi = tid >> 2;  // body of this quad
    const int c = tid & 3;    // component of this lane (3: none)
    const long long T = A.total;
    const int b0 = d.body0, n = d.n, nm = d.nm;
    const bool valid = bi < n;
    const bool comp = c < 3 && valid;
    const bool massive = bi < nm;
    const bool deleted = valid && A.deleted[b0 + bi] != 0;
    const bool in_norm = c < 3 && (d.norm_all != 0 ? valid : massive);
    const bool wave_massive = (tid & ~63) < 4 * nm;
    const double gm_b = massive ? A.gm[b0 + bi] : 0.0;
    const long long gi = b0 + (valid ? bi : 0);
    Coef2 kc{1.5, 1.875};
    asm volatile("" : "+v"(kc.c1), "+v"(kc.c2));

    // ---- load this lane's component of the state
    double y0 = 0.0, dy0 = 0.0, a0 = 0.0, b6p = 0.0, bb[7], gg[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) bb[j] = gg[j] = 0.0;
    if (comp) {
        const double *px = c == 0 ? A.x : (c == 1 ? A.y : A.z);
        const double *pv = c == 0 ? A.vx : (c == 1 ? A.vy : A.vz);
        y0 = px[gi];
        dy0 = pv[gi];
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            bb[j] = A.state[(long long)(c * 7 + j) * T + gi];
            gg[j] = A.state[(long long)(21 + c * 7 + j) * T + gi];
        }
        a0 = A.state[(long long)(42 + c) * T + gi];
        b6p = A.state[(long long)(45 + c) * T + gi];
    }
    const double ext_c = (A.ext != nullptr && comp) ? A.ext[c * T + gi] : 0.0;

    int buf = 0;
    // acceleration component c at the predicted component yp of every lane
    auto forces = [&](double yp) __attribute__((always_inline)) -> double {
        const double px = quad_dpp<kQuadBcast0>(yp), py = quad_dpp<kQuadBcast1>(yp), pz = quad_dpp<kQuadBcast2>(yp);
        double *S = src + buf * NB * 4;
        buf ^= 1;
        if (massive && c == 0) {
            S[4 * bi] = px;
            S[4 * bi + 1] = py;
            S[4 * bi + 2] = pz;
            S[4 * bi + 3] = gm_b;
        }
        __syncthreads();
        double ax = 0.0, ay = 0.0, az = 0.0;
        // massive waves: FP64 with the self-pair removed for massive sinks,
        // test particles of the same wave as in the other branch (one loop,
        // no divergence)
        if (wave_massive) sum_stride4<FM, true, (NT >= 1024 ? 1 : 2)>(S, c, nm, bi, massive, px, py, pz, ax, ay, az, kc);
        else sum_stride4<FM, false, (NT >= 1024 ? 1 : 2)>(S, c, nm, bi, false, px, py, pz, ax, ay, az, kc);
        ax = quad_sum(ax);
        ay = quad_sum(ay);
        az = quad_sum(az);
        double a = c == 0 ? ax : (c == 1 ? ay : az);
        if (deleted) a = 0.0;
        // lane 3 and idle quads keep a zero state
        return comp ? (A.ext != nullptr ? a + ext_c : a) : 0.0;
    };

    double t = ctl->t, h = ctl->h;
    const double t_end = ctl->t_end;
    int rejects = ctl->rejects, warned = ctl->warned;
    long long n_evals = 0, n_att = 0, n_acc = 0, n_pc = 0, n_pu = 0, n_pr = 0;
    int status = ST_RUNNING;
    int prop_pending = ctl->pad;

    // ---- fresh start: a0 = a(y0), initial_time_step (integrator_gauss_radau15.c:9-54,270-273)
    if (ctl->phase == 0) {
        a0 = forces(y0);
        n_evals += 1;
        double d0 = in_norm ? fabs(y0) : -DBL_MAX, d1 = in_norm ? fabs(a0) : -DBL_MAX;
        wg_max2<NT>(d0, d1, red, red_par, nw);
        const double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
        const double acc = forces(y0 + dt0 * dy0);
        n_evals += 1;
        double d2 = in_norm ? fabs(acc - a0) : -DBL_MAX, dummy = -DBL_MAX;
        wg_max2<NT>(d2, dummy, red, red_par, nw);
        d2 = d2 / dt0;
        double dt1;
        if (fmax(d1, d2) <= 1e-15) dt1 = fmax(1e-6, dt0 * 1e-3);
        else dt1 = pow_pos(0.01 / fmax(d1, d2), 1.0 / 16);
        h = fmin(100 * dt0, dt1);
        b6p = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j) bb[j] = gg[j] = 0.0;
        rejects = 0;
        warned = 0;
    }

    // ---- the step loop (integrator_gauss_radau15.c:279-368)
    // attempts this launch may make before the unit's attempt count reaches
    // ctl->stop_attempt (when set)
    long long stop_at = 0x7fffffffffffffffll;
    if (ctl->stop_attempt > 0) stop_at = ctl->stop_attempt > ctl->n_attempts ? ctl->stop_attempt - ctl->n_attempts : 0;
    while (n_att < A.max_attempts && n_att < stop_at) {
        if (n_att > 0 && (n_att & 3) == 0 && (A.budget_ticks > 0 || A.deadline_tick > 0 || A.tail_units > 0)) {
            if (tid == 0) {
                const long long now = wall_clock64();
                bool stop = (A.budget_ticks > 0 && now - t_enter > A.budget_ticks) ||
                            (A.deadline_tick > 0 && now > A.deadline_tick);
                if (!stop && A.tail_units > 0)
                    stop = __hip_atomic_load(A.running, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= A.tail_units;
                flag[2] = stop ? 1.0 : 0.0;
            }
            __syncthreads();
            if (flag[2] != 0.0) break;
        }
        int pc_used = 0;
        double pc_last = DBL_MAX;
        double acc = 0.0;
        for (int ipc = 0; ipc < 12; ++ipc) {
            pc_used = ipc + 1;
#define ABIE_QNODE(IH)                                                              \
    {                                                                               \
        CKTab *K = launder();                                                  \
        const double s = K->h[IH];                                                  \
        acc = forces(pred_pos<true>(K, y0, dy0, a0, bb, s, h * s));                 \
        update_gb<IH, true>(launder(), acc, a0, gg, bb);                            \
    }
            ABIE_QNODE(1) ABIE_QNODE(2) ABIE_QNODE(3) ABIE_QNODE(4) ABIE_QNODE(5) ABIE_QNODE(6) ABIE_QNODE(7)
#undef ABIE_QNODE
            n_evals += 7;
            double mdb = in_norm ? fabs(bb[6] - b6p) : -DBL_MAX, mf = in_norm ? fabs(acc) : -DBL_MAX;
            wg_max2<NT>(mdb, mf, red, red_par, nw);
            const double pc_err = mdb / mf;
            if (pc_err < kTolpc) break;
            // stop once the correction no longer shrinks (Rein & Spiegel 2015)
            if (ipc >= 2 && pc_err >= pc_last) break;
            pc_last = pc_err;
            b6p = bb[6];
        }
        n_pc += pc_used;
        n_att += 1;

        // advance: predictor at s = 1, final evaluation
        CKTab *KA = launder();
        const double yp = pred_pos<true>(KA, y0, dy0, a0, bb, 1.0, h * 1.0);
        const double dy = pred_vel1<true>(KA, dy0, a0, bb, h * 1.0);
        acc = forces(yp);
        n_evals += 1;
        double mb = in_norm ? fabs(bb[6]) : -DBL_MAX, ma = in_norm ? fabs(acc) : -DBL_MAX;
        wg_max2<NT>(mb, ma, red, red_par, nw);
        const double estim = (ma > 0.0) ? mb / ma : 0.0;
        const double err = step_err<CT>(estim);
        const double dtreq = h / err;
        const double h_att = h, t_att = t;
        const int rej_before = rejects;
        bool accepted = false, finished = false;

        if (err <= 1 || rejects > 100) {
            accepted = true;
            t += h;
            rejects = 0;
            if (t > t_end) finished = true;
            y0 = yp;
            dy0 = dy;
            a0 = acc;
            b6p = bb[6];
            n_acc += 1;
        } else {
            rejects += 1;
        }
        // max |v_ij|^2 / r_ij^2 over the massive pairs of the new state (the
        // proposer's encounter timescale, and logged): each massive body's
        // (x, v) goes to the LDS source area, free here, and each quad takes
        // the max over its row, the four lanes over a quarter of it each
        double rate2 = 0.0;
        if (accepted && ((CT == ABIE_CTRL_LEARNED && A.mlp != nullptr) || A.log != nullptr)) {
            const double X = quad_dpp<kQuadBcast0>(y0), Y = quad_dpp<kQuadBcast1>(y0), Z = quad_dpp<kQuadBcast2>(y0);
            const double VX = quad_dpp<kQuadBcast0>(dy0), VY = quad_dpp<kQuadBcast1>(dy0),
                         VZ = quad_dpp<kQuadBcast2>(dy0);
            double *L = src;
            if (massive && c == 0) {
                L[6 * bi] = X; L[6 * bi + 1] = Y; L[6 * bi + 2] = Z;
                L[6 * bi + 3] = VX; L[6 * bi + 4] = VY; L[6 * bi + 5] = VZ;
            }
            __syncthreads();
            double best = 0.0, unused = 0.0;
            if (massive) {
#pragma unroll 4
                for (int j = c; j < nm; j += 4) {
                    if (j == bi) continue;
                    const double dx = X - L[6 * j], dyy = Y - L[6 * j + 1], dz = Z - L[6 * j + 2];
                    const double vx = VX - L[6 * j + 3], vy = VY - L[6 * j + 4], vz = VZ - L[6 * j + 5];
                    best = fmax(best, pair_rate2(dx, dyy, dz, vx, vy, vz));
                }
            }
            wg_max2<NT>(best, unused, red, red_par, nw);
            rate2 = best;
        }

        // step size (integrator_gauss_radau15.c:349-367)
        double h_next_plain;
        if (dtreq / h > 1.0 / kFac) h_next_plain = h / kFac;
        else if (dtreq < 1.e-12) h_next_plain = h * kFac;
        else h_next_plain = dtreq;
        double h_next = h_next_plain;
        double q = dtreq / h_att;
        if (q > DBL_MAX) q = 1.0 / kFac;  // only when estim = 0 (free motion, all b = 0)

        if (CT == ABIE_CTRL_LEARNED && accepted && A.mlp != nullptr) {
            if (tid < 64) {
                const PropIn f{h_att, err, dtreq, ctl->prev_h, ctl->prev_err, ctl->prev_dtreq, rate2, A.mratio[sys],
                               pc_used, ctl->prev_pc, rej_before, nm, d.nt_sys};
                float xi;
                const float yl = proposer_eval(mlp_w, f, (float *)(red + 64), &xi);
                if (tid == 0) flag[0] = (double)yl;
                if (A.log != nullptr) {
                    if (tid < ABIE_MLP_IN) red[kLdsFeat + tid] = (double)xi;
                    if (tid == 0) red[kLdsFeat + ABIE_MLP_IN] = (double)yl;
                }
            }
            __syncthreads();
            const double ratio = exp10(fmin(fmax(flag[0], kPropLo), kPropHi));
            double hp = dtreq * ratio;
            hp = fmin(hp, h_att / kFac);
            hp = fmax(hp, h_att * kFac);
            q = q * (hp / h_next_plain);
            h_next = hp;
            n_pu += 1;
        }
        if (prop_pending && !accepted) n_pr += 1;
        prop_pending = (CT == ABIE_CTRL_LEARNED && accepted && A.mlp != nullptr);

        if (A.log != nullptr && tid == 0 && d.primary) {
            const unsigned long long slot = atomicAdd(A.log_len, 1ull);
            if ((long long)slot < A.log_cap) {
                double *r = A.log + slot * ABIE_STEPLOG_FIELDS;
                r[0] = d.sys_id; r[1] = t_att; r[2] = h_att; r[3] = estim; r[4] = err; r[5] = dtreq;
                r[6] = accepted; r[7] = pc_used; r[8] = rej_before; r[9] = ctl->prev_h; r[10] = ctl->prev_err;
                r[11] = ctl->prev_dtreq; r[12] = ctl->prev_pc; r[13] = nm; r[14] = d.nt_sys; r[15] = h_next;
                r[16] = rate2; r[17] = A.mratio[sys];
                const bool proposed = CT == ABIE_CTRL_LEARNED && accepted && A.mlp != nullptr;
#pragma unroll
                for (int i = 0; i <= ABIE_MLP_IN; ++i) r[18 + i] = proposed ? red[kLdsFeat + i] : __builtin_nan("");
            }
        }
        if (accepted) {
            if (tid == 0) {
                ctl->prev_h = h_att;
                ctl->prev_err = err;
                ctl->prev_dtreq = dtreq;
                ctl->prev_pc = pc_used;
            }
            refine(launder(), bb, q);
        }
        // a non-finite step or time never reaches t_end: stop the unit with
        // an error instead of stepping on (the value is workgroup-uniform)
        if (!(h_next > 0.0) || !(fabs(t) <= DBL_MAX)) {
            status = ST_ERROR;
            break;
        }
        h = h_next;
        if (h < kHmin) {
            if (!warned && tid == 0 && d.primary)
                printf("Warning! Timestep %g being too small! Imposing a minimum timestep of %g, t = %g\n", h, kHmin,
                       t);
            warned = 1;
            h = kHmin;
        }
        if (finished) {
            status = ST_DONE;
            break;
        }
    }

    // ---- store this lane's component of the state
    if (comp) {
        long long g = gi;
        asm volatile("" : "+v"(g));
        double *px = c == 0 ? A.x : (c == 1 ? A.y : A.z);
        double *pv = c == 0 ? A.vx : (c == 1 ? A.vy : A.vz);
        double *ps = A.state;
        px[g] = y0;
        pv[g] = dy0;
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            ps[(long long)(c * 7 + j) * T + g] = bb[j];
            ps[(long long)(21 + c * 7 + j) * T + g] = gg[j];
        }
        ps[(long long)(42 + c) * T + g] = a0;
        ps[(long long)(45 + c) * T + g] = b6p;
    }
    if (tid == 0) {
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

// Instantiates the quad form's group sizes for one (force mode, controller).
template <int FM, int CT>
hipError_t launch_gr15q_fm_ct(int q, const GR15Args &a, int nb, size_t lds, hipStream_t st) {
    switch (q) {
        case 0: hipLaunchKernelGGL((gr15q_kernel<64, FM, CT>), dim3(nb), dim3(64), lds, st, a); break;
        case 1: hipLaunchKernelGGL((gr15q_kernel<128, FM, CT>), dim3(nb), dim3(128), lds, st, a); break;
        case 2: hipLaunchKernelGGL((gr15q_kernel<256, FM, CT>), dim3(nb), dim3(256), lds, st, a); break;
        case 3: hipLaunchKernelGGL((gr15q_kernel<512, FM, CT>), dim3(nb), dim3(512), lds, st, a); break;
        case 4: hipLaunchKernelGGL((gr15q_kernel<1024, FM, CT>), dim3(nb), dim3(1024), lds, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace abie
