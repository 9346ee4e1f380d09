// runtime.hip -- host side of libabie_hip.so: device memory, the
// reference-compatible C ABI (src/common.c:200-413 semantics) and the
// ensemble / force entry points of abie_hip.h.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <numeric>
#include <thread>
#include <string>
#include <vector>

#include "abie_hip.h"
#include "devmath.hpp"
#include "force.hpp"
#include "gr15.hpp"

namespace abie {

// ---------------------------------------------------------------- errors
static thread_local char g_errbuf[512] = "";

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errbuf, sizeof g_errbuf, fmt, ap);
    va_end(ap);
    fprintf(stderr, "libabie_hip: %s\n", g_errbuf);
    return code;
}

// for the entry points of other translation units
int report_error(int code, const char *msg) { return fail(code, "%s", msg); }

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return fail(ABIE_HIP_EHIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------- device memory
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t want) {
        if (want <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// index of the one-lane (256 threads, one body each) configuration
static int one_lane_256() {
    for (int c = 0; c < kNumCfg; ++c)
        if (kCfgs[c].nt == 256 && kCfgs[c].bpt == 1) return c;
    return 0;
}

static int device_cu_count(int dev) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) return 256;
    return v;
}

static int force_mode_from_env() {
    const char *s = getenv("ABIE_HIP_FORCE_MODE");
    if (!s) return ABIE_FORCE_FP64;
    if (!strcmp(s, "exact")) return ABIE_FORCE_FP64_EXACT;
    if (!strcmp(s, "mixed")) return ABIE_FORCE_MIXED;
    return ABIE_FORCE_FP64;
}

// Engine order of n bodies: massive (m > 0) ascending, then the rest.
static int engine_order(const double *m, int n, std::vector<int> &perm) {
    perm.resize(n);
    int k = 0;
    for (int i = 0; i < n; ++i)
        if (m[i] > 0.0) perm[k++] = i;
    const int nm = k;
    for (int i = 0; i < n; ++i)
        if (!(m[i] > 0.0)) perm[k++] = i;
    return nm;
}

}  // namespace abie

using namespace abie;

// ===================================================================== ensemble
// The split tail of a hybrid run: child units that each carry every massive
// body of a parent unit and a share of its test particles, in body arrays of
// their own.  Test particles never act on the massive bodies or on the step
// under the massive-only controllers, so the children evolve the parent's
// massive bodies and step sequence exactly, with fewer bodies each: lower
// latency per evaluation on the CUs that the finished units left idle.
struct TailSet {
    int n = 0;                          // child units
    long long total = 0;                // child bodies
    std::vector<int> parent;            // per child: parent unit
    std::vector<SysDesc> desc;          // per child (body0 into these arrays)
    std::vector<long long> gmap, smap;  // per child body: parent body (gather); parent body or -1 (scatter)
    DBuf<double> x, y, z, vx, vy, vz, gm, m, rad, state, mratio;
    DBuf<int> orig, inv, order;
    DBuf<unsigned char> del;
    DBuf<SysDesc> sys;
    DBuf<SysCtl> ctl;
    DBuf<long long> d_gmap, d_smap;
    DBuf<SysCtl> merged;               // parents' merged control blocks (ts_finish)
    DBuf<int> merged_idx;
    void release() {
        for (auto *b : {&x, &y, &z, &vx, &vy, &vz, &gm, &m, &rad, &state, &mratio}) b->release();
        orig.release(); inv.release(); order.release(); del.release(); sys.release(); ctl.release();
        d_gmap.release(); d_smap.release(); merged.release(); merged_idx.release();
    }
};

struct abie_ensemble {
    int dev = 0;
    int n_sys = 0;
    long long total = 0;
    double G = 0;
    int force_mode = ABIE_FORCE_FP64, ctrl = ABIE_CTRL_PLAIN;
    int events = 1;
    double ce_dist = 0.0;
    int max_ce = 0, max_coll = 0;
    int max_attempts = 1 << 30;       // per launch
    int link_above = 1 << 30;         // plain controller: link systems of more bodies than this
    int link_cap = 0;                 // bodies per linked slice (0: the slice size)
    long long clock_khz = 0;          // device wall-clock rate
    int n_units = 0;                   // work units (a system, or a slice of one)
    std::vector<int> offsets;          // [n_sys+1] original layout
    std::vector<long long> eng2orig;   // engine index -> original global index (-1: replica)
    std::vector<int> unit_sys, sys_unit0;
    std::vector<SysDesc> desc;         // per unit
    std::vector<int> order;            // grouped by cfg
    // quad form (four lanes per body) of the massive-only controllers: 0 off
    // (one lane per body throughout), 1 hybrid (one-lane form while the GPU
    // is full, the quad form once no more units run than there are CUs), 2
    // the quad form from the start.  Off when the event scan is asked for.
    int quad = 1;
    bool tail = false;                 // hybrid: the quad form has taken over
    std::vector<int> cur_cfg;          // per unit: the form of its next launch
    std::vector<int> alt_quad;         // per unit: its quad group size, or -1
    std::vector<int> lpt;              // units that are not linked slices, longest first
    DBuf<int> d_running;
    DBuf<long long> d_clock;
    bool tail_split = true;            // hybrid: split the tail's units over idle CUs
    bool promote = true;               // hybrid: stragglers move to the quad form early
    bool trace_sched = false;          // ABIE_HIP_TRACE_SCHED: one stderr line per launch round
    TailSet ts;
    std::vector<double> mratio;        // per unit (host copy of d_mratio)
    int group_off[kNumAllCfg + 1] = {0};
    // Launch lanes: four streams created one after the other, which gives
    // them distinct hardware queues (HIP's default is four per process).
    // Launches on streams that share a queue run one after the other, so a
    // round's launches are assigned to the lanes, heaviest first.
    static constexpr int kLanes = 4;
    hipStream_t streams[kLanes] = {};
    int lanes_distinct = 0;            // lanes that passed the concurrency probe
    DBuf<int> d_live;                  // running units' indices (ens_run's tail rounds)
    DBuf<SysCtl> d_live_ctl;
    std::vector<double> tau_sys;       // shortest orbital time scale per system at the start
    DBuf<unsigned> d_probe;
    DBuf<SysDesc> d_sys;
    DBuf<int> d_order, d_orig, d_inv;
    DBuf<SysCtl> d_ctl;
    DBuf<double> d_x, d_y, d_z, d_vx, d_vy, d_vz, d_gm, d_m, d_rad, d_ext, d_state, d_scan, d_ring_ce,
        d_ring_coll, d_log, d_energy;
    DBuf<unsigned char> d_del;
    DBuf<unsigned long long> d_log_len;
    DBuf<float> d_mlp;
    DBuf<double> d_mratio;             // per unit
    // linked systems (plain controller, larger than one workgroup): each runs
    // as one cooperative launch of its slices; {cfg, first entry in order,
    // slices}.  Their units are not in the configuration groups.
    std::vector<std::array<int, 3>> links;
    DBuf<unsigned long long> d_link_count;  // per unit (used at each system's link_base)
    DBuf<double> d_link_part;               // [2][n_units][2]
    DBuf<int> d_sys_inv;                    // bodies of each linked system by original index
    long long log_cap = 0;
    bool has_ext = false, has_mlp = false;
};

static int ens_build(abie_ensemble *e, const int *offsets, const double *pos, const double *vel, const double *m,
                     const double *r, const double *ext) {
    const int S = e->n_sys;
    e->offsets.assign(offsets, offsets + S + 1);
    // bodies per slice when a system is split (massive-only controllers).
    // Small slices bound the time of the longest unit, which sets the
    // ensemble's makespan: 4,096 planetary systems of 3..64 massive bodies
    // and 0..2,048 test particles ran 0.2 yr at 512, 811, 1,055 system-yr/s
    // with 1024-, 512- and 256-body slices (128: 1,020).
    // A lone system too: 1 star + 4 planets + 100,000 test particles, 100 yr,
    // 2.44 us per evaluation with 256-body slices (399 units), 2.53 with 201
    // (511), 2.57 with 400 or 512, 2.79 with 160 and 3.46 with 128 (814
    // units of two waves, placed unevenly over the CUs).
    int cap = 256;
    if (const char *sl = getenv("ABIE_HIP_SLICE")) cap = std::max(128, atoi(sl));
    // plain controller: systems of more than link_above bodies run as linked
    // slices even when one workgroup would hold them (default: above 256
    // bodies for a lone system, otherwise only when no workgroup holds it)
    int link_above = e->link_above;
    if (S == 1 && link_above == (1 << 30)) link_above = 256;  // a lone system: linked slices over more CUs
    if (const char *la = getenv("ABIE_HIP_LINK_ABOVE")) link_above = atoi(la);
    // linked slices of a lone system: 128 bodies (below)
    const int link_cap = e->link_cap > 0 ? e->link_cap : getenv("ABIE_HIP_SLICE") ? cap : S == 1 ? 128 : cap;
    const bool auto_link_cap = getenv("ABIE_HIP_SLICE") == nullptr;
    const int n_cu = device_cu_count(e->dev);
    link_above = std::max(link_above, link_cap);
    // units: a system that fits one workgroup is one unit; with the
    // massive-only controllers a larger system is split into slices that
    // each carry every massive body plus a share of the test particles (the
    // slices evolve the massive bodies identically, so they stay in step).
    struct Unit {
        int sys, nm, t0, t1;  // test-particle range [t0, t1) in engine order
        int link_n, link_rank, cfg;  // slices of a linked system (plain controller); cfg forced if >= 0
    };
    std::vector<Unit> units;
    std::vector<std::vector<int>> perms(S);
    std::vector<int> nms(S);
    for (int s = 0; s < S; ++s) {
        const int o = offsets[s], n = offsets[s + 1] - offsets[s];
        if (n <= 0) return fail(ABIE_HIP_EINVAL, "system %d has no bodies", s);
        const int nm = engine_order(m + o, n, perms[s]);
        nms[s] = nm;
        const int nt = n - nm;
        // (with at most one massive body the slices control their steps on
        // their own bodies, norm_all; the lone star then moves at constant
        // velocity in each slice, equal to rounding)
        const bool slice = e->ctrl != ABIE_CTRL_PLAIN && n > cap && nm <= 512 && cap - nm >= 64;
        // the plain controller's norms cover every body: a system larger than
        // one workgroup runs as linked slices that combine their maxima
        // slice size: keep a linked system to about 256 slices (1 star + 4
        // planets + 100,000 test particles ran 4.5 / 5.7 / 6.9 / 4.0
        // system-yr/s with 128 / 256 / 512 / 1,024-body slices; the
        // configuration-0 system, 1,005 bodies, runs fastest with 128), then
        // grow it until every slice fits on the GPU at once
        int lcap = link_cap;
        if (auto_link_cap)
            while (lcap < 512 && n > 256 * lcap) lcap *= 2;
        while (lcap - nm < std::min(64, lcap / 2) && lcap < 1024) lcap *= 2;
        auto n_slices = [&](int c) { return (nt + (c - nm) - 1) / std::max(1, c - nm); };
        while (lcap < 1024 && pick_cfg(lcap, nm) >= 0 &&
               n_slices(lcap) > n_cu * gr15_blocks_per_cu(pick_cfg(lcap, nm)))
            lcap *= 2;
        const bool link = e->ctrl == ABIE_CTRL_PLAIN && (pick_cfg(n, nm) < 0 || n > link_above) && n > lcap &&
                          lcap - nm >= std::min(64, lcap / 2) && nt > 0 && pick_cfg(lcap, nm) >= 0;
        if (!slice && !link && pick_cfg(n, nm) >= 0) {
            units.push_back({s, nm, nm, n, 1, 0, -1});
        } else if (slice || link) {
            const int per = (link ? lcap : cap) - nm;
            const int k = (nt + per - 1) / per;
            const int chunk = (nt + k - 1) / k;
            const int lcfg = link ? pick_cfg(nm + chunk, nm) : -1;  // one kernel for all slices
            if (link && (k < 2 || lcfg < 0))
                return fail(ABIE_HIP_ETOOBIG, "system %d: %d bodies / %d massive cannot be linked", s, n, nm);
            if (link && k > n_cu * gr15_blocks_per_cu(lcfg))
                return fail(ABIE_HIP_ETOOBIG, "system %d: %d bodies need %d linked slices of %d bodies, more "
                            "than the %d CUs hold at once", s, n, k, nm + chunk, n_cu);
            for (int q = 0; q < k; ++q)
                units.push_back({s, nm, nm + q * chunk, std::min(n, nm + (q + 1) * chunk), link ? k : 1,
                                 link ? q : 0, lcfg});
        } else {
            return fail(ABIE_HIP_ETOOBIG,
                        "system %d: %d bodies / %d massive exceeds one workgroup (max %d bodies, %d massive) and "
                        "cannot be sliced (a slice carries every massive body plus at least 64 test particles)",
                        s, n, nm, max_bodies_one_group(), src_cap(1024));
        }
    }
    // engine form of a unit that is not a linked slice: the quad form under
    // the massive-only controllers when the unit fits it, else the
    // configuration of the tuning rules
    auto quad_of = [&](int n) {
        return (e->ctrl != ABIE_CTRL_PLAIN && e->quad != 0 && n <= kQuadMaxBodies) ? pick_quad(n) : -1;
    };
    // the hybrid schedule runs the one-lane units of the massive-only
    // controllers (at most 256 bodies) in one launch, so that none waits on
    // a hardware queue behind another shape's: in the shape the tuning
    // rules give them when that is the same for all (a lone system's
    // slices: 1 star + 4 planets + 100,000 test particles in 128-body
    // slices fit the GPU at once as (128, 1) groups, not as (256, 1)),
    // else in the (256, 1) shape
    auto hybrid_small = [&](int n, int nm) {
        return e->quad == 1 && quad_of(n) >= 0 && cfg_fits(one_lane_256(), n, nm);
    };
    int small_cfg = -1;
    for (const Unit &un : units) {
        const int n = un.nm + (un.t1 - un.t0);
        if (un.cfg >= 0 || !hybrid_small(n, un.nm)) continue;
        int c = pick_cfg(n, un.nm);
        if (c < 0) c = one_lane_256();
        small_cfg = (small_cfg == -1 || small_cfg == c) ? c : one_lane_256();
    }
    auto unit_cfg = [&](int n, int nm) {
        const int q = quad_of(n);
        if (q >= 0 && e->quad == 2) return kNumCfg + q;
        if (hybrid_small(n, nm)) return small_cfg;
        return pick_cfg(n, nm);
    };
    const int U = (int)units.size();
    e->n_units = U;
    long long T = 0;
    for (auto &u : units) T += u.nm + (u.t1 - u.t0);
    e->total = T;
    e->desc.resize(U);
    e->unit_sys.resize(U);
    e->sys_unit0.assign(S, -1);
    e->eng2orig.assign(T, -1);
    std::vector<double> x(T), y(T), z(T), vx(T), vy(T), vz(T), gm(T), mm(T), rad(T), ex;
    std::vector<int> orig(T), inv(T, 0);
    std::vector<unsigned char> del(T);
    if (ext) ex.resize(3 * T);
    long long g0 = 0;
    for (int u = 0; u < U; ++u) {
        const Unit &un = units[u];
        const int s = un.sys, o = offsets[s];
        const bool first = e->sys_unit0[s] < 0;
        if (first) e->sys_unit0[s] = u;
        e->unit_sys[u] = s;
        const int n = un.nm + (un.t1 - un.t0);
        e->desc[u] = SysDesc{(int)g0, n, un.nm, un.cfg >= 0 ? un.cfg : unit_cfg(n, un.nm), (offsets[s + 1] - o) - un.nm,
                              s, first ? 1 : 0, un.nm <= 1 ? 1 : 0, un.link_n, un.link_rank, u - un.link_rank,
                              offsets[s + 1] - o, 0};
        for (int p = 0; p < n; ++p) {
            const int ep = p < un.nm ? p : un.t0 + (p - un.nm);  // engine position in the whole system
            const int i = perms[s][ep];
            const long long g = g0 + p, gi = o + i;
            if (first || p >= un.nm) e->eng2orig[g] = gi;
            x[g] = pos[3 * gi]; y[g] = pos[3 * gi + 1]; z[g] = pos[3 * gi + 2];
            vx[g] = vel[3 * gi]; vy[g] = vel[3 * gi + 1]; vz[g] = vel[3 * gi + 2];
            mm[g] = m[gi];
            gm[g] = (p < un.nm) ? e->G * m[gi] : 0.0;
            rad[g] = r ? r[gi] : 0.0;
            del[g] = m[gi] < 0.0;
            orig[g] = i;
            if (ext)
                for (int c = 0; c < 3; ++c) ex[c * T + g] = ext[3 * gi + c];
        }
        // the unit's bodies in ascending original index (the order of the
        // reference's event loop, common.c:119-123)
        std::vector<int> byorig(n);
        std::iota(byorig.begin(), byorig.end(), 0);
        std::sort(byorig.begin(), byorig.end(), [&](int a, int b) { return orig[g0 + a] < orig[g0 + b]; });
        for (int q = 0; q < n; ++q) inv[g0 + q] = byorig[q];
        g0 += n;
    }
    // linked systems: their bodies (the primary slice's massive bodies once)
    // by ascending original index, the order of the exact event pass
    std::vector<int> sysinv;
    for (int u = 0; u < U;) {
        const int ln = e->desc[u].link_n;
        if (ln <= 1) {
            ++u;
            continue;
        }
        const int first = (int)sysinv.size();
        std::vector<std::pair<int, int>> lst;  // (original index, body slot)
        for (int r = 0; r < ln; ++r) {
            const SysDesc &dr = e->desc[u + r];
            for (int p = (r == 0 ? 0 : dr.nm); p < dr.n; ++p) lst.push_back({orig[dr.body0 + p], dr.body0 + p});
        }
        std::sort(lst.begin(), lst.end());
        for (auto &x : lst) sysinv.push_back(x.second);
        for (int r = 0; r < ln; ++r) e->desc[u + r].sys_inv0 = first;
        u += ln;
    }
    // Work per unit for the start order (the queue of units waiting for a
    // group slot, longest first): evaluations per model year taken as
    // proportional to 1 / tau, tau = min over massive pairs of
    // sqrt(r^3 / G (m_i + m_j)) at the initial positions (the shortest
    // orbital time scale, which sets the step), times the work of an
    // evaluation, n (N_m + 8) (8 sources' worth of per-body step work).  A
    // system with close pairs needs 30-100 times the median evaluations per
    // model year; ordered by size alone it starts late and ends the run
    // alone.  (4,096 planetary systems, learned proposer, mixed forces:
    // 0.2 / 1 / 10 yr in 0.222 / 0.738 / 6.84 s against 0.235 / 0.955 /
    // 8.17 s by size.)  ABIE_HIP_ORDER=0: by size, n N_m.
    std::vector<double> uwork(U);
    {
        int order_mode = 1;
        if (const char *s = getenv("ABIE_HIP_ORDER")) order_mode = atoi(s);
        std::vector<double> &tau = e->tau_sys;
        tau.assign(S, 1.0);
        // (a lone system's units share one time scale, and a system of more
        // than 1,024 massive bodies would cost O(N_m^2) host time here: then
        // the order is by size)
        if (S == 1) order_mode = 0;
        for (int sy = 0; sy < S && order_mode == 1; ++sy)
            if (nms[sy] > 1024) order_mode = 0;
        for (int sy = 0; sy < S && order_mode == 1; ++sy) {
            // engine order: the system's massive bodies first
            const int o = offsets[sy], nm = nms[sy];
            const std::vector<int> &pm = perms[sy];
            double best = DBL_MAX;
            for (int a = 0; a < nm; ++a) {
                const long long i = o + pm[a];
                for (int b = a + 1; b < nm; ++b) {
                    const long long j = o + pm[b];
                    const double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                                 dz = pos[3 * i + 2] - pos[3 * j + 2];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    best = std::min(best, r2 * std::sqrt(r2) / (e->G * (m[i] + m[j])));
                }
            }
            tau[sy] = best < DBL_MAX && best > 0.0 ? std::sqrt(best) : 1.0;
        }
        for (int u = 0; u < U; ++u)
            uwork[u] = order_mode == 1 ? (double)e->desc[u].n * (e->desc[u].nm + 8) / tau[e->unit_sys[u]]
                                       : (double)e->desc[u].n * e->desc[u].nm;
    }
    // LPT order inside each configuration group (largest work first)
    e->order.clear();
    for (int c = 0; c < kNumAllCfg; ++c) {
        e->group_off[c] = (int)e->order.size();
        std::vector<int> grp;
        for (int u = 0; u < U; ++u)
            if (e->desc[u].cfg == c && e->desc[u].link_n <= 1) grp.push_back(u);
        std::stable_sort(grp.begin(), grp.end(), [&](int a, int b) { return uwork[a] > uwork[b]; });
        e->order.insert(e->order.end(), grp.begin(), grp.end());
    }
    e->group_off[kNumAllCfg] = (int)e->order.size();
    e->lpt.assign(e->order.begin(), e->order.end());
    std::stable_sort(e->lpt.begin(), e->lpt.end(), [&](int a, int b) { return uwork[a] > uwork[b]; });
    e->cur_cfg.resize(U);
    e->alt_quad.resize(U);
    for (int u = 0; u < U; ++u) {
        e->cur_cfg[u] = e->desc[u].cfg;
        // a unit built with the tree of the quad form can move to it (one-lane
        // units of one body per thread under the massive-only controllers)
        const int q = quad_of(e->desc[u].n);
        e->alt_quad[u] = (q >= 0 && e->desc[u].link_n <= 1 &&
                          (cfg_is_quad(e->desc[u].cfg) || kCfgs[e->desc[u].cfg].bpt == 1)) ? q : -1;
    }
    e->tail = false;
    e->links.clear();
    for (int u = 0; u < U;) {
        const SysDesc &du = e->desc[u];
        if (du.link_n <= 1) {
            ++u;
            continue;
        }
        e->links.push_back({du.cfg, (int)e->order.size(), du.link_n});
        for (int r = 0; r < du.link_n; ++r) e->order.push_back(u + r);
        u += du.link_n;
    }

    HIP_TRY(e->d_sys.reserve(U));
    HIP_TRY(e->d_order.reserve(U));
    HIP_TRY(e->d_ctl.reserve(U));
    HIP_TRY(e->d_energy.reserve(U));
    for (auto *b : {&e->d_x, &e->d_y, &e->d_z, &e->d_vx, &e->d_vy, &e->d_vz, &e->d_gm, &e->d_m, &e->d_rad})
        HIP_TRY(b->reserve(T));
    HIP_TRY(e->d_orig.reserve(T));
    HIP_TRY(e->d_inv.reserve(T));
    HIP_TRY(e->d_del.reserve(T));
    HIP_TRY(e->d_state.reserve((size_t)kStateDoubles * T));
    HIP_TRY(e->d_scan.reserve(4 * T));  // x, y, z of the event scan, then the candidate flags
    HIP_TRY(hipMemset(e->d_scan.p, 0, 4 * T * sizeof(double)));
    HIP_TRY(e->d_log_len.reserve(1));
    HIP_TRY(e->d_link_count.reserve(U));
    HIP_TRY(e->d_link_part.reserve((size_t)4 * U));
    HIP_TRY(e->d_sys_inv.reserve(std::max<size_t>(1, sysinv.size())));
    HIP_TRY(hipMemset(e->d_link_count.p, 0, U * sizeof(unsigned long long)));
    if (!sysinv.empty())
        HIP_TRY(hipMemcpy(e->d_sys_inv.p, sysinv.data(), sysinv.size() * sizeof(int), hipMemcpyHostToDevice));
    e->has_ext = ext != nullptr;
    if (ext) {
        HIP_TRY(e->d_ext.reserve(3 * T));
        HIP_TRY(hipMemcpy(e->d_ext.p, ex.data(), 3 * T * sizeof(double), hipMemcpyHostToDevice));
    }
    auto up = [&](double *d, const std::vector<double> &h) {
        return hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
    };
    HIP_TRY(up(e->d_x.p, x)); HIP_TRY(up(e->d_y.p, y)); HIP_TRY(up(e->d_z.p, z));
    HIP_TRY(up(e->d_vx.p, vx)); HIP_TRY(up(e->d_vy.p, vy)); HIP_TRY(up(e->d_vz.p, vz));
    HIP_TRY(up(e->d_gm.p, gm)); HIP_TRY(up(e->d_m.p, mm)); HIP_TRY(up(e->d_rad.p, rad));
    HIP_TRY(hipMemcpy(e->d_orig.p, orig.data(), T * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_inv.p, inv.data(), T * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_del.p, del.data(), T, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_sys.p, e->desc.data(), U * sizeof(SysDesc), hipMemcpyHostToDevice));
    // mass ratio of each unit's system: (sum of the other massive masses) / largest
    {
        std::vector<double> mr(U, 0.0);
        for (int u = 0; u < U; ++u) {
            const int s = e->unit_sys[u], o = offsets[s], n = offsets[s + 1] - o;
            double mx = 0.0, sum = 0.0;
            for (int i = 0; i < n; ++i)
                if (m[o + i] > 0.0) {
                    mx = std::max(mx, m[o + i]);
                    sum += m[o + i];
                }
            mr[u] = mx > 0.0 ? (sum - mx) / mx : 0.0;
        }
        HIP_TRY(e->d_mratio.reserve(U));
        HIP_TRY(hipMemcpy(e->d_mratio.p, mr.data(), U * sizeof(double), hipMemcpyHostToDevice));
        e->mratio = mr;
    }
    HIP_TRY(hipMemcpy(e->d_order.p, e->order.data(), U * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(e->d_state.p, 0, (size_t)kStateDoubles * T * sizeof(double)));
    std::vector<SysCtl> ctl(U);
    for (auto &c : ctl) {
        memset(&c, 0, sizeof c);
        c.status = ST_RUNNING;
    }
    HIP_TRY(hipMemcpy(e->d_ctl.p, ctl.data(), U * sizeof(SysCtl), hipMemcpyHostToDevice));
    if (!e->streams[0] && e->group_off[kNumAllCfg] > 0) {
        // four lanes on distinct hardware queues: each new stream joins the
        // lanes if all of them then run a probe kernel at the same time;
        // one that shares a queue with a lane is kept aside (holding its
        // queue, so that the next stream goes elsewhere) and released after.
        // (Not for an ensemble of linked slices only, such as the drop-in's
        // engine: those launch on one lane.)
        int khz = 0;
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->dev));
        HIP_TRY(e->d_probe.reserve(2));
        std::vector<hipStream_t> aside;
        int have = 0;
        for (int attempt = 0; attempt < 16 && have < abie_ensemble::kLanes; ++attempt) {
            HIP_TRY(hipStreamCreateWithFlags(&e->streams[have], hipStreamNonBlocking));
            bool ok = false;
            HIP_TRY(streams_concurrent(e->streams, have + 1, e->d_probe.p, (long long)(2 * (long long)khz), &ok));
            if (ok) {
                ++have;
            } else {
                aside.push_back(e->streams[have]);
                e->streams[have] = nullptr;
            }
        }
        for (auto st : aside) (void)hipStreamDestroy(st);
        e->lanes_distinct = have;
    }
    for (auto &st : e->streams)
        if (!st) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    return 0;
}

static int ens_rings(abie_ensemble *e, int max_ce, int max_coll) {
    e->max_ce = max_ce;
    e->max_coll = max_coll;
    // Without rings nothing observable depends on the close-encounter /
    // collision scan (the event counters only stop a system through the
    // rings' limits), so the O(n^2) pass after each accepted step is skipped.
    e->events = (max_ce > 0 || max_coll > 0) ? 1 : 0;
    if (max_ce > 0) {
        HIP_TRY(e->d_ring_ce.reserve((size_t)4 * max_ce * e->n_units));
        HIP_TRY(hipMemset(e->d_ring_ce.p, 0, (size_t)4 * max_ce * e->n_units * sizeof(double)));
    }
    if (max_coll > 0) {
        HIP_TRY(e->d_ring_coll.reserve((size_t)4 * max_coll * e->n_units));
        HIP_TRY(hipMemset(e->d_ring_coll.p, 0, (size_t)4 * max_coll * e->n_units * sizeof(double)));
    }
    return 0;
}

static GR15Args ens_args(abie_ensemble *e) {
    GR15Args a;
    memset(&a, 0, sizeof a);
    a.sys = e->d_sys.p;
    a.order = e->d_order.p;
    a.ctl = e->d_ctl.p;
    a.x = e->d_x.p; a.y = e->d_y.p; a.z = e->d_z.p;
    a.vx = e->d_vx.p; a.vy = e->d_vy.p; a.vz = e->d_vz.p;
    a.gm = e->d_gm.p;
    a.rad = e->d_rad.p;
    a.ext = e->has_ext ? e->d_ext.p : nullptr;
    a.orig = e->d_orig.p;
    a.inv = e->d_inv.p;
    a.deleted = e->d_del.p;
    a.state = e->d_state.p;
    a.scan = e->d_scan.p;
    a.total = e->total;
    a.ce_dist = e->ce_dist;
    a.max_ce = e->max_ce;
    a.max_coll = e->max_coll;
    a.ring_ce = e->max_ce > 0 ? e->d_ring_ce.p : nullptr;
    a.ring_coll = e->max_coll > 0 ? e->d_ring_coll.p : nullptr;
    a.log = e->log_cap > 0 ? e->d_log.p : nullptr;
    a.log_len = e->d_log_len.p;
    a.log_cap = e->log_cap;
    a.mlp = e->has_mlp ? e->d_mlp.p : nullptr;
    a.mratio = e->d_mratio.p;
    a.max_attempts = e->max_attempts;
    // units of 8+ massive bodies use the prefetching source loop: a lone
    // unit of 20 + 236 bodies 3.07 -> 2.53 us per evaluation, of 12 + 244
    // 2.27 -> 2.01 us, 2,048 of them +2-4 % evaluations per second; same bits
    a.prefetch = 8;
    if (const char *s = getenv("ABIE_HIP_PREFETCH")) a.prefetch = std::max(0, atoi(s));
    a.events = e->events;
    a.link_count = e->d_link_count.p;
    a.link_part = e->d_link_part.p;
    a.sys_inv = e->d_sys_inv.p;
    a.n_units = e->n_units;
    // a grid barrier waits at most 5 s (the slices of one launch are all
    // resident and do the same work between barriers)
    a.spin_ticks = (long long)(5e3 * (double)e->clock_khz);
    return a;
}

static int ens_set_t_end(abie_ensemble *e, double t_end, const double *t_start) {
    std::vector<SysCtl> ctl(e->n_units);
    HIP_TRY(hipMemcpy(ctl.data(), e->d_ctl.p, e->n_units * sizeof(SysCtl), hipMemcpyDeviceToHost));
    for (int u = 0; u < e->n_units; ++u) {
        ctl[u].t_end = t_end;
        if (t_start) ctl[u].t = t_start[e->unit_sys[u]];
        if (ctl[u].status == ST_DONE && ctl[u].t <= t_end) ctl[u].status = ST_RUNNING;
        ctl[u].start_tick = ctl[u].done_tick = 0;
    }
    HIP_TRY(hipMemcpy(e->d_ctl.p, ctl.data(), e->n_units * sizeof(SysCtl), hipMemcpyHostToDevice));
    return 0;
}

static BodyArrays main_bodies(abie_ensemble *e) {
    return BodyArrays{e->d_x.p, e->d_y.p, e->d_z.p, e->d_vx.p, e->d_vy.p, e->d_vz.p, e->d_gm.p, e->d_m.p,
                      e->d_rad.p, e->d_state.p, e->d_orig.p, e->d_inv.p, e->d_del.p, e->total};
}

static BodyArrays tail_bodies(TailSet &t) {
    return BodyArrays{t.x.p, t.y.p, t.z.p, t.vx.p, t.vy.p, t.vz.p, t.gm.p, t.m.p, t.rad.p, t.state.p,
                      t.orig.p, t.inv.p, t.del.p, t.total};
}

// The arguments of a launch over the split tail's children.
static GR15Args tail_args(abie_ensemble *e, const GR15Args &base) {
    TailSet &t = e->ts;
    GR15Args a = base;
    a.sys = t.sys.p;
    a.ctl = t.ctl.p;
    a.order = t.order.p;
    a.x = t.x.p; a.y = t.y.p; a.z = t.z.p;
    a.vx = t.vx.p; a.vy = t.vy.p; a.vz = t.vz.p;
    a.gm = t.gm.p;
    a.rad = t.rad.p;
    a.ext = nullptr;
    a.orig = t.orig.p;
    a.inv = t.inv.p;
    a.deleted = t.del.p;
    a.state = t.state.p;
    a.scan = nullptr;
    a.total = t.total;
    a.mratio = t.mratio.p;
    a.n_units = t.n;
    return a;
}

// Move the running units of the quad form (those in *only, if given) to the
// tail set, split over at most n_cu units in all (one per CU).  Each child
// carries all of its parent's massive bodies and a share of its test
// particles (at least kTailMinTP), and makes exactly the parent's steps (test
// particles never enter the massive-only norms).  A unit's evaluation
// latency steps with the wave rows each SIMD runs, ceil(n / 64) for n bodies
// at four lanes per body, so a split only pays where its children need fewer
// rows than the parent: the unit with the most remaining time -- attempts
// still to come (rem[u]) times its rows times (16 + its massive bodies) --
// repeatedly moves to the next lower row count while the CUs allow it.
// max_rows > 0: every unit of the set is first split to at most that many
// rows (one group size), most remaining time first; a unit the CUs left
// cannot split that far stays out.  Copies the children's bodies out of the
// ensemble's arrays; split[u] = 1 for the units moved.  Units whose norms
// cover every body (norm_all) move whole: their step depends on every body.
constexpr int kTailMinTP = 4;
// The quad group a unit of n bodies runs in at the tail: 512 threads up to
// 128 bodies, 1,024 beyond (waves without bodies end at the kernel's start).
// Two group sizes at most, so the tail is at most two concurrent launches,
// on the two streams created first: launches on streams that share a
// hardware queue run one after the other.
inline int tail_q(int n) { return n <= kQuadNT[3] / 4 ? 3 : 4; }
// longest round at the tail (a round also ends when a quarter of its units finish)
constexpr double kTailRoundSeconds = 0.2;
static int ts_plan(abie_ensemble *e, const std::vector<SysCtl> &ctl, const std::vector<double> &rem, int n_cu,
                   std::vector<char> &split, const std::vector<char> *only = nullptr, int max_rows = 0) {
    TailSet &t = e->ts;
    t.n = 0;
    t.total = 0;
    t.parent.clear(); t.desc.clear(); t.gmap.clear(); t.smap.clear();
    split.assign(e->n_units, 0);
    std::vector<int> run;
    for (int u : e->lpt)
        if (ctl[u].status == ST_RUNNING && e->alt_quad[u] >= 0 && (only == nullptr || (*only)[u]))
            run.push_back(u);
    auto rows = [](int n) { return (n + 63) / 64; };
    std::vector<int> kk(e->n_units, 1);
    int extra = n_cu - (int)run.size();
    if (max_rows > 0) {
        // every unit of the set runs in at most max_rows wave rows (a single
        // group size): most remaining time first, a unit the CUs left
        // cannot split that far stays out of the set
        std::stable_sort(run.begin(), run.end(), [&](int a, int b) { return rem[a] > rem[b]; });
        std::vector<int> keep;
        int used = 0;
        for (int u : run) {
            const SysDesc &d = e->desc[u];
            int want = 1;
            if (rows(d.n) > max_rows) {
                const int nt = d.n - d.nm, room = 64 * max_rows - d.nm;
                want = (!d.norm_all && room >= kTailMinTP && nt > 0) ? (nt + room - 1) / room : 0;
            }
            if (want == 0 || used + want > n_cu) continue;
            kk[u] = want;
            used += want;
            keep.push_back(u);
        }
        run.swap(keep);
        extra = n_cu - used;
    }
    if (run.empty()) return 0;
    auto child_n = [&](int u) {
        const SysDesc &d = e->desc[u];
        const int nt = d.n - d.nm;
        return d.nm + (nt > 0 ? (nt + kk[u] - 1) / kk[u] : 0);
    };
    std::vector<int> cand;
    for (int u : run) {
        const SysDesc &d = e->desc[u];
        if (!d.norm_all && d.link_n <= 1 && d.n - d.nm >= 2 * kTailMinTP && rows(child_n(u)) > 1) cand.push_back(u);
    }
    while (extra > 0 && !cand.empty()) {
        int bi = 0;
        double bt = -1.0;
        for (int i = 0; i < (int)cand.size(); ++i) {
            const SysDesc &d = e->desc[cand[i]];
            const double tt = rem[cand[i]] * rows(child_n(cand[i])) * (16.0 + d.nm);
            if (tt > bt) {
                bt = tt;
                bi = i;
            }
        }
        const int u = cand[bi];
        const SysDesc &d = e->desc[u];
        const int nt = d.n - d.nm;
        const int r = rows(child_n(u)) - 1;  // the next lower row count
        const int room = 64 * r - d.nm;                              // test particles per child there
        const int want = room >= kTailMinTP ? (nt + room - 1) / room : 0;
        if (r < 1 || want == 0 || want - kk[u] > extra) {
            cand.erase(cand.begin() + bi);  // no lower row count within reach
            continue;
        }
        extra -= want - kk[u];
        kk[u] = want;
    }
    // every unit that can run in the quad form moves to the tail's arrays,
    // split or whole (kk = 1), so that the tail is one launch per group size
    std::vector<SysCtl> cctl;
    for (int u : run) {
        const SysDesc &d = e->desc[u];
        const int nt = d.n - d.nm;
        const int chunk = nt > 0 ? (nt + kk[u] - 1) / kk[u] : 0;
        const int k = nt > 0 ? (nt + chunk - 1) / chunk : 1;
        split[u] = 1;
        for (int c = 0; c < k; ++c) {
            const int t0 = d.nm + c * chunk, t1 = std::min(d.n, t0 + chunk);
            const int n = d.nm + (t1 - t0);
            SysDesc cd = d;
            cd.body0 = (int)t.total;
            cd.n = n;
            cd.cfg = kNumCfg + tail_q(n);
            cd.primary = (d.primary && c == 0) ? 1 : 0;
            cd.link_n = 1;
            cd.link_rank = 0;
            cd.link_base = t.n;
            cd.sys_n = n;
            cd.sys_inv0 = 0;
            for (int p = 0; p < d.nm; ++p) {
                t.gmap.push_back(d.body0 + p);
                t.smap.push_back(c == 0 ? d.body0 + p : -1);  // the massive bodies come back from child 0
            }
            for (int p = t0; p < t1; ++p) {
                t.gmap.push_back(d.body0 + p);
                t.smap.push_back(d.body0 + p);
            }
            t.total += n;
            t.desc.push_back(cd);
            t.parent.push_back(u);
            cctl.push_back(ctl[u]);
            cctl.back().stop_attempt = 0;
            ++t.n;
        }
    }
    if (t.n == 0) return 0;
    const long long T = t.total;
    for (auto *b : {&t.x, &t.y, &t.z, &t.vx, &t.vy, &t.vz, &t.gm, &t.m, &t.rad}) HIP_TRY(b->reserve(T));
    HIP_TRY(t.state.reserve((size_t)kStateDoubles * T));
    HIP_TRY(t.orig.reserve(T)); HIP_TRY(t.inv.reserve(T)); HIP_TRY(t.del.reserve(T));
    HIP_TRY(t.sys.reserve(t.n)); HIP_TRY(t.ctl.reserve(t.n)); HIP_TRY(t.order.reserve(t.n));
    HIP_TRY(t.mratio.reserve(t.n));
    HIP_TRY(t.d_gmap.reserve(T)); HIP_TRY(t.d_smap.reserve(T));
    std::vector<double> mr(t.n);
    for (int c = 0; c < t.n; ++c) mr[c] = e->mratio[t.parent[c]];
    HIP_TRY(hipMemcpy(t.sys.p, t.desc.data(), t.n * sizeof(SysDesc), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.ctl.p, cctl.data(), t.n * sizeof(SysCtl), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.mratio.p, mr.data(), t.n * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.d_gmap.p, t.gmap.data(), T * sizeof(long long), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.d_smap.p, t.smap.data(), T * sizeof(long long), hipMemcpyHostToDevice));
    HIP_TRY(launch_body_gather(t.d_gmap.p, T, main_bodies(e), tail_bodies(t), e->streams[0]));
    HIP_TRY(hipStreamSynchronize(e->streams[0]));
    return 0;
}

// Launch the children listed in `which` (grouped by group size), with
// arguments `a` over the tail's arrays.
static int ts_launch(abie_ensemble *e, const GR15Args &a, const std::vector<int> &which) {
    TailSet &t = e->ts;
    std::vector<int> ord;
    int goff[kNumQuad + 1];
    for (int q = 0; q < kNumQuad; ++q) {
        goff[q] = (int)ord.size();
        for (int c : which)
            if (t.desc[c].cfg == kNumCfg + q) ord.push_back(c);
    }
    goff[kNumQuad] = (int)ord.size();
    if (ord.empty()) return 0;
    HIP_TRY(hipMemcpy(t.order.p, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice));
    for (int q = 0; q < kNumQuad; ++q) {
        const int nb = goff[q + 1] - goff[q];
        if (nb == 0) continue;
        GR15Args ac = a;
        ac.order = t.order.p + goff[q];
        HIP_TRY(launch_gr15q(q, e->force_mode, e->ctrl, ac, nb, e->streams[q == kNumQuad - 1 ? 1 : 0]));
    }
    return 0;
}

// After a round: bring the children of each parent to one attempt count
// (a child that stopped earlier runs on to its siblings' count; all of them
// make the same steps), then copy them back into the ensemble's arrays and
// the parent's control block, and drop them.
static int ts_finish(abie_ensemble *e, const GR15Args &base) {
    TailSet &t = e->ts;
    if (t.n == 0) return 0;
    std::vector<SysCtl> cc(t.n);
    for (int pass = 0;; ++pass) {
        HIP_TRY(hipMemcpy(cc.data(), t.ctl.p, t.n * sizeof(SysCtl), hipMemcpyDeviceToHost));
        std::vector<long long> target(e->n_units, -1);
        for (int c = 0; c < t.n; ++c) target[t.parent[c]] = std::max(target[t.parent[c]], cc[c].n_attempts);
        std::vector<int> lag;
        for (int c = 0; c < t.n; ++c)
            if (cc[c].n_attempts < target[t.parent[c]]) {
                if (cc[c].status != ST_RUNNING || pass > 0)
                    return fail(ABIE_HIP_EHIP, "split unit %d: a child stopped at attempt %lld of %lld", t.parent[c],
                                cc[c].n_attempts, target[t.parent[c]]);
                cc[c].stop_attempt = target[t.parent[c]];
                lag.push_back(c);
            }
        if (lag.empty()) break;
        HIP_TRY(hipMemcpy(t.ctl.p, cc.data(), t.n * sizeof(SysCtl), hipMemcpyHostToDevice));
        GR15Args a = tail_args(e, base);
        a.deadline_tick = 0;
        a.budget_ticks = 0;
        a.tail_units = 0;
        a.running = nullptr;
        int rc = ts_launch(e, a, lag);
        if (rc) return rc;
        for (auto st : e->streams) HIP_TRY(hipStreamSynchronize(st));
    }
    HIP_TRY(launch_body_scatter(t.d_smap.p, t.total, tail_bodies(t), main_bodies(e), e->streams[0]));
    // each parent's control block from its children (child 0's, with the
    // earliest start and the latest finish), written back in one scatter
    std::vector<SysCtl> merged;
    std::vector<int> idx;
    for (int c = 0; c < t.n;) {
        const int u = t.parent[c];
        SysCtl pc = cc[c];
        int c1 = c;
        for (; c1 < t.n && t.parent[c1] == u; ++c1) {
            if (cc[c1].start_tick > 0 && (pc.start_tick == 0 || cc[c1].start_tick < pc.start_tick))
                pc.start_tick = cc[c1].start_tick;
            pc.done_tick = std::max(pc.done_tick, cc[c1].done_tick);
        }
        pc.stop_attempt = 0;
        merged.push_back(pc);
        idx.push_back(u);
        c = c1;
    }
    const int np = (int)merged.size();
    HIP_TRY(t.merged.reserve(np));
    HIP_TRY(t.merged_idx.reserve(np));
    HIP_TRY(hipMemcpyAsync(t.merged.p, merged.data(), np * sizeof(SysCtl), hipMemcpyHostToDevice, e->streams[0]));
    HIP_TRY(hipMemcpyAsync(t.merged_idx.p, idx.data(), np * sizeof(int), hipMemcpyHostToDevice, e->streams[0]));
    HIP_TRY(launch_ctl_scatter(t.merged.p, t.merged_idx.p, np, e->d_ctl.p, e->streams[0]));
    HIP_TRY(hipStreamSynchronize(e->streams[0]));
    t.n = 0;
    return 0;
}

// Launch the running units, grouped by engine form and longest first, until
// all systems stop or time is up.  Under the hybrid schedule the one-lane
// units stop (state kept) once no more units run than there are CUs, and
// continue in the quad form: the same arithmetic, lower latency per unit.
// Once no more than n_cu / 8 units run, those are split over the CUs in
// proportion to their remaining work (ts_plan); the children run to the end
// of the run (or its deadline) and are then merged back.
static int ens_run(abie_ensemble *e, double max_seconds, int *done_out) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<SysCtl> ctl(e->n_units);
    int done = 0;
    HIP_TRY(hipSetDevice(e->dev));
    if (e->clock_khz == 0) {
        int khz = 0;
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->dev));
        e->clock_khz = khz;
    }
    // a per-group wall-clock budget only when asked for (tests of relaunches)
    double group_seconds = 0.0;
    if (const char *s = getenv("ABIE_HIP_LAUNCH_SECONDS")) group_seconds = atof(s);
    HIP_TRY(e->d_running.reserve(1));
    HIP_TRY(e->d_clock.reserve(1));
    const int n_cu = device_cu_count(e->dev);
    const bool hybrid = e->ctrl != ABIE_CTRL_PLAIN && e->quad == 1;
    const bool splitting = hybrid && e->tail_split && !e->has_ext;
    const int n_plain_units = e->group_off[kNumAllCfg];  // units that are not linked slices
    if (hybrid) {  // every run starts in the one-lane form (the same bits either way)
        e->tail = false;
        for (int u = 0; u < e->n_units; ++u) e->cur_cfg[u] = e->desc[u].cfg;
    }
    std::vector<int> ord;
    std::vector<char> split;
    std::vector<double> t_run0, rem;          // per unit: model time and attempts at the start of the run
    std::vector<long long> att_run0;
    ord.reserve(n_plain_units);
    // Stragglers (hybrid, before the tail): the run proceeds in rounds of a
    // growing length (0.25 s, then as long as the run so far, at most 1 s);
    // after each, a unit whose projected time to t_end (its model time left
    // over its progress rate while it ran in the round) exceeds 4 times the
    // median is promoted, most remaining time first, while it stays slower
    // than the median: it runs in the quad form, in a tail set of at most a
    // quarter of the CUs (one per unit, the rest splitting the units with
    // the most remaining time by the rows rule), launched ahead of the
    // one-lane units.  Its trajectory does not change (the same bits in
    // either form).
    const bool promoting = splitting && e->promote;
    int promote_cus = std::max(1, n_cu / 4);
    if (const char *s = getenv("ABIE_HIP_PROMOTE_CUS")) promote_cus = std::max(1, std::min(n_cu, atoi(s)));
    // (as many units as CUs: the 4,096-system ensemble ran 100 yr in 66.4 s
    // against 68.8 s with half as many, and 1 star + 4 planets + 100,000
    // test particles 1,000 yr in 9.42 s against 9.91 s)
    int promote_units = promote_cus;
    if (const char *s = getenv("ABIE_HIP_PROMOTE_UNITS")) promote_units = std::max(1, std::min(promote_cus, atoi(s)));
    double promote_factor = 4.0, round0 = 0.25;  // tests force promotions with short rounds
    if (const char *s = getenv("ABIE_HIP_PROMOTE_FACTOR")) promote_factor = atof(s);
    if (const char *s = getenv("ABIE_HIP_ROUND_SECONDS")) round0 = std::max(1e-3, atof(s));
    // Probe round (hybrid schedule with promotion, when the units of some
    // shape outnumber the groups of that shape the GPU holds at once; the
    // attempt count from ABIE_HIP_PROBE_ATTEMPTS, else chosen at the run's
    // start below): a first round of n attempts per unit
    // measures each one's model time and run time per attempt; the queue is
    // then re-sorted by the projected run time left (LPT on measured
    // times), and the promotion below can judge every unit after that
    // round.  Units stop and resume between rounds with the same bits.  The
    // start order from the initial orbital time scales (ens_build) costs
    // nothing but misses a system that turns heavy later: 4,096 planetary
    // systems, learned proposer, mixed forces, 0.2 / 1 / 10 / 100 / 1,000
    // yr in 0.26 / 0.86 / 6.98 / 66.7 / 829 s with 16 probe attempts
    // against 0.222 / 0.738 / 6.84 / 67.7 / 945 s without.
    int probe_attempts = 0, probe_env = -1;
    bool probe_fits = false;  // some shape's units outnumber its resident groups
    if (const char *s = getenv("ABIE_HIP_PROBE_ATTEMPTS")) probe_env = std::max(0, atoi(s));
    if (promoting) {
        std::vector<long long> per_cfg(kNumAllCfg, 0);
        for (int u = 0; u < e->n_units; ++u)
            if (e->desc[u].link_n <= 1) ++per_cfg[e->cur_cfg[u]];
        for (int c = 0; c < kNumCfg; ++c) probe_fits = probe_fits || per_cfg[c] > (long long)n_cu * gr15_blocks_per_cu(c);
    }
    bool probe_order = true;
    if (const char *s = getenv("ABIE_HIP_PROBE_ORDER")) probe_order = atoi(s) != 0;
    std::vector<char> promoted(e->n_units, 0);
    std::vector<double> t_prev(e->n_units, 0.0);
    std::vector<long long> busy_prev(e->n_units, 0);
    double el_prev = 0.0, round_s = round0;
    // control blocks each round: all of them, or, once few units run (the
    // tail), only those that ran in the last round (a finished unit's block
    // no longer changes), gathered on the device and read back in one copy
    std::vector<int> live;
    std::vector<SysCtl> live_ctl;
    for (int round = 0;; ++round) {
        if (round == 0 || live.empty() || (long long)live.size() * 8 > e->n_units) {
            HIP_TRY(hipMemcpy(ctl.data(), e->d_ctl.p, e->n_units * sizeof(SysCtl), hipMemcpyDeviceToHost));
        } else {
            const int nl = (int)live.size();
            HIP_TRY(e->d_live.reserve(nl));
            HIP_TRY(e->d_live_ctl.reserve(nl));
            HIP_TRY(hipMemcpy(e->d_live.p, live.data(), nl * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRY(launch_ctl_gather(e->d_ctl.p, e->d_live.p, nl, e->d_live_ctl.p, e->streams[0]));
            live_ctl.resize(nl);
            HIP_TRY(hipMemcpyAsync(live_ctl.data(), e->d_live_ctl.p, nl * sizeof(SysCtl), hipMemcpyDeviceToHost,
                                   e->streams[0]));
            HIP_TRY(hipStreamSynchronize(e->streams[0]));
            for (int i = 0; i < nl; ++i) ctl[live[i]] = live_ctl[i];
        }
        live.clear();
        for (int u = 0; u < e->n_units; ++u)
            if (ctl[u].status == ST_RUNNING) live.push_back(u);
        if (round == 0) {
            t_run0.resize(e->n_units);
            att_run0.resize(e->n_units);
            for (int u = 0; u < e->n_units; ++u) {
                t_run0[u] = ctl[u].t;
                att_run0[u] = ctl[u].n_attempts;
            }
            // by default a long run probes (median span of at least 1,000
            // shortest orbital time scales: 4,096 planetary systems ran
            // 1,000 yr in 829 s with the probe round, 945 s without, where
            // one system that turns heavy after ~24 yr, and ranks mid-queue
            // by its time scale, started late), a short one does not
            // (10 yr, median 154 time scales: 6.84 s without, 6.98 s with)
            if (promoting && probe_fits) {
                if (probe_env >= 0) {
                    probe_attempts = probe_env;
                } else {
                    std::vector<double> orbits;
                    for (int u = 0; u < e->n_units; ++u)
                        if (ctl[u].status == ST_RUNNING && e->desc[u].link_n <= 1)
                            orbits.push_back((ctl[u].t_end - ctl[u].t) / e->tau_sys[e->unit_sys[u]]);
                    if (!orbits.empty()) {
                        std::nth_element(orbits.begin(), orbits.begin() + orbits.size() / 2, orbits.end());
                        probe_attempts = orbits[orbits.size() / 2] >= 1000.0 ? 16 : 0;
                    }
                }
            }
        }
        int n_run = 0, n_run_units = 0;
        for (int u = 0; u < e->n_units; ++u)
            if (ctl[u].status == ST_RUNNING) {
                ++n_run;
                if (e->desc[u].link_n <= 1) ++n_run_units;
            }
        if (n_run == 0) {
            done = 1;
            break;
        }
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (max_seconds > 0 && el >= max_seconds) break;
        if (round == 1 && probe_attempts > 0 && probe_order) {
            std::vector<double> key(e->n_units, 0.0);
            double worst = 0.0;
            for (int u = 0; u < e->n_units; ++u) {
                const double att = (double)(ctl[u].n_attempts - att_run0[u]), dt = ctl[u].t - t_run0[u];
                if (ctl[u].status != ST_RUNNING || att <= 0) continue;
                const double busy = (double)(ctl[u].busy_ticks - busy_prev[u]);  // ticks in the first round
                key[u] = dt > 0 ? (ctl[u].t_end - ctl[u].t) / (dt / att) * (busy / att) : -1.0;
                worst = std::max(worst, key[u]);
            }
            for (auto &k : key)
                if (k < 0) k = 2.0 * worst + 1.0;  // no model time yet: first
            std::stable_sort(e->lpt.begin(), e->lpt.end(), [&](int a, int b) { return key[a] > key[b]; });
        }
        split.assign(e->n_units, 0);
        if (promoting && !e->tail && round > 0 && el > el_prev) {
            // only units that ran in the last round have a rate (with more
            // units than the GPU holds at once, many have not started yet)
            const double w = el - el_prev;
            std::vector<double> pu(e->n_units, -1.0), all;
            // projected run time left: model time left over the rate per
            // tick the unit was running in the last round (units that ran for
            // less than a tenth of the round are not judged)
            const double min_ticks = (round == 1 && probe_attempts > 0) ? 1.0 : 0.1 * w * 1e3 * (double)e->clock_khz;
            for (int u = 0; u < e->n_units; ++u) {
                if (ctl[u].status != ST_RUNNING || e->desc[u].link_n > 1) continue;
                const double busy = (double)(ctl[u].busy_ticks - busy_prev[u]);
                if (busy < min_ticks) continue;
                const double dt = ctl[u].t - t_prev[u];
                pu[u] = dt > 0 ? (ctl[u].t_end - ctl[u].t) * busy / dt : 1e300;
                all.push_back(pu[u]);
            }
            if (!all.empty()) {
                std::nth_element(all.begin(), all.begin() + all.size() / 2, all.end());
                const double med = all[all.size() / 2];
                // the set is ranked afresh each round: a unit above the
                // factor joins, a promoted one (measured in the faster form)
                // stays while it is slower than the median and ranks at twice
                // its projection; the rest go back to the one-lane form
                int live = 0;
                std::vector<std::pair<double, int>> cand;
                for (int u = 0; u < e->n_units; ++u) {
                    if (pu[u] < 0) {
                        live += promoted[u] && ctl[u].status == ST_RUNNING;  // not judged: kept
                        continue;
                    }
                    const bool was = promoted[u] != 0;
                    promoted[u] = 0;
                    if (e->alt_quad[u] < 0) continue;
                    if (was ? pu[u] > med : pu[u] > promote_factor * med) cand.push_back({pu[u] * (was ? 2.0 : 1.0), u});
                }
                std::sort(cand.rbegin(), cand.rend());
                // one CU per promoted unit; CUs left over split the ones
                // with the most remaining time (ts_plan)
                for (auto &cu : cand) {
                    if (live >= promote_units) break;
                    promoted[cu.second] = 1;
                    ++live;
                }
            }
        }
        for (int u = 0; u < e->n_units; ++u) {
            t_prev[u] = ctl[u].t;
            busy_prev[u] = ctl[u].busy_ticks;
        }
        el_prev = el;
        if (hybrid && !e->tail && n_run_units <= n_cu) {
            e->tail = true;
            for (int u = 0; u < e->n_units; ++u)
                if (e->alt_quad[u] >= 0) e->cur_cfg[u] = kNumCfg + tail_q(e->desc[u].n);
        }
        // remaining attempts of each unit from its progress so far in this
        // run (a unit that has not advanced yet counts as the one with the
        // most progress to make)
        auto remaining = [&]() {
            rem.assign(e->n_units, 0.0);
            double worst = 1.0;
            for (int u = 0; u < e->n_units; ++u) {
                const double span = ctl[u].t_end - t_run0[u], done_t = ctl[u].t - t_run0[u];
                const double att = (double)(ctl[u].n_attempts - att_run0[u]);
                if (ctl[u].status == ST_RUNNING && span > 0 && done_t > 0 && att > 0) {
                    rem[u] = att * std::max(0.0, span - done_t) / done_t;
                    worst = std::max(worst, rem[u]);
                } else {
                    rem[u] = -1.0;
                }
            }
            for (auto &r : rem)
                if (r < 0) r = worst;
        };
        // at the tail every round moves the running units to the tail set,
        // split by the rows rule over the CUs they leave idle (re-planned as
        // units finish); before it, the promoted units alone, within their
        // quarter of the CUs (two launches on lanes 0 and 1, beside the
        // one-lane launch on lane 2)
        const bool tail_rounds = splitting && e->tail;
        bool any_promoted = false;
        if (promoting && !e->tail)
            for (int u = 0; u < e->n_units && !any_promoted; ++u)
                any_promoted = promoted[u] && ctl[u].status == ST_RUNNING;
        if (tail_rounds) {
            remaining();
            int rc = ts_plan(e, ctl, rem, n_cu, split);
            if (rc) return rc;
        } else if (any_promoted) {
            remaining();
            int rc = ts_plan(e, ctl, rem, promote_cus, split, &promoted);
            if (rc) return rc;
        }
        int goff[kNumAllCfg + 1];
        ord.clear();
        for (int c = 0; c < kNumAllCfg; ++c) {
            goff[c] = (int)ord.size();
            for (int u : e->lpt)
                if (e->cur_cfg[u] == c && ctl[u].status == ST_RUNNING && !split[u]) ord.push_back(u);
        }
        goff[kNumAllCfg] = (int)ord.size();
        if (!ord.empty())
            HIP_TRY(hipMemcpy(e->d_order.p, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice));
        const int n_launch = n_run - (n_run_units - (int)ord.size()) + e->ts.n;  // split parents run as children
        HIP_TRY(hipMemcpy(e->d_running.p, &n_launch, sizeof(int), hipMemcpyHostToDevice));
        GR15Args a = ens_args(e);
        if (round == 0 && probe_attempts > 0) a.max_attempts = std::min(a.max_attempts, probe_attempts);
        a.budget_ticks = (long long)(group_seconds * 1e3 * (double)e->clock_khz);
        a.running = e->d_running.p;
        // one-lane units yield to the quad form at the tail; at the tail a
        // round ends once a quarter of its units have finished (their CUs go
        // to the others' children in the next round's plan)
        a.tail_units = !hybrid ? 0 : !e->tail ? n_cu : tail_rounds ? 3 * n_launch / 4 : 0;
        const bool rounds = (promoting && !e->tail) || tail_rounds;
        if (max_seconds > 0 || rounds) {
            long long now = 0;
            HIP_TRY(read_device_clock(&now, e->d_clock.p, e->streams[0]));
            double left = max_seconds > 0 ? max_seconds - el : 1e30;
            if (tail_rounds) {
                left = std::min(left, kTailRoundSeconds);
            } else if (rounds) {
                left = std::min(left, round_s);
                round_s = std::min(1.0, std::max(round0, el + round_s));
            }
            a.deadline_tick = now + (long long)(left * 1e3 * (double)e->clock_khz);
        }
        if (e->trace_sched) {
            int np = 0;
            for (int u = 0; u < e->n_units; ++u) np += promoted[u] && ctl[u].status == ST_RUNNING;
            fprintf(stderr, "[abie sched] round %d t %.3f s: %d units running, %d promoted, %d launched in "
                    "the arrays, %d in the tail set (%lld bodies), tail %d, %d distinct lanes\n", round, el,
                    n_run_units, np, (int)ord.size(), e->ts.n, e->ts.total, (int)e->tail, e->lanes_distinct);
            if (const char *ts = getenv("ABIE_HIP_TRACE_SYS")) {
                const int sy = atoi(ts);
                if (sy >= 0 && sy < e->n_sys) {
                    const int u = e->sys_unit0[sy];
                    int kids = 0, kid_n = 0;
                    for (int c = 0; c < e->ts.n; ++c)
                        if (e->ts.parent[c] == u) {
                            ++kids;
                            kid_n = std::max(kid_n, e->ts.desc[c].n);
                        }
                    fprintf(stderr, "[abie sched]   system %d unit %d: t %.4f, %lld evals, %d bodies, cfg %d, %d "
                            "children of up to %d bodies\n", sy, u, ctl[u].t, ctl[u].n_evals, e->desc[u].n,
                            e->cur_cfg[u], kids, kid_n);
                }
            }
        }
        // the tail set first: promoted stragglers take their CUs while the
        // GPU is idle (a quad group of up to 512 threads needs a CU's whole
        // register file, which a CU running one-lane groups never frees while
        // more of them wait), and only then do the one-lane units fill the rest
        if (e->ts.n > 0) {
            std::vector<int> all(e->ts.n);
            std::iota(all.begin(), all.end(), 0);
            int rc = ts_launch(e, tail_args(e, a), all);
            if (rc) return rc;
            if (ord.size() > 0) std::this_thread::sleep_for(std::chrono::microseconds(300));
        }
        // the groups' launches over the lanes, heaviest first (work per
        // evaluation, sum of n * N_m), each to the least loaded lane; the
        // tail set's launches hold lanes 0 and 1
        {
            std::vector<std::pair<double, int>> grp;
            for (int c = 0; c < kNumAllCfg; ++c) {
                double w = 0.0;
                for (int i = goff[c]; i < goff[c + 1]; ++i) w += (double)e->desc[ord[i]].n * std::max(1, e->desc[ord[i]].nm);
                if (goff[c + 1] > goff[c]) grp.push_back({w, c});
            }
            std::sort(grp.rbegin(), grp.rend());
            double load[abie_ensemble::kLanes] = {0.0, 0.0, 0.0, 0.0};
            if (e->ts.n > 0) load[0] = load[1] = 1e300;
            for (auto &g : grp) {
                const int c = g.second, nb = goff[c + 1] - goff[c];
                int l = 0;
                for (int k = 1; k < abie_ensemble::kLanes; ++k)
                    if (load[k] < load[l]) l = k;
                load[l] += g.first;
                GR15Args ac = a;
                ac.order = e->d_order.p + goff[c];
                if (cfg_is_quad(c)) HIP_TRY(launch_gr15q(c - kNumCfg, e->force_mode, e->ctrl, ac, nb, e->streams[l]));
                else HIP_TRY(launch_gr15(c, e->force_mode, e->ctrl, ac, nb, e->streams[l]));
            }
        }
        for (auto st : e->streams) HIP_TRY(hipStreamSynchronize(st));
        if (e->ts.n > 0) {
            int rc = ts_finish(e, a);
            if (rc) return rc;
        }
        // linked systems one at a time, each a cooperative launch of its slices
        // (their entries of d_order follow the other units' and are never rewritten)
        for (auto &lk : e->links) {
            if (ctl[e->order[lk[1]]].status != ST_RUNNING) continue;
            GR15Args ac = a;
            ac.order = e->d_order.p + lk[1];
            const hipError_t err = launch_gr15(lk[0], e->force_mode, e->ctrl, ac, lk[2], e->streams[0], true);
            if (err != hipSuccess)
                return fail(ABIE_HIP_EHIP, "cooperative launch of %d linked slices failed: %s (the slices of one "
                            "system must be resident together)", lk[2], hipGetErrorString(err));
            HIP_TRY(hipStreamSynchronize(e->streams[0]));
        }
    }
    if (done_out) *done_out = done;
    return 0;
}

extern "C" {

const char *abie_hip_last_error(void) { return g_errbuf; }

int abie_hip_set_device(int device) {
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return fail(ABIE_HIP_ENODEV, "no HIP device");
    if (device < 0 || device >= cnt) return fail(ABIE_HIP_ENODEV, "device %d out of range (%d devices)", device, cnt);
    HIP_TRY(hipSetDevice(device));
    return 0;
}

int abie_ensemble_create(abie_ensemble **out, int n_sys, const int *offsets, const double *pos, const double *vel,
                         const double *m, const double *r, double G, int force_mode, int controller) {
    if (!out || n_sys <= 0 || !offsets || !pos || !vel || !m) return fail(ABIE_HIP_EINVAL, "ensemble_create: bad arguments");
    if (!gr15_combo_supported(force_mode, controller))
        return fail(ABIE_HIP_EINVAL, "ensemble_create: force mode %d with controller %d is not built "
                    "(exact forces run with the plain controller, mixed forces with the massive-only ones)",
                    force_mode, controller);
    auto *e = new abie_ensemble();
    (void)hipGetDevice(&e->dev);
    e->n_sys = n_sys;
    e->G = G;
    e->force_mode = force_mode;
    e->ctrl = controller;
    // the quad form has no close-encounter scan; ABIE_HIP_QUAD=0 turns it off
    const char *ev_env = getenv("ABIE_HIP_EVENTS");
    if (const char *q = getenv("ABIE_HIP_QUAD")) e->quad = std::min(2, std::max(0, atoi(q)));
    if (const char *q = getenv("ABIE_HIP_TAIL_SPLIT")) e->tail_split = atoi(q) != 0;
    if (const char *q = getenv("ABIE_HIP_PROMOTE")) e->promote = atoi(q) != 0;
    if (const char *q = getenv("ABIE_HIP_TRACE_SCHED")) e->trace_sched = atoi(q) != 0;
    if (ev_env && atoi(ev_env) != 0) e->quad = 0;
    int rc = ens_build(e, offsets, pos, vel, m, r, nullptr);
    if (rc == 0) rc = ens_rings(e, 0, 0);
    if (rc != 0) {
        abie_ensemble_destroy(e);
        return rc;
    }
    if (const char *s = getenv("ABIE_HIP_MAX_ATTEMPTS")) e->max_attempts = std::max(1, atoi(s));
    if (const char *s = getenv("ABIE_HIP_EVENTS")) e->events = atoi(s);
    *out = e;
    return 0;
}

void abie_ensemble_destroy(abie_ensemble *e) {
    if (!e) return;
    // nothing may still run on the streams whose buffers are freed and which
    // are destroyed below (a cooperative launch of linked slices included)
    (void)hipSetDevice(e->dev);
    for (auto &s : e->streams)
        if (s) (void)hipStreamSynchronize(s);
    for (auto *b : {&e->d_x, &e->d_y, &e->d_z, &e->d_vx, &e->d_vy, &e->d_vz, &e->d_gm, &e->d_m, &e->d_rad,
                    &e->d_ext, &e->d_state, &e->d_scan, &e->d_ring_ce, &e->d_ring_coll, &e->d_log, &e->d_energy})
        b->release();
    e->d_sys.release(); e->d_order.release(); e->d_orig.release(); e->d_inv.release();
    e->d_ctl.release(); e->d_del.release(); e->d_log_len.release(); e->d_mlp.release(); e->d_mratio.release();
    e->d_link_count.release(); e->d_link_part.release(); e->d_sys_inv.release();
    e->d_running.release(); e->d_clock.release(); e->d_probe.release(); e->d_live.release(); e->d_live_ctl.release();
    e->ts.release();
    for (auto &s : e->streams)
        if (s) (void)hipStreamDestroy(s);
    delete e;
}

int abie_ensemble_run(abie_ensemble *e, double t_end, double max_seconds, int *done) {
    if (!e) return fail(ABIE_HIP_EINVAL, "ensemble_run: null handle");
    int rc = ens_set_t_end(e, t_end, nullptr);
    if (rc) return rc;
    return ens_run(e, max_seconds, done);
}

int abie_ensemble_get_state(abie_ensemble *e, double *pos, double *vel, double *t, int *status) {
    if (!e) return fail(ABIE_HIP_EINVAL, "ensemble_get_state: null handle");
    const long long T = e->total;
    if (pos || vel) {
        std::vector<double> a(T), b(T), c(T);
        auto put = [&](double *dst, DBuf<double> &x, DBuf<double> &y, DBuf<double> &z) -> int {
            HIP_TRY(hipMemcpy(a.data(), x.p, T * sizeof(double), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(b.data(), y.p, T * sizeof(double), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(c.data(), z.p, T * sizeof(double), hipMemcpyDeviceToHost));
            for (long long g = 0; g < T; ++g) {
                const long long i = e->eng2orig[g];
                if (i < 0) continue;
                dst[3 * i] = a[g]; dst[3 * i + 1] = b[g]; dst[3 * i + 2] = c[g];
            }
            return 0;
        };
        if (pos && put(pos, e->d_x, e->d_y, e->d_z)) return ABIE_HIP_EHIP;
        if (vel && put(vel, e->d_vx, e->d_vy, e->d_vz)) return ABIE_HIP_EHIP;
    }
    if (t || status) {
        std::vector<SysCtl> ctl(e->n_units);
        HIP_TRY(hipMemcpy(ctl.data(), e->d_ctl.p, e->n_units * sizeof(SysCtl), hipMemcpyDeviceToHost));
        for (int s = 0; s < e->n_sys; ++s) {
            const SysCtl &c = ctl[e->sys_unit0[s]];
            if (t) t[s] = c.t;
            int st = c.status == ST_DONE ? 0 : c.status;
            for (int u = 0; u < e->n_units; ++u)  // any slice's event or error wins
                if (e->unit_sys[u] == s && ctl[u].status != ST_DONE && ctl[u].status != ST_RUNNING) st = ctl[u].status;
            if (status) status[s] = st;
        }
    }
    return 0;
}

int abie_ensemble_get_counters(abie_ensemble *e, long long *cnt) {
    if (!e || !cnt) return fail(ABIE_HIP_EINVAL, "ensemble_get_counters: bad arguments");
    std::vector<SysCtl> ctl(e->n_units);
    HIP_TRY(hipMemcpy(ctl.data(), e->d_ctl.p, e->n_units * sizeof(SysCtl), hipMemcpyDeviceToHost));
    for (int s = 0; s < e->n_sys; ++s) {
        const SysCtl &u = ctl[e->sys_unit0[s]];
        long long *c = cnt + 6 * s;
        c[0] = u.n_evals; c[1] = u.n_attempts; c[2] = u.n_accepted;
        c[3] = u.n_pc; c[4] = u.n_prop_used; c[5] = u.n_prop_rej;
    }
    return 0;
}

int abie_ensemble_get_finish_times(abie_ensemble *e, double *seconds) {
    if (!e || !seconds) return fail(ABIE_HIP_EINVAL, "ensemble_get_finish_times: bad arguments");
    std::vector<SysCtl> ctl(e->n_units);
    HIP_TRY(hipMemcpy(ctl.data(), e->d_ctl.p, e->n_units * sizeof(SysCtl), hipMemcpyDeviceToHost));
    long long t0 = 0;
    for (auto &c : ctl)
        if (c.start_tick > 0 && (t0 == 0 || c.start_tick < t0)) t0 = c.start_tick;
    const double khz = e->clock_khz > 0 ? (double)e->clock_khz : 1.0;
    for (int s = 0; s < e->n_sys; ++s) seconds[s] = 0.0;
    for (int u = 0; u < e->n_units; ++u) {
        const SysCtl &c = ctl[u];
        // a unit that stopped before this run (its launches return at once) counts as 0
        const double f = (c.done_tick > 0 && t0 > 0) ? (double)(c.done_tick - t0) / (khz * 1e3)
                         : (c.status == ST_RUNNING ? -1.0 : 0.0);
        double &d = seconds[e->unit_sys[u]];
        d = (f < 0 || d < 0) ? -1.0 : std::max(d, f);
    }
    return 0;
}

int abie_ensemble_energy(abie_ensemble *e, double *energy) {
    if (!e || !energy) return fail(ABIE_HIP_EINVAL, "ensemble_energy: bad arguments");
    HIP_TRY(launch_energy(e->d_sys.p, e->n_units, e->d_x.p, e->d_y.p, e->d_z.p, e->d_vx.p, e->d_vy.p, e->d_vz.p,
                          e->d_m.p, e->G, e->d_energy.p, 0));
    std::vector<double> eu(e->n_units);
    HIP_TRY(hipMemcpy(eu.data(), e->d_energy.p, e->n_units * sizeof(double), hipMemcpyDeviceToHost));
    for (int s = 0; s < e->n_sys; ++s) energy[s] = eu[e->sys_unit0[s]];
    return 0;
}

int abie_ensemble_set_steplog(abie_ensemble *e, long long capacity) {
    if (!e || capacity < 0) return fail(ABIE_HIP_EINVAL, "ensemble_set_steplog: bad arguments");
    e->log_cap = capacity;
    if (capacity > 0) HIP_TRY(e->d_log.reserve((size_t)capacity * ABIE_STEPLOG_FIELDS));
    HIP_TRY(hipMemset(e->d_log_len.p, 0, sizeof(unsigned long long)));
    return 0;
}

int abie_ensemble_get_steplog(abie_ensemble *e, double *records, long long *n_records) {
    if (!e || !n_records) return fail(ABIE_HIP_EINVAL, "ensemble_get_steplog: bad arguments");
    unsigned long long len = 0;
    HIP_TRY(hipMemcpy(&len, e->d_log_len.p, sizeof len, hipMemcpyDeviceToHost));
    const long long n = std::min<long long>((long long)len, e->log_cap);
    if (records && n > 0)
        HIP_TRY(hipMemcpy(records, e->d_log.p, (size_t)n * ABIE_STEPLOG_FIELDS * sizeof(double), hipMemcpyDeviceToHost));
    *n_records = n;
    return 0;
}

int abie_ensemble_set_proposer(abie_ensemble *e, const float *w, int n_weights) {
    if (!e) return fail(ABIE_HIP_EINVAL, "ensemble_set_proposer: null handle");
    if (!w) {
        e->has_mlp = false;
        return 0;
    }
    if (n_weights != ABIE_MLP_NPARAM) return fail(ABIE_HIP_EINVAL, "proposer needs %d weights, got %d", ABIE_MLP_NPARAM, n_weights);
    HIP_TRY(e->d_mlp.reserve(ABIE_MLP_NPARAM));
    HIP_TRY(hipMemcpy(e->d_mlp.p, w, ABIE_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    e->has_mlp = true;
    return 0;
}

}  // extern "C"

// ===================================================================== reference ABI
namespace {

struct GlobalSim {
    bool inited = false;
    int dev = -1;
    double G = 0, C = 0;
    size_t nmax = 0, max_ce = 0, max_coll = 0;
    size_t n = 0;
    std::vector<double> pos, vel, m, r, ext;
    bool ext_on = false;
    double t = 0.0, ce_dist = 0.0;
    size_t n_ce = 0, n_coll = 0;
    std::vector<double> buf_ce, buf_coll;
    int force_mode = ABIE_FORCE_FP64;
    bool mode_set = false;
    // AoS force path buffers
    DBuf<double> d_pos, d_m, d_x, d_y, d_z, d_gm, d_ax, d_ay, d_az, d_acc, d_ext, d_part;
    DBuf<int> d_perm;
    hipStream_t stream = nullptr;
    int n_cu = 256;
    // integrator_gr's engine, kept between calls while nothing it was built
    // from (bodies, masses, radii, extra forces) changes: set_state,
    // set_additional_forces and finalize_code drop it
    abie_ensemble *gr_cache = nullptr;
};

GlobalSim g;

void gr_cache_drop() {
    if (g.gr_cache) {
        abie_ensemble_destroy(g.gr_cache);
        g.gr_cache = nullptr;
    }
}

int check_device() {
    if (!g.inited) return fail(ABIE_HIP_ESTATE, "initialize_code has not been called");
    HIP_TRY(hipSetDevice(g.dev));
    return 0;
}

// Host AoS accelerations through the device (common.c:82-114 semantics).
int aos_accel(const double *pos, size_t N, double G, const double *m, double *acc, bool with_ext) {
    if (N == 0) return 0;
    int rc = check_device();
    if (rc) return rc;
    std::vector<int> perm;
    const int nm = engine_order(m, (int)N, perm);
    bool any_deleted = false;
    for (size_t i = 0; i < N; ++i) any_deleted |= (m[i] < 0.0);
    HIP_TRY(g.d_pos.reserve(3 * N)); HIP_TRY(g.d_m.reserve(N)); HIP_TRY(g.d_perm.reserve(N));
    for (auto *b : {&g.d_x, &g.d_y, &g.d_z, &g.d_gm, &g.d_ax, &g.d_ay, &g.d_az}) HIP_TRY(b->reserve(N));
    HIP_TRY(g.d_acc.reserve(3 * N));
    hipStream_t st = g.stream;
    HIP_TRY(hipMemcpyAsync(g.d_pos.p, pos, 3 * N * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(g.d_m.p, m, N * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(g.d_perm.p, perm.data(), N * sizeof(int), hipMemcpyHostToDevice, st));
    if (any_deleted) HIP_TRY(hipMemcpyAsync(g.d_acc.p, acc, 3 * N * sizeof(double), hipMemcpyHostToDevice, st));
    const double *dext = nullptr;
    if (with_ext && g.ext_on && g.ext.size() >= 3 * N) {
        HIP_TRY(g.d_ext.reserve(3 * N));
        HIP_TRY(hipMemcpyAsync(g.d_ext.p, g.ext.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, st));
        dext = g.d_ext.p;
    }
    HIP_TRY(launch_pack(g.d_pos.p, g.d_m.p, g.d_perm.p, (int)N, nm, G, g.d_x.p, g.d_y.p, g.d_z.p, g.d_gm.p, st));
    AccelArgs a{g.d_x.p, g.d_y.p, g.d_z.p, g.d_gm.p, g.d_x.p, g.d_y.p, g.d_z.p, g.d_ax.p, g.d_ay.p, g.d_az.p,
                nm, (int)N, nm, g.force_mode};
    AccelPlan pl = plan_accel(nm, (int)N, g.force_mode, g.n_cu);
    if (pl.parts > 1) HIP_TRY(g.d_part.reserve((size_t)pl.parts * 3 * N));
    HIP_TRY(launch_accel(a, pl, g.d_part.p, st));
    HIP_TRY(launch_unpack(g.d_perm.p, g.d_m.p, (int)N, g.d_ax.p, g.d_ay.p, g.d_az.p, dext, g.d_acc.p, st));
    HIP_TRY(hipMemcpyAsync(acc, g.d_acc.p, 3 * N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

// GR15 of the global simulation through a one-system ensemble.
// GR15 from (pos, vel) to t_end as one single-system ensemble.  cache: the
// arrays are the global simulation's own and the engine is kept for the next
// call; its device state then already holds these positions and velocities
// (it wrote them back at the end of the previous call), so a later call only
// resets the step controller -- every call restarts GR15 from a fresh initial
// step, as the reference's does, and the trajectory is the same bits.
int run_gr15(double *pos, double *vel, const double *m, const double *r, size_t N, double G, double t,
             double t_end, double *t_out, bool cache) {
    if (N == 0) return fail(ABIE_HIP_EINVAL, "integrator_gr: no bodies (the reference never terminates here)");
    if (g.C > 0.0) return fail(ABIE_HIP_EUNSUP, "post-Newtonian forces (C > 0) are not part of the device path");
    // The reference's controller takes its norms over every body, test
    // particles included, and the seventh divided difference of its error
    // estimate amplifies single-precision noise in their forces by ~1e4: the
    // mixed mode is built for the massive-only controllers of the ensemble
    // API only.  calculate_accelerations still honours it.
    if (g.force_mode == ABIE_FORCE_MIXED)
        return fail(ABIE_HIP_EUNSUP, "integrator_gr: mixed forces run only under the massive-only controllers "
                    "(abie_ensemble_*); the reference's controller takes fp64 or exact forces");
    int rc = check_device();
    if (rc) return rc;
    abie_ensemble *e = cache ? g.gr_cache : nullptr;
    if (e == nullptr) {
        e = new abie_ensemble();
        e->dev = g.dev;
        e->n_sys = 1;
        e->G = G;
        e->ctrl = ABIE_CTRL_PLAIN;
        // a lone system of more than 256 bodies runs as linked slices of 128
        // over several CUs: the configuration-0 system (1,005 bodies) ran
        // 10 yr at 116.6 system-yr/s against 49.7 in one workgroup
        e->link_above = 256;
        e->link_cap = 128;
        if (const char *sl = getenv("ABIE_HIP_SLICE")) e->link_cap = std::max(64, atoi(sl));
        const int offs[2] = {0, (int)N};
        const double *ext = (g.ext_on && g.ext.size() >= 3 * N) ? g.ext.data() : nullptr;
        rc = ens_build(e, offs, pos, vel, m, r, ext);
        if (rc == 0) rc = ens_rings(e, (int)g.max_ce, (int)g.max_coll);
        if (rc != 0) {
            abie_ensemble_destroy(e);
            return rc;
        }
        if (cache) g.gr_cache = e;
    }
    e->force_mode = g.force_mode;
    e->ce_dist = g.ce_dist;
    e->max_attempts = 1 << 30;
    if (const char *s = getenv("ABIE_HIP_MAX_ATTEMPTS")) e->max_attempts = std::max(1, atoi(s));
    // carry the event counters and rings of the global simulation
    if (rc == 0) {
        SysCtl c;
        memset(&c, 0, sizeof c);
        c.status = ST_RUNNING;
        c.t = t;
        c.t_end = t_end;
        c.n_ce = (long long)g.n_ce;
        c.n_coll = (long long)g.n_coll;
        // every unit (the slices of a linked system) restarts, grid-barrier
        // epochs and counters from zero
        std::vector<SysCtl> cs(e->n_units, c);
        if (hipMemcpy(e->d_ctl.p, cs.data(), e->n_units * sizeof(SysCtl), hipMemcpyHostToDevice) != hipSuccess)
            rc = ABIE_HIP_EHIP;
        if (rc == 0 && hipMemset(e->d_link_count.p, 0, e->n_units * sizeof(unsigned long long)) != hipSuccess)
            rc = ABIE_HIP_EHIP;
        if (rc == 0 && g.max_ce > 0)
            if (hipMemcpy(e->d_ring_ce.p, g.buf_ce.data(), 4 * g.max_ce * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                rc = ABIE_HIP_EHIP;
        if (rc == 0 && g.max_coll > 0)
            if (hipMemcpy(e->d_ring_coll.p, g.buf_coll.data(), 4 * g.max_coll * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                rc = ABIE_HIP_EHIP;
    }
    int done = 0;
    if (rc == 0) rc = ens_run(e, 0.0, &done);
    int ret = 0;
    if (rc == 0) {
        SysCtl c;
        if (hipMemcpy(&c, e->d_ctl.p, sizeof c, hipMemcpyDeviceToHost) != hipSuccess) rc = ABIE_HIP_EHIP;
        if (rc == 0) rc = abie_ensemble_get_state(e, pos, vel, nullptr, nullptr);
        if (rc == 0) {
            *t_out = c.t;
            g.n_ce = (size_t)c.n_ce;
            g.n_coll = (size_t)c.n_coll;
            if (g.max_ce > 0)
                (void)hipMemcpy(g.buf_ce.data(), e->d_ring_ce.p, 4 * g.max_ce * sizeof(double), hipMemcpyDeviceToHost);
            if (g.max_coll > 0)
                (void)hipMemcpy(g.buf_coll.data(), e->d_ring_coll.p, 4 * g.max_coll * sizeof(double), hipMemcpyDeviceToHost);
            ret = c.status == ST_CE ? ABIE_EXIT_MAX_N_CE_EXCEEDED
                  : c.status == ST_COLL ? ABIE_EXIT_MAX_N_COLLISIONS_EXCEEDED : ABIE_EXIT_NORMAL;
            if (c.status < 0) rc = fail(ABIE_HIP_EHIP, "integrator_gr: device engine reported an error");
        }
    }
    if (!cache) abie_ensemble_destroy(e);
    else if (rc != 0) gr_cache_drop();  // keep no engine in an unknown state
    return rc ? rc : ret;
}

}  // namespace

extern "C" {

int initialize_code(double G, double C, size_t N_MAX, size_t MAX_N_CE, size_t MAX_N_COLLISIONS, int deviceID) {
    if (g.inited) return 0;  // once per process (common.c:225-227)
    printf("Initializing the code..., deviceID = %d", deviceID);
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) {
        printf("\n");
        return fail(ABIE_HIP_ENODEV, "no HIP device found; libabie_hip.so has no CPU path");
    }
    if (deviceID < 0 || deviceID >= cnt) {
        printf("\n");
        return fail(ABIE_HIP_ENODEV, "deviceID %d invalid (%d devices); use the reference libabie.so for the CPU path",
                    deviceID, cnt);
    }
    HIP_TRY(hipSetDevice(deviceID));
    g.dev = deviceID;
    g.n_cu = device_cu_count(deviceID);
    if (!g.stream) HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    g.G = G;
    g.C = C;
    g.nmax = N_MAX;
    g.max_ce = MAX_N_CE;
    g.max_coll = MAX_N_COLLISIONS;
    g.pos.assign(3 * N_MAX, 0.0);
    g.vel.assign(3 * N_MAX, 0.0);
    g.m.assign(N_MAX, 0.0);
    g.r.assign(N_MAX, 0.0);
    if (g.ext.size() < 3 * N_MAX) g.ext.assign(3 * N_MAX, 0.0);
    g.buf_ce.assign(4 * MAX_N_CE, 0.0);
    g.buf_coll.assign(4 * MAX_N_COLLISIONS, 0.0);
    // a fresh simulation: counters, clock and extra forces start empty (the
    // reference initialises once per process, so it never re-enters here)
    g.n_ce = 0;
    g.n_coll = 0;
    g.t = 0.0;
    g.n = 0;
    g.ext_on = false;
    std::fill(g.ext.begin(), g.ext.end(), 0.0);
    if (!g.mode_set) g.force_mode = force_mode_from_env();
    g.inited = true;
    printf("Initialized.\n");
    return 0;
}

int finalize_code(void) {
    gr_cache_drop();
    if (g.stream) {
        (void)hipSetDevice(g.dev);
        (void)hipStreamSynchronize(g.stream);
    }
    for (auto *b : {&g.d_pos, &g.d_m, &g.d_x, &g.d_y, &g.d_z, &g.d_gm, &g.d_ax, &g.d_ay, &g.d_az, &g.d_acc, &g.d_ext, &g.d_part})
        b->release();
    g.d_perm.release();
    if (g.stream) (void)hipStreamDestroy(g.stream);
    g.stream = nullptr;
    g.inited = false;
    g.n = 0;
    g.ext_on = false;
    g.ext.clear();
    return 0;
}

void set_state(double *pos, double *vel, double *m, double *r, size_t N, double G, double C) {
    gr_cache_drop();
    if (N > g.nmax) {
        g.nmax = N;
        g.pos.resize(3 * N); g.vel.resize(3 * N); g.m.resize(N); g.r.resize(N);
        if (g.ext.size() < 3 * N) g.ext.resize(3 * N, 0.0);
    }
    std::copy(pos, pos + 3 * N, g.pos.begin());
    std::copy(vel, vel + 3 * N, g.vel.begin());
    std::copy(m, m + N, g.m.begin());
    std::copy(r, r + N, g.r.begin());
    g.n = N;
    g.G = G;
    g.C = C;
}

int get_state(double *pos, double *vel, double *m, double *r) {
    std::copy(g.pos.begin(), g.pos.begin() + 3 * g.n, pos);
    std::copy(g.vel.begin(), g.vel.begin() + 3 * g.n, vel);
    std::copy(g.m.begin(), g.m.begin() + g.n, m);
    std::copy(g.r.begin(), g.r.begin() + g.n, r);
    return (int)g.n;
}

double get_model_time(void) { return g.t; }
void set_close_encounter_distance(double d) { g.ce_dist = d; }
double get_close_encounter_distance(void) { return g.ce_dist; }
void get_close_encounter_buffer(double *buf) { std::copy(g.buf_ce.begin(), g.buf_ce.end(), buf); }
void get_collision_buffer(double *buf) { std::copy(g.buf_coll.begin(), g.buf_coll.end(), buf); }
void reset_close_encounter_buffer(void) {
    g.n_ce = 0;
    std::fill(g.buf_ce.begin(), g.buf_ce.end(), 0.0);
}
void reset_collision_buffer(void) {
    g.n_coll = 0;
    std::fill(g.buf_coll.begin(), g.buf_coll.end(), 0.0);
}

size_t set_additional_forces(size_t N, double *ext_acc) {
    gr_cache_drop();
    g.ext_on = true;
    if (g.ext.size() < 3 * N) g.ext.resize(3 * N, 0.0);
    std::copy(ext_acc, ext_acc + 3 * N, g.ext.begin());
    return 0;
}

double calculate_energy(void) {
    if (check_device() != 0 || g.n == 0) return 0.0;
    const size_t N = g.n;
    DBuf<SysDesc> sd;
    DBuf<double> buf, out;
    double e = 0.0;
    if (sd.reserve(1) != hipSuccess || buf.reserve(7 * N) != hipSuccess || out.reserve(1) != hipSuccess) {
        fail(ABIE_HIP_ENOMEM, "calculate_energy: allocation failed");
        return 0.0;
    }
    std::vector<double> h(7 * N);
    for (size_t i = 0; i < N; ++i) {
        for (int c = 0; c < 3; ++c) {
            h[c * N + i] = g.pos[3 * i + c];
            h[(3 + c) * N + i] = g.vel[3 * i + c];
        }
        h[6 * N + i] = g.m[i];
    }
    SysDesc d{0, (int)N, 0, 0, 0, 0, 1, 0};
    bool ok = hipMemcpy(sd.p, &d, sizeof d, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(buf.p, h.data(), 7 * N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              launch_energy(sd.p, 1, buf.p, buf.p + N, buf.p + 2 * N, buf.p + 3 * N, buf.p + 4 * N, buf.p + 5 * N,
                            buf.p + 6 * N, g.G, out.p, 0) == hipSuccess &&
              hipMemcpy(&e, out.p, sizeof e, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) fail(ABIE_HIP_EHIP, "calculate_energy: device call failed");
    sd.release(); buf.release(); out.release();
    return e;
}

int integrator_gr(double t, double t_end, double dt) {
    (void)dt;
    if (!g.inited) return fail(ABIE_HIP_ESTATE, "initialize_code has not been called");
    double t_out = t;
    const int ret = run_gr15(g.pos.data(), g.vel.data(), g.m.data(), g.r.data(), g.n, g.G, t, t_end, &t_out, true);
    if (ret >= 0) g.t = t_out;
    return ret;
}

size_t integrator_gauss_radau15(double *pos, double *vel, double *m, double *r, size_t N, double G, double t,
                                double t_end, double dt) {
    (void)dt;
    double t_out = t;
    const int ret = run_gr15(pos, vel, m, r, N, G, t, t_end, &t_out, false);
    if (ret >= 0) g.t = t_out;
    return (size_t)(long long)ret;
}

size_t calculate_accelerations(const double *pos, const double *vel, size_t N, double G, const double *m,
                               const double *r, double *acc) {
    (void)vel;
    (void)r;
    if (g.C > 0.0) return (size_t)(long long)fail(ABIE_HIP_EUNSUP, "post-Newtonian forces (C > 0) are not part of the device path");
    return (size_t)(long long)aos_accel(pos, N, G, m, acc, true);
}

size_t ode_n_body_second_order(const double *pos, size_t N, double G, const double *m, const double *r, double *acc) {
    (void)r;
    return (size_t)(long long)aos_accel(pos, N, G, m, acc, false);
}

size_t ode_n_body_second_order_gpu(const double *pos, size_t N, double G, const double *m, const double *r,
                                   double *acc) {
    return ode_n_body_second_order(pos, N, G, m, r, acc);
}

size_t check_collisions_close_encounters(const double *pos, const double *radii, size_t N, double t) {
    if (check_device() != 0) return (size_t)(long long)ABIE_HIP_ESTATE;
    if (N == 0) return 0;
    DBuf<double> dpos, drad, dce, dco;
    DBuf<long long> dcnt;
    DBuf<int> dret, dflag;
    long long cnt[2] = {(long long)g.n_ce, (long long)g.n_coll};
    int ret = 0;
    bool ok = dpos.reserve(3 * N) == hipSuccess && drad.reserve(N) == hipSuccess && dcnt.reserve(2) == hipSuccess &&
              dret.reserve(2) == hipSuccess && dce.reserve(4 * std::max<size_t>(g.max_ce, 1)) == hipSuccess &&
              dco.reserve(4 * std::max<size_t>(g.max_coll, 1)) == hipSuccess && dflag.reserve(N) == hipSuccess;
    const int flags0[2] = {0, 0};
    ok = ok && hipMemcpy(dret.p, flags0, sizeof flags0, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dpos.p, pos, 3 * N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(drad.p, radii, N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dcnt.p, cnt, sizeof cnt, hipMemcpyHostToDevice) == hipSuccess &&
         (g.max_ce == 0 || hipMemcpy(dce.p, g.buf_ce.data(), 4 * g.max_ce * sizeof(double), hipMemcpyHostToDevice) == hipSuccess) &&
         (g.max_coll == 0 || hipMemcpy(dco.p, g.buf_coll.data(), 4 * g.max_coll * sizeof(double), hipMemcpyHostToDevice) == hipSuccess);
    ok = ok && hipMemsetAsync(dflag.p, 0, N * sizeof(int), g.stream) == hipSuccess &&
         launch_events(dpos.p, drad.p, (int)N, t, g.ce_dist, dcnt.p, (int)g.max_ce, (int)g.max_coll, dce.p, dco.p,
                       dret.p, dflag.p, g.stream) == hipSuccess &&
         hipStreamSynchronize(g.stream) == hipSuccess &&
         hipMemcpy(cnt, dcnt.p, sizeof cnt, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(&ret, dret.p + 1, sizeof ret, hipMemcpyDeviceToHost) == hipSuccess &&
         (g.max_ce == 0 || hipMemcpy(g.buf_ce.data(), dce.p, 4 * g.max_ce * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess) &&
         (g.max_coll == 0 || hipMemcpy(g.buf_coll.data(), dco.p, 4 * g.max_coll * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess);
    dpos.release(); drad.release(); dce.release(); dco.release(); dcnt.release(); dret.release(); dflag.release();
    if (!ok) return (size_t)(long long)fail(ABIE_HIP_EHIP, "check_collisions_close_encounters: device call failed");
    g.n_ce = (size_t)cnt[0];
    g.n_coll = (size_t)cnt[1];
    return (size_t)ret;
}

int abie_gr15_num_configs(void) { return kNumCfg; }

int abie_gr15_config(int cfg, int *threads, int *bodies_per_thread) {
    if (cfg < 0 || cfg >= kNumCfg) return fail(ABIE_HIP_EINVAL, "no configuration %d", cfg);
    if (threads) *threads = kCfgs[cfg].nt;
    if (bodies_per_thread) *bodies_per_thread = kCfgs[cfg].bpt;
    return 0;
}

int abie_gr15_set_tuning(int n_rules, const int *max_bodies, const int *cfg) {
    if (set_tuning(n_rules, max_bodies, cfg) != 0) return fail(ABIE_HIP_EINVAL, "bad tuning rules");
    return 0;
}

double abie_hip_step_root7(double x) { return abie::root7(x); }

int abie_hip_set_force_mode(int mode) {
    if (mode < 0 || mode > 2) return fail(ABIE_HIP_EINVAL, "bad force mode %d", mode);
    g.force_mode = mode;
    g.mode_set = true;
    return 0;
}

void gpu_init(int N, int deviceID) {
    if (!g.inited) initialize_code(g.G, g.C, (size_t)std::max(N, 0), 1, 1, deviceID);
}

void gpu_finalize(void) { finalize_code(); }

int integrator_rk(double, double, double) { return fail(ABIE_HIP_EUNSUP, "RungeKutta is outside the device path"); }
int integrator_wh(double, double, double) { return fail(ABIE_HIP_EUNSUP, "WisdomHolman is outside the device path"); }
void integrator_runge_kutta(double *, double *, double *, double *, size_t, double, double, double, double) {
    fail(ABIE_HIP_EUNSUP, "RungeKutta is outside the device path");
}
void integrator_wisdom_holman(double *, double *, double *, double *, size_t, double, double, double, double) {
    fail(ABIE_HIP_EUNSUP, "WisdomHolman is outside the device path");
}
size_t ode_n_body_first_order(double *, size_t, double, const double *, double *) {
    return (size_t)(long long)fail(ABIE_HIP_EUNSUP, "ode_n_body_first_order is outside the device path");
}

// ---------------------------------------------------------------- SoA force
int abie_hip_accel_soa(const double *sx, const double *sy, const double *sz, const double *gm, int n_src,
                       const double *px, const double *py, const double *pz, int n_sink, int n_self, double *ax,
                       double *ay, double *az, int mode, void *stream) {
    if (n_src < 0 || n_sink < 0 || n_self < 0 || n_self > n_src || n_self > n_sink)
        return fail(ABIE_HIP_EINVAL, "accel_soa: bad sizes (n_src %d, n_sink %d, n_self %d)", n_src, n_sink, n_self);
    if (mode < 0 || mode > 3) return fail(ABIE_HIP_EINVAL, "accel_soa: bad mode %d", mode);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    AccelArgs a{sx, sy, sz, gm, px, py, pz, ax, ay, az, n_src, n_sink, n_self, mode};
    AccelPlan pl = plan_accel(n_src, n_sink, mode, device_cu_count(dev));
    const hipStream_t st = (hipStream_t)stream;
    const size_t need = plan_scratch_doubles(pl, n_sink);
    if (need == 0) {
        HIP_TRY(launch_accel(a, pl, nullptr, st));
        return 0;
    }
    // The source-split partials live in a stream-ordered allocation: calls on
    // different streams get their own buffers, and the pool reuses the
    // memory once the stream has passed the free.
    double *part = nullptr;
    HIP_TRY(hipMallocAsync((void **)&part, need * sizeof(double), st));
    const hipError_t le = launch_accel(a, pl, part, st);
    const hipError_t fe = hipFreeAsync(part, st);
    if (le != hipSuccess) return fail(ABIE_HIP_EHIP, "accel_soa: launch failed: %s", hipGetErrorString(le));
    HIP_TRY(fe);
    return 0;
}

int abie_hip_time_accel_soa(const double *sx, const double *sy, const double *sz, const double *gm, int n_src,
                            const double *px, const double *py, const double *pz, int n_sink, int n_self,
                            double *ax, double *ay, double *az, int mode, void *stream, int reps,
                            double *ms_per_launch, double *ms_each, double *ms_each_median, double *shader_ghz) {
    if (reps <= 0 || !ms_per_launch) return fail(ABIE_HIP_EINVAL, "time_accel_soa: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    std::vector<hipEvent_t> ev(ms_each ? 2 * reps : 2);
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
    auto destroy = [&]() {
        for (auto &e : ev) (void)hipEventDestroy(e);
    };
    int rc = 0;
    HIP_TRY(hipEventRecord(ev[0], st));
    for (int r = 0; r < reps && !rc; ++r)
        rc = abie_hip_accel_soa(sx, sy, sz, gm, n_src, px, py, pz, n_sink, n_self, ax, ay, az, mode, stream);
    if (rc) {
        destroy();
        return rc;
    }
    HIP_TRY(hipEventRecord(ev[1], st));
    HIP_TRY(hipEventSynchronize(ev[1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
    *ms_per_launch = ms / reps;
    if (ms_each) {
        for (int r = 0; r < reps && !rc; ++r) {
            HIP_TRY(hipEventRecord(ev[2 * r], st));
            rc = abie_hip_accel_soa(sx, sy, sz, gm, n_src, px, py, pz, n_sink, n_self, ax, ay, az, mode, stream);
            HIP_TRY(hipEventRecord(ev[2 * r + 1], st));
        }
        if (rc) {
            destroy();
            return rc;
        }
        HIP_TRY(hipEventSynchronize(ev[2 * reps - 1]));
        std::vector<double> each(reps);
        double tot = 0.0;
        for (int r = 0; r < reps; ++r) {
            float m1 = 0;
            HIP_TRY(hipEventElapsedTime(&m1, ev[2 * r], ev[2 * r + 1]));
            each[r] = m1;
            tot += m1;
        }
        *ms_each = tot / reps;
        if (ms_each_median) {
            std::sort(each.begin(), each.end());
            *ms_each_median = reps % 2 ? each[reps / 2] : 0.5 * (each[reps / 2 - 1] + each[reps / 2]);
        }
    }
    destroy();
    if (shader_ghz) {
        // back-to-back launches of the clocked instance of the same kernel;
        // per launch the median over waves of shader cycles per wall second
        *shader_ghz = 0.0;
        int dev = 0, khz = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        AccelArgs a{sx, sy, sz, gm, px, py, pz, ax, ay, az, n_src, n_sink, n_self, mode};
        const AccelPlan pl = plan_accel(n_src, n_sink, mode, device_cu_count(dev));
        const int nw = clocked_waves(a, pl);
        // as many launches as the timed run (the clock settles over the first
        // milliseconds of a load); the last nl of them keep their ticks
        const int nl = std::min(reps, 8);
        if (nw > 0 && khz > 0) {
            const size_t per = 4 * (size_t)nw;
            long long *ticks = nullptr;
            HIP_TRY(hipMalloc((void **)&ticks, per * nl * sizeof(long long)));
            hipError_t e = hipSuccess;
            for (int r = 0; r < reps && e == hipSuccess; ++r)
                e = launch_accel_clocked(a, pl, ticks + per * (r % nl), st);  // nl consecutive r: distinct slots
            std::vector<long long> h(per * nl);
            if (e == hipSuccess) e = hipMemcpyAsync(h.data(), ticks, h.size() * sizeof(long long), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            (void)hipFree(ticks);
            if (e != hipSuccess) return fail(ABIE_HIP_EHIP, "time_accel_soa: clocked launches: %s", hipGetErrorString(e));
            std::vector<double> launch_ghz, wave_ghz;
            for (int r = 0; r < nl; ++r) {
                wave_ghz.clear();
                for (int w = 0; w < nw; ++w) {
                    const long long *t = h.data() + per * r + 4 * w;
                    const double wall_s = (double)(t[1] - t[0]) / (khz * 1e3);
                    if (wall_s > 0) wave_ghz.push_back((double)(t[3] - t[2]) / wall_s / 1e9);
                }
                if (wave_ghz.empty()) continue;
                std::nth_element(wave_ghz.begin(), wave_ghz.begin() + wave_ghz.size() / 2, wave_ghz.end());
                launch_ghz.push_back(wave_ghz[wave_ghz.size() / 2]);
            }
            if (!launch_ghz.empty()) {
                std::sort(launch_ghz.begin(), launch_ghz.end());
                *shader_ghz = launch_ghz[launch_ghz.size() / 2];
            }
        }
    }
    return 0;
}

}  // extern "C"
