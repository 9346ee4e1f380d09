// gr15.hpp -- the batched, device-resident Gauss-Radau15 engine (gr15.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "abie_hip.h"

namespace abie {

// One independent system: bodies [body0, body0 + n) of the flat arrays, in
// engine order (massive bodies first, ascending original index, then the
// rest, ascending).  nm = number of massive bodies (m > 0).  A system split
// into several units (test-particle slices) has one SysDesc per unit: nt_sys
// is the whole system's test-particle count, sys_id its index and primary is
// 1 on the unit whose massive bodies are reported.  norm_all = 1 makes a
// massive-only controller use every body in its norms (systems with at most
// one massive body, whose massive norm would be 0/0).
//
// Linked units (link_n > 1): a system too large for one workgroup under the
// plain controller runs as link_n slices, units link_base .. link_base +
// link_n - 1 (link_rank 0 is the primary), each carrying every massive body
// and a share of the test particles.  The plain controller's norms are over
// all bodies, so every maximum is combined over the slices through a grid
// barrier (the launch is cooperative: all slices resident); sys_n is the
// system's body count and sys_inv0 the offset of its bodies, by ascending
// original index, in GR15Args::sys_inv.
struct SysDesc {
    int body0, n, nm, cfg;
    int nt_sys, sys_id, primary, norm_all;
    int link_n, link_rank, link_base, sys_n, sys_inv0;
};

// Per-system control block, device-resident; persists across launches.
struct SysCtl {
    double t, h, t_end;
    int status;   // ST_*
    int phase;    // 0: needs the initial step, 1: stepping
    int rejects;  // consecutive rejections (step_loop_count)
    int warned;
    long long n_evals, n_attempts, n_accepted, n_pc, n_prop_used, n_prop_rej;
    long long n_ce, n_coll;  // event counters (sim.n_close_encounters, ...)
    double prev_h, prev_err, prev_dtreq;  // history used by the step proposer
    int prev_pc, pad;
    long long sync_epoch;  // grid reductions done by a linked unit (the same on every slice)
    // device wall-clock ticks (constant-rate clock) of the unit's first
    // launch of the current run and of the moment it stopped running; 0
    // until set (cleared by the host at the start of each run)
    long long start_tick, done_tick;
    // a unit stops (state kept) once its attempt count reaches this, 0: no
    // such stop (the host brings the slices of a split unit to one attempt)
    long long stop_attempt;
    // device wall-clock ticks the unit has spent in launches, summed (its
    // progress rate for the host's straggler promotion)
    long long busy_ticks;
};

enum : int { ST_RUNNING = 0, ST_CE = 1, ST_COLL = 2, ST_DONE = 3, ST_ERROR = -1 };

// Number of persistent doubles per body: b[7][3], g[7][3], a0[3], b6_prev[3].
constexpr int kStateDoubles = 48;

struct GR15Args {
    const SysDesc *sys;
    const int *order;  // block -> system index
    SysCtl *ctl;
    // flat body arrays (engine order)
    double *x, *y, *z, *vx, *vy, *vz;  // positions / velocities, updated in place
    const double *gm;                  // G*m for massive bodies, else 0
    const double *rad;                 // radii
    const double *ext;                 // [3][total] constant extra acceleration, or null
    const int *orig;                   // original index within the system
    const int *inv;                    // the unit's engine indices by ascending original index
    const unsigned char *deleted;      // m < 0
    double *state;                     // [kStateDoubles][total]
    double *scan;                      // [4][total] event-scan positions x, y, z and candidate flags
    long long total;
    // events (per system rings, 4 doubles per event)
    double ce_dist;
    int max_ce, max_coll;
    double *ring_ce, *ring_coll;  // [n_sys][4*max]
    // step log
    double *log;
    unsigned long long *log_len;
    long long log_cap;
    // learned proposer (FP32 weights) or null
    const float *mlp;
    const double *mratio;  // per unit: (sum of the other massive masses) / largest mass
    int max_attempts;  // per launch
    long long budget_ticks;  // wall-clock ticks a group may run (checked every 4 attempts); 0: no limit
    long long deadline_tick; // device wall clock at which every unit stops (state kept); 0: none
    int *running;            // units not yet stopped for good in this run; each decrements it once
    int tail_units;          // units of the massive-only controllers stop (state kept) once
                             // *running <= tail_units (the host then moves them to the quad form,
                             // or splits them finer); 0: never
    int prefetch;            // units with at least this many massive bodies use the prefetching
                             // source loop (massive-only controllers); 0: never
    int events;        // 0: skip the close-encounter / collision scan
    // linked units: one arrival counter per system (at its link_base),
    // partial maxima [2][n_units][2] (two parities), bodies of each linked
    // system by ascending original index (global body slots), and the
    // longest a grid barrier may wait before the unit reports ST_ERROR
    unsigned long long *link_count;
    double *link_part;
    const int *sys_inv;
    int n_units;
    long long spin_ticks;
};

// Engine configurations: threads per block and bodies per thread.
struct GR15Cfg {
    int nt, bpt;
};
// Massive bodies one group holds in its LDS source list (<= 64 KB of LDS).
constexpr int src_cap(int nt) { return nt < 1008 ? nt : 1008; }
constexpr int kNumCfg = 11;
extern const GR15Cfg kCfgs[kNumCfg];
int pick_cfg(int n, int nm);  // by the tuning rules; -1 if no configuration fits
bool cfg_fits(int cfg, int n, int nm);
int max_bodies_one_group();
// Tuning rules: a system of n bodies uses the first rule with n <= max_n whose
// configuration fits it.
int set_tuning(int n_rules, const int *max_n, const int *cfg);
size_t gr15_lds_bytes(int cfg);
// Workgroups of configuration cfg that one CU holds at once (its waves per
// SIMD target and its LDS): the bound on the slices of one cooperative launch.
int gr15_blocks_per_cu(int cfg);

// The quad form (gr15_quad.hpp, massive-only controllers): four lanes per
// body, groups of kQuadNT[q] threads holding kQuadNT[q] / 4 bodies.  A unit
// in quad group size q carries the configuration index kNumCfg + q.
constexpr int kNumQuad = 5;
constexpr int kQuadNT[kNumQuad] = {64, 128, 256, 512, 1024};
constexpr int kQuadMaxBodies = 256;
constexpr int kNumAllCfg = kNumCfg + kNumQuad;
inline bool cfg_is_quad(int cfg) { return cfg >= kNumCfg && cfg < kNumAllCfg; }
// the smallest quad group size that holds n bodies, or -1
int pick_quad(int n);
size_t gr15q_lds_bytes(int q);
hipError_t launch_gr15q(int q, int force_mode, int ctrl, const GR15Args &a, int n_blocks, hipStream_t st,
                        size_t min_lds = 0);

// The learned proposer's weights are staged in LDS after the 128 doubles of
// scratch when a kernel of the learned controller has a network: the network
// then reads them from LDS instead of fetching 20 KB through the caches in
// rounds of dependent loads once per accepted attempt.
constexpr size_t kMlpLdsBytes = (((size_t)ABIE_MLP_NPARAM * sizeof(float) + 255) / 256) * 256;
inline size_t mlp_lds_bytes(const GR15Args &a, int ctrl) {
    return (ctrl == ABIE_CTRL_LEARNED && a.mlp != nullptr) ? kMlpLdsBytes : 0;
}

bool gr15_combo_supported(int force_mode, int ctrl);
// coop: a cooperative launch (every block resident at once, or an error),
// for the slices of one linked system.
hipError_t launch_gr15(int cfg, int force_mode, int ctrl, const GR15Args &a, int n_blocks, hipStream_t st,
                       bool coop = false);
// flags[0]: hit flag, 0 on entry; flags[1] receives 0, 1 or 2; evflag: n
// ints of candidate flags, 0 on entry.
hipError_t launch_events(const double *pos, const double *rad, int n, double t, double cd, long long *cnt,
                         int max_ce, int max_coll, double *ring_ce, double *ring_coll, int *flags, int *evflag,
                         hipStream_t st);
// The per-body arrays of a set of units (the ensemble's, or the split tail's).
struct BodyArrays {
    double *x, *y, *z, *vx, *vy, *vz, *gm, *m, *rad, *state;
    int *orig, *inv;
    unsigned char *del;
    long long total;  // bodies; the state is [kStateDoubles][total]
};
// dst body i <- src body map[i], every field (i < n).
hipError_t launch_body_gather(const long long *map, long long n, const BodyArrays &src, const BodyArrays &dst,
                              hipStream_t st);
// dst body map[i] <- src body i for positions, velocities and the state,
// where map[i] >= 0 (i < n).
// dst[i] = src[idx[i]], i < n, on stream st
hipError_t launch_ctl_gather(const SysCtl *src, const int *idx, int n, SysCtl *dst, hipStream_t st);
// dst[idx[i]] = src[i], i < n, on stream st
hipError_t launch_ctl_scatter(const SysCtl *src, const int *idx, int n, SysCtl *dst, hipStream_t st);
hipError_t launch_body_scatter(const long long *map, long long n, const BodyArrays &src, const BodyArrays &dst,
                               hipStream_t st);

// The device's constant-rate wall clock (wall_clock64) into *host_tick.
hipError_t read_device_clock(long long *host_tick, long long *dev_buf, hipStream_t st);
// *ok: the n streams ran a probe kernel each at the same time (no two share a
// hardware queue); dev_buf holds 2 unsigned ints of scratch; each probe waits
// at most wait_ticks of the constant-rate clock.
hipError_t streams_concurrent(const hipStream_t *st, int n, unsigned *dev_buf, long long wait_ticks, bool *ok);
hipError_t launch_energy(const SysDesc *sys, int n_sys, const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz, const double *m, double G,
                         double *energy, hipStream_t st);

}  // namespace abie
