// gr15.hip -- configuration table, dispatch, energy and event kernels of
// the GR15 engine (the engine itself is gr15_impl.hpp).
#include <algorithm>
#include <utility>
#include <vector>

#include "devmath.hpp"
#include "gr15.hpp"

namespace abie {

const GR15Cfg kCfgs[kNumCfg] = {{64, 1},  {128, 1}, {256, 1}, {256, 2}, {512, 1}, {256, 4},
                                 {512, 2}, {1024, 1}, {512, 3}, {512, 4}, {512, 5}};

bool cfg_fits(int c, int n, int nm) {
    return c >= 0 && c < kNumCfg && n <= kCfgs[c].nt * kCfgs[c].bpt && nm <= src_cap(kCfgs[c].nt);
}

int max_bodies_one_group() {
    int m = 0;
    for (int c = 0; c < kNumCfg; ++c) m = std::max(m, kCfgs[c].nt * kCfgs[c].bpt);
    return m;
}

// Default rules: the smallest group that holds the system, favouring
// configurations whose state fits the register file.
static std::vector<std::pair<int, int>> g_rules = {{64, 0}, {128, 1}, {256, 2}, {512, 3}, {1024, 5},
                                                   {1536, 8}, {2048, 9}, {2560, 10}};

int set_tuning(int n_rules, const int *max_n, const int *cfg) {
    if (n_rules <= 0 || !max_n || !cfg) return -1;
    std::vector<std::pair<int, int>> r;
    for (int i = 0; i < n_rules; ++i) {
        if (cfg[i] < 0 || cfg[i] >= kNumCfg) return -1;
        r.push_back({max_n[i], cfg[i]});
    }
    g_rules = r;
    return 0;
}

int pick_cfg(int n, int nm) {
    for (auto &r : g_rules)
        if (n <= r.first && cfg_fits(r.second, n, nm)) return r.second;
    // no rule: the smallest configuration that fits
    int best = -1;
    for (int c = 0; c < kNumCfg; ++c)
        if (cfg_fits(c, n, nm) && (best < 0 || kCfgs[c].nt * kCfgs[c].bpt < kCfgs[best].nt * kCfgs[best].bpt))
            best = c;
    return best;
}

// LDS: sources [2][NT] x {x, y, z, gm}, then 128 doubles of scratch:
// [0, 32) reductions, [40, 51) broadcast flags and linked-slice maxima,
// [64, 96) proposer hidden layer, [96, 109) proposer inputs and output.
size_t gr15_lds_bytes(int cfg) { return (size_t)2 * src_cap(kCfgs[cfg].nt) * 4 * sizeof(double) + 128 * sizeof(double); }

// waves per SIMD each configuration's kernel is built for (the WPE template
// argument of launch_gr15_fm_ct), 4 SIMDs and 160 KB of LDS per CU
static const int kCfgWpe[kNumCfg] = {2, 3, 3, 1, 3, 1, 2, 4, 2, 2, 2};

int gr15_blocks_per_cu(int cfg) {
    if (cfg < 0 || cfg >= kNumCfg) return 0;
    const int by_waves = 4 * kCfgWpe[cfg] / (kCfgs[cfg].nt / 64);
    const int by_lds = (int)((size_t)160 * 1024 / gr15_lds_bytes(cfg));
    return std::min(by_waves, by_lds);
}


// one translation unit per (force mode, controller) pair, compiled in parallel
hipError_t launch_gr15_fp64_plain(int, const GR15Args &, int, size_t, hipStream_t, bool);
hipError_t launch_gr15_fp64_massive(int, const GR15Args &, int, size_t, hipStream_t, bool);
hipError_t launch_gr15_fp64_learned(int, const GR15Args &, int, size_t, hipStream_t, bool);
hipError_t launch_gr15_exact_plain(int, const GR15Args &, int, size_t, hipStream_t, bool);
hipError_t launch_gr15_mixed_massive(int, const GR15Args &, int, size_t, hipStream_t, bool);
hipError_t launch_gr15_mixed_learned(int, const GR15Args &, int, size_t, hipStream_t, bool);

hipError_t launch_gr15q_fp64_massive(int, const GR15Args &, int, size_t, hipStream_t);
hipError_t launch_gr15q_fp64_learned(int, const GR15Args &, int, size_t, hipStream_t);
hipError_t launch_gr15q_mixed_massive(int, const GR15Args &, int, size_t, hipStream_t);
hipError_t launch_gr15q_mixed_learned(int, const GR15Args &, int, size_t, hipStream_t);

int pick_quad(int n) {
    for (int q = 0; q < kNumQuad; ++q)
        if (n <= kQuadNT[q] / 4) return q;
    return -1;
}

// LDS of the quad form: source lists [2][NT / 4] x {x, y, z, gm}, then the
// same 128 doubles of scratch as gr15_kernel
size_t gr15q_lds_bytes(int q) { return (size_t)2 * (kQuadNT[q] / 4) * 4 * sizeof(double) + 128 * sizeof(double); }

hipError_t launch_gr15q(int q, int fm, int ctrl, const GR15Args &a, int nb, hipStream_t st, size_t min_lds) {
    if (nb <= 0) return hipSuccess;
    if (q < 0 || q >= kNumQuad) return hipErrorInvalidValue;
    const size_t lds = std::max(gr15q_lds_bytes(q) + mlp_lds_bytes(a, ctrl), min_lds);
    if (fm == ABIE_FORCE_FP64 && ctrl == ABIE_CTRL_MASSIVE) return launch_gr15q_fp64_massive(q, a, nb, lds, st);
    if (fm == ABIE_FORCE_FP64 && ctrl == ABIE_CTRL_LEARNED) return launch_gr15q_fp64_learned(q, a, nb, lds, st);
    if (fm == ABIE_FORCE_MIXED && ctrl == ABIE_CTRL_MASSIVE) return launch_gr15q_mixed_massive(q, a, nb, lds, st);
    if (fm == ABIE_FORCE_MIXED && ctrl == ABIE_CTRL_LEARNED) return launch_gr15q_mixed_learned(q, a, nb, lds, st);
    return hipErrorInvalidValue;
}

__global__ void device_clock_kernel(long long *out) { *out = wall_clock64(); }

__global__ void body_gather_kernel(const long long *__restrict__ map, long long n, BodyArrays src, BodyArrays dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = map[i];
    dst.x[i] = src.x[j]; dst.y[i] = src.y[j]; dst.z[i] = src.z[j];
    dst.vx[i] = src.vx[j]; dst.vy[i] = src.vy[j]; dst.vz[i] = src.vz[j];
    dst.gm[i] = src.gm[j]; dst.m[i] = src.m[j]; dst.rad[i] = src.rad[j];
    dst.orig[i] = src.orig[j]; dst.inv[i] = src.inv[j]; dst.del[i] = src.del[j];
    for (int r = 0; r < kStateDoubles; ++r) dst.state[r * dst.total + i] = src.state[r * src.total + j];
}

__global__ void body_scatter_kernel(const long long *__restrict__ map, long long n, BodyArrays src, BodyArrays dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = map[i];
    if (j < 0) return;
    dst.x[j] = src.x[i]; dst.y[j] = src.y[i]; dst.z[j] = src.z[i];
    dst.vx[j] = src.vx[i]; dst.vy[j] = src.vy[i]; dst.vz[j] = src.vz[i];
    for (int r = 0; r < kStateDoubles; ++r) dst.state[r * dst.total + j] = src.state[r * src.total + i];
}

hipError_t launch_body_gather(const long long *map, long long n, const BodyArrays &src, const BodyArrays &dst,
                              hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(body_gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, map, n, src, dst);
    return hipGetLastError();
}

// dst[idx[i]] = src[i] for i < n (control blocks of split units merged back
// from their children, one launch instead of one copy per unit)
__global__ void ctl_scatter_kernel(const SysCtl *__restrict__ src, const int *__restrict__ idx, int n,
                                   SysCtl *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}

hipError_t launch_ctl_scatter(const SysCtl *src, const int *idx, int n, SysCtl *dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ctl_scatter_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, src, idx, n, dst);
    return hipGetLastError();
}

// dst[i] = src[idx[i]] for i < n (the control blocks of the units still
// running, read back without the rest)
__global__ void ctl_gather_kernel(const SysCtl *__restrict__ src, const int *__restrict__ idx, int n,
                                  SysCtl *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

hipError_t launch_ctl_gather(const SysCtl *src, const int *idx, int n, SysCtl *dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ctl_gather_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, src, idx, n, dst);
    return hipGetLastError();
}

hipError_t launch_body_scatter(const long long *map, long long n, const BodyArrays &src, const BodyArrays &dst,
                               hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(body_scatter_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, map, n, src, dst);
    return hipGetLastError();
}

// Concurrency probe of n streams: one single-wave kernel per stream counts
// itself in and waits until all n have arrived, or until wait_ticks of the
// constant-rate clock have passed, when it flags *late.  Launches on
// streams that share a hardware queue run one after the other, so the first
// of two such probes times out.  Every wait ends.
__global__ void stream_probe_kernel(unsigned *count, unsigned n, long long wait_ticks, unsigned *late) {
    if (threadIdx.x != 0) return;
    __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long t0 = wall_clock64();
    while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n) {
        if (wall_clock64() - t0 > wait_ticks) {
            __hip_atomic_fetch_or(late, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

hipError_t streams_concurrent(const hipStream_t *st, int n, unsigned *dev_buf, long long wait_ticks, bool *ok) {
    hipError_t e = hipMemset(dev_buf, 0, 2 * sizeof(unsigned));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        hipLaunchKernelGGL(stream_probe_kernel, dim3(1), dim3(64), 0, st[i], dev_buf, (unsigned)n, wait_ticks,
                           dev_buf + 1);
        e = hipGetLastError();
    }
    for (int i = 0; i < n; ++i) {
        const hipError_t f = hipStreamSynchronize(st[i]);
        if (e == hipSuccess) e = f;
    }
    unsigned late = 1;
    if (e == hipSuccess) e = hipMemcpy(&late, dev_buf + 1, sizeof(unsigned), hipMemcpyDeviceToHost);
    *ok = e == hipSuccess && late == 0;
    return e;
}

hipError_t read_device_clock(long long *host_tick, long long *dev_buf, hipStream_t st) {
    hipLaunchKernelGGL(device_clock_kernel, dim3(1), dim3(1), 0, st, dev_buf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(host_tick, dev_buf, sizeof(long long), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

bool gr15_combo_supported(int fm, int ctrl) {
    if (fm == ABIE_FORCE_FP64) return ctrl >= 0 && ctrl <= 2;
    if (fm == ABIE_FORCE_FP64_EXACT) return ctrl == ABIE_CTRL_PLAIN;
    if (fm == ABIE_FORCE_MIXED) return ctrl == ABIE_CTRL_MASSIVE || ctrl == ABIE_CTRL_LEARNED;
    return false;
}

hipError_t launch_gr15(int cfg, int fm, int ctrl, const GR15Args &a, int nb, hipStream_t st, bool coop) {
    if (nb <= 0) return hipSuccess;
    if (cfg < 0 || cfg >= kNumCfg) return hipErrorInvalidValue;
    const size_t lds = gr15_lds_bytes(cfg) + mlp_lds_bytes(a, ctrl);
    if (fm == ABIE_FORCE_FP64) {
        if (ctrl == ABIE_CTRL_PLAIN) return launch_gr15_fp64_plain(cfg, a, nb, lds, st, coop);
        if (ctrl == ABIE_CTRL_MASSIVE) return launch_gr15_fp64_massive(cfg, a, nb, lds, st, coop);
        return launch_gr15_fp64_learned(cfg, a, nb, lds, st, coop);
    }
    if (fm == ABIE_FORCE_FP64_EXACT && ctrl == ABIE_CTRL_PLAIN) return launch_gr15_exact_plain(cfg, a, nb, lds, st, coop);
    if (fm == ABIE_FORCE_MIXED) {
        if (ctrl == ABIE_CTRL_MASSIVE) return launch_gr15_mixed_massive(cfg, a, nb, lds, st, coop);
        if (ctrl == ABIE_CTRL_LEARNED) return launch_gr15_mixed_learned(cfg, a, nb, lds, st, coop);
    }
    return hipErrorInvalidValue;
}

// check_collisions_close_encounters (common.c:116-158) for one system, in two
// kernels.  events_filter_kernel: block b owns bodies j in [256 b, 256 b +
// 256) and streams the bodies k >= 256 b through LDS tiles, testing each
// unordered pair once with a test that can only over-report (r^2 in the
// reference's operation order against cd^2 or (r_j + r_k)^2 with a 2^-50
// slack for the rounding of sqrt); any pass sets *hit, and a wave with a
// pass in a tile tests that tile again and flags both bodies of every
// passing pair in evflag (zero on entry).  events_exact_kernel: only if
// *hit, one wave redoes the reference's ordered (j, k) loop exactly over the
// flagged bodies (an event pair passes the filter's test, so the skipped
// pairs hold no event), j serially and the k > j in chunks of 64 lanes, each event's ring slot from
// a ballot prefix count (an event is skipped when a later one of the same
// chunk takes its slot), and updates the counters (cnt[0] = close encounters,
// cnt[1] = collisions) and the rings as the serial loop would.
constexpr int kEvTile = 256;

__device__ __forceinline__ bool ev_pass(double xj, double yj, double zj, double rj, double tx, double ty,
                                        double tz, double tr, double cd2) {
    const double dx = xj - tx, dy = yj - ty, dz = zj - tz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double rr = rj + tr;
    return (r2 <= cd2) || (rr > 0 && r2 <= rr * rr * (1.0 + 0x1p-50));
}

__global__ __launch_bounds__(kEvTile) void events_filter_kernel(const double *__restrict__ pos,
                                                                const double *__restrict__ rad, int n, double cd,
                                                                int *hit, int *evflag) {
    __shared__ double tx[kEvTile], ty[kEvTile], tz[kEvTile], tr[kEvTile];
    const int j = blockIdx.x * kEvTile + threadIdx.x;
    const bool valid = j < n;
    const int jc = valid ? j : n - 1;
    const double xj = pos[3 * jc], yj = pos[3 * jc + 1], zj = pos[3 * jc + 2], rj = rad[jc];
    const double cd2 = cd * cd * (1.0 + 0x1p-50);
    bool mine = false;
    for (int k0 = blockIdx.x * kEvTile; k0 < n; k0 += kEvTile) {
        const int cnt = min(kEvTile, n - k0);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const int k = k0 + threadIdx.x;
            tx[threadIdx.x] = pos[3 * k];
            ty[threadIdx.x] = pos[3 * k + 1];
            tz[threadIdx.x] = pos[3 * k + 2];
            tr[threadIdx.x] = rad[k];
        }
        __syncthreads();
        bool tile_hit = false;
        for (int q = 0; q < cnt; ++q)
            tile_hit |= valid && (k0 + q > j) && ev_pass(xj, yj, zj, rj, tx[q], ty[q], tz[q], tr[q], cd2);
        if (__ballot(tile_hit)) {  // rare: flag the passing pairs of this tile
            for (int q = 0; q < cnt; ++q)
                if (valid && (k0 + q > j) && ev_pass(xj, yj, zj, rj, tx[q], ty[q], tz[q], tr[q], cd2)) {
                    evflag[j] = 1;
                    evflag[k0 + q] = 1;
                }
        }
        mine |= tile_hit;
    }
    if (valid && mine) atomicOr(hit, 1);
}

__global__ __launch_bounds__(64) void events_exact_kernel(const double *__restrict__ pos,
                                                          const double *__restrict__ rad, int n, double t, double cd,
                                                          long long *cnt, int max_ce, int max_coll, double *ring_ce,
                                                          double *ring_coll, const int *hit, int *ret,
                                                          const int *evflag) {
    const int lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    long long nce = cnt[0], nco = cnt[1];
    if (*hit) {
        for (int j = 0; j < n; ++j) {
            if (evflag[j] == 0) continue;  // no candidate pair (the same for every lane)
            const double xj = pos[3 * j], yj = pos[3 * j + 1], zj = pos[3 * j + 2], rj = rad[j];
            for (int c = j + 1; c < n; c += 64) {
                const int k = c + lane;
                bool ce = false, co = false;
                double d = 0.0;
                if (k < n && evflag[k] != 0) {
                    const double dx = xj - pos[3 * k], dy = yj - pos[3 * k + 1], dz = zj - pos[3 * k + 2];
                    d = sqrt(dx * dx + dy * dy + dz * dz);
                    const double rr = rj + rad[k];
                    ce = d <= cd;
                    co = rr > 0 && d <= rr;
                }
                const unsigned long long mce = __ballot(ce), mco = __ballot(co);
                if (mce | mco) {
                    const long long ece = nce + __popcll(mce), eco = nco + __popcll(mco);
                    if (ce && max_ce > 0) {
                        const long long idx = nce + __popcll(mce & below);
                        if (idx >= ece - max_ce) {
                            const long long w = 4LL * max_ce, o = 4 * idx;
                            ring_ce[o % w] = t; ring_ce[(o + 1) % w] = j; ring_ce[(o + 2) % w] = k;
                            ring_ce[(o + 3) % w] = d;
                        }
                    }
                    if (co && max_coll > 0) {
                        const long long idx = nco + __popcll(mco & below);
                        if (idx >= eco - max_coll) {
                            const long long w = 4LL * max_coll, o = 4 * idx;
                            ring_coll[o % w] = t; ring_coll[(o + 1) % w] = j; ring_coll[(o + 2) % w] = k;
                            ring_coll[(o + 3) % w] = d;
                        }
                    }
                    nce = ece;
                    nco = eco;
                    // the next chunk's writes may reuse these slots
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                }
            }
        }
    }
    if (lane == 0) {
        cnt[0] = nce;
        cnt[1] = nco;
        *ret = (max_ce > 0 && nce >= max_ce) ? 1 : (max_coll > 0 && nco >= max_coll) ? 2 : 0;
    }
}

// flags[0] must be 0 on entry (hit flag); flags[1] receives 0, 1 or 2;
// evflag (n ints) must be 0 on entry.
hipError_t launch_events(const double *pos, const double *rad, int n, double t, double cd, long long *cnt,
                         int max_ce, int max_coll, double *ring_ce, double *ring_coll, int *flags, int *evflag,
                         hipStream_t st) {
    if (n <= 0) return hipErrorInvalidValue;
    const int nb = (n + kEvTile - 1) / kEvTile;
    hipLaunchKernelGGL(events_filter_kernel, dim3(nb), dim3(kEvTile), 0, st, pos, rad, n, cd, flags, evflag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(events_exact_kernel, dim3(1), dim3(64), 0, st, pos, rad, n, t, cd, cnt, max_ce, max_coll,
                       ring_ce, ring_coll, flags, flags + 1, evflag);
    return hipGetLastError();
}

// calculate_energy (common.c:340-357) per system: one block per system; each
// thread forms body i's term (kinetic, then its pair terms in ascending j),
// and the block sums the per-body terms in a fixed tree order.
__global__ __launch_bounds__(256) void energy_kernel(const SysDesc *sys, const double *x, const double *y,
                                                     const double *z, const double *vx, const double *vy,
                                                     const double *vz, const double *m, double G, double *out) {
    __shared__ double part[256];
    const SysDesc d = sys[blockIdx.x];
    double acc = 0.0;
    for (int i = threadIdx.x; i < d.n; i += 256) {
        const int gi = d.body0 + i;
        const double mi = m[gi];
        if (mi == 0) continue;
        const double v = sqrt(vx[gi] * vx[gi] + vy[gi] * vy[gi] + vz[gi] * vz[gi]);
        double e = 0.5 * mi * (v * v);
        for (int j = 0; j < d.n; ++j) {
            const int gj = d.body0 + j;
            if (j == i || m[gj] == 0) continue;
            const double dx = x[gi] - x[gj], dy = y[gi] - y[gj], dz = z[gi] - z[gj];
            e -= 0.5 * G * mi * m[gj] / sqrt(dx * dx + dy * dy + dz * dz);
        }
        acc += e;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

hipError_t launch_energy(const SysDesc *sys, int n_sys, const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz, const double *m, double G,
                         double *energy, hipStream_t st) {
    if (n_sys <= 0) return hipSuccess;
    hipLaunchKernelGGL(energy_kernel, dim3(n_sys), dim3(256), 0, st, sys, x, y, z, vx, vy, vz, m, G, energy);
    return hipGetLastError();
}

}  // namespace abie
