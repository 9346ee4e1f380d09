// noalloc.hip -- abie_hip_accel_soa_noalloc: the SoA force evaluation for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  The source-split partial sums that
// abie_hip_accel_soa takes from a stream-ordered allocation live here in a
// file-scope device array instead, one copy per device, held until the
// process ends.  The kernels are the same (force.hip), so the results are
// the same bits.
#include <stdio.h>

#include <mutex>

#include "abie_hip.h"
#include "force.hpp"

namespace abie {

int report_error(int code, const char *msg);  // runtime.hip: sets abie_hip_last_error

// 2^21 doubles = 16 MiB per device.  plan_accel splits the sources only while
// the sink blocks number fewer than 4 per CU, so a plan needs
// parts * 3 * n_sink < 3 * 2 * (4 * n_cu) * 256 doubles: 12 MiB on a 256-CU
// MI355X.
constexpr size_t kScratchDoubles = (size_t)1 << 21;
constexpr int kMaxDevices = 64;

// Host code never uses this name as an address (it is a host-side stand-in):
// scratch_ptr() looks up the calling device's copy once and keeps it.
static __device__ __align__(256) double scratch[kScratchDoubles];

static hipError_t scratch_ptr(int dev, double **ptr) {
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (addr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = static_cast<double *>(p);
    }
    *ptr = addr[dev];
    return hipSuccess;
}

}  // namespace abie

using namespace abie;

extern "C" int abie_hip_accel_soa_noalloc(const double *sx, const double *sy, const double *sz, const double *gm,
                                          int n_src, const double *px, const double *py, const double *pz,
                                          int n_sink, int n_self, double *ax, double *ay, double *az, int mode,
                                          void *stream) {
    char msg[256];
    if (n_src < 0 || n_sink < 0 || n_self < 0 || n_self > n_src || n_self > n_sink) {
        snprintf(msg, sizeof msg, "accel_soa_noalloc: bad sizes (n_src %d, n_sink %d, n_self %d)", n_src, n_sink,
                 n_self);
        return report_error(ABIE_HIP_EINVAL, msg);
    }
    if (mode < 0 || mode > 3) {
        snprintf(msg, sizeof msg, "accel_soa_noalloc: bad mode %d", mode);
        return report_error(ABIE_HIP_EINVAL, msg);
    }
    int dev = 0, n_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) {
        snprintf(msg, sizeof msg, "accel_soa_noalloc: %s", hipGetErrorString(e));
        return report_error(ABIE_HIP_EHIP, msg);
    }
    if (dev < 0 || dev >= kMaxDevices) {
        snprintf(msg, sizeof msg, "accel_soa_noalloc: device %d beyond the %d the scratch table holds", dev,
                 kMaxDevices);
        return report_error(ABIE_HIP_ENODEV, msg);
    }
    AccelArgs a{sx, sy, sz, gm, px, py, pz, ax, ay, az, n_src, n_sink, n_self, mode};
    const AccelPlan pl = plan_accel(n_src, n_sink, mode, n_cu > 0 ? n_cu : 256);
    const size_t need = plan_scratch_doubles(pl, n_sink);
    if (need > kScratchDoubles) {
        snprintf(msg, sizeof msg, "accel_soa_noalloc: %d sources x %d sinks need %zu doubles of scratch, more "
                 "than the %zu the library holds per device", n_src, n_sink, need, kScratchDoubles);
        return report_error(ABIE_HIP_ETOOBIG, msg);
    }
    double *part = nullptr;
    if (need > 0) {
        e = scratch_ptr(dev, &part);
        if (e != hipSuccess) {
            snprintf(msg, sizeof msg, "accel_soa_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
            return report_error(ABIE_HIP_EHIP, msg);
        }
    }
    e = launch_accel(a, pl, part, (hipStream_t)stream);
    if (e != hipSuccess) {
        snprintf(msg, sizeof msg, "accel_soa_noalloc: launch failed: %s", hipGetErrorString(e));
        return report_error(ABIE_HIP_EHIP, msg);
    }
    return 0;
}
