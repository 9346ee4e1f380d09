// force.hpp -- launch interface of the all-pairs force kernels (force.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "abie_hip.h"

namespace abie {

struct AccelArgs {
    const double *sx, *sy, *sz, *gm;  // sources, SoA, gm = G*m
    const double *px, *py, *pz;       // sinks, SoA
    double *ax, *ay, *az;             // accelerations, SoA
    int n_src, n_sink, n_self;        // sinks [0, n_self) are sources [0, n_self)
    int mode;                         // ABIE_FORCE_*
};

struct AccelPlan {
    int spt = 4;          // sinks per thread
    int parts = 1;        // source-range split (partials summed in order)
    int scalar_src = 0;   // 1: sources through scalar loads (accel_scalar_kernel)
};

AccelPlan plan_accel(int n_src, int n_sink, int mode, int n_cu);
// Doubles of source-split scratch a plan needs (0 when it does not split).
inline size_t plan_scratch_doubles(const AccelPlan &pl, int n_sink) {
    return pl.parts > 1 ? (size_t)pl.parts * 3 * (size_t)n_sink : 0;
}
// `part` must hold parts*3*n_sink doubles when plan.parts > 1.
hipError_t launch_accel(const AccelArgs &a, const AccelPlan &plan, double *part, hipStream_t st);
// The scalar-source unsplit plan (spt 1, parts 1) with per-wave shader-cycle
// and wall-clock ticks (4 per wave) recorded; clocked_waves: the waves such a
// launch records, 0 for plans it does not cover.
int clocked_waves(const AccelArgs &a, const AccelPlan &plan);
hipError_t launch_accel_clocked(const AccelArgs &a, const AccelPlan &plan, long long *ticks, hipStream_t st);

hipError_t launch_pack(const double *pos, const double *m, const int *perm, int n, int n_src, double G,
                       double *x, double *y, double *z, double *gm, hipStream_t st);
hipError_t launch_unpack(const int *perm, const double *m, int n, const double *ax, const double *ay,
                         const double *az, const double *ext, double *acc, hipStream_t st);

}  // namespace abie
