/*
 * abie_hip.h -- C ABI of libabie_hip.so, the MI355X (gfx950) backend of
 * ABIE's all-pairs gravity path and of the Gauss-Radau15 integrator that
 * drives it.
 *
 * Part 1 keeps the reference's C ABI (names, argument meaning, return codes)
 * so that abie/clibabie.py can bind this library in place of libabie.so when
 * deviceID >= 0.  Each declaration cites the reference symbol it replaces.
 * All pointers in part 1 are caller-owned HOST arrays (AoS, C-contiguous,
 * double); the library copies them and never keeps them past the call.
 *
 * Part 2 is project-defined: device-pointer entry points (SoA, explicit HIP
 * stream) for the force kernel on its own, the batched GR15 ensemble engine
 * and the learned step proposer.  They return 0 on success and a negative
 * ABIE_HIP_E* code on error; they never exit the process.
 *
 * Threading: part 1 drives one process-global simulation, as the reference
 * does (src/common.h:70); calls must not overlap.  Part 2 objects are
 * independent handles; one handle must not be used by two threads at once.
 */
#ifndef ABIE_HIP_H
#define ABIE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes of integrator_gr (src/common.h:6-8). */
#define ABIE_EXIT_NORMAL 0
#define ABIE_EXIT_MAX_N_CE_EXCEEDED 1
#define ABIE_EXIT_MAX_N_COLLISIONS_EXCEEDED 2

/* Errors (negative; never returned by the reference, which exits instead). */
#define ABIE_HIP_EINVAL   (-1)   /* bad argument                              */
#define ABIE_HIP_ENODEV   (-2)   /* no usable HIP device / deviceID invalid   */
#define ABIE_HIP_ENOMEM   (-3)   /* device or host allocation failed          */
#define ABIE_HIP_EHIP     (-4)   /* a HIP runtime call failed                 */
#define ABIE_HIP_EUNSUP   (-5)   /* feature outside this backend (PN, RK, WH) */
#define ABIE_HIP_ESTATE   (-6)   /* called before initialize_code / set_state */
#define ABIE_HIP_ETOOBIG  (-7)   /* system larger than the engine supports    */

/* ======================= part 1: reference ABI ========================== */

/* src/common.c:226-275 initialize_code.  deviceID selects the GPU; unlike the
 * reference, deviceID < 0 or an invalid ID is an error (ABIE_HIP_ENODEV):
 * this library has no CPU path, the reference's libabie.so is the CPU path.
 * Like the reference it initialises once per process; later calls return 0. */
int initialize_code(double G, double C, size_t N_MAX, size_t MAX_N_CE,
                    size_t MAX_N_COLLISIONS, int deviceID);

/* src/common.c:382-398 finalize_code.  Frees host and device memory; a later
 * initialize_code starts afresh. */
int finalize_code(void);

/* src/common.c:309-325 set_state / :327-338 get_state (returns N).  N larger
 * than N_MAX grows the buffers (the reference overflows them). */
void set_state(double *pos_vec, double *vel_vec, double *m_vec, double *r_vec,
               size_t N, double G, double C);
int get_state(double *pos_vec, double *vel_vec, double *m_vec, double *r_vec);

/* src/common.c:200-202, 277-307 */
double get_model_time(void);
void set_close_encounter_distance(double d);
double get_close_encounter_distance(void);
void get_close_encounter_buffer(double *buf_ce);         /* 4*MAX_N_CE      */
void get_collision_buffer(double *buf_collision);        /* 4*MAX_N_COLL    */
void reset_close_encounter_buffer(void);
void reset_collision_buffer(void);

/* src/common.c:364-380 set_additional_forces: constant extra acceleration
 * [3N] added after gravity on every evaluation. */
size_t set_additional_forces(size_t N, double *ext_acc);

/* src/common.c:340-357 calculate_energy (massive bodies only), on the GPU. */
double calculate_energy(void);

/* src/common.c:400-403 integrator_gr -> src/integrator_gauss_radau15.c:221-380.
 * GR15 on the GPU from the state of set_state; dt is ignored (as in the
 * reference).  Ends on the first accepted step past t_end.  Returns 0, 1
 * (close encounters), 2 (collisions) or a negative error. */
int integrator_gr(double t, double t_end, double dt);

/* src/integrator_gauss_radau15.c:221 integrator_gauss_radau15: the same GR15
 * on caller arrays (pos, vel updated in place). */
size_t integrator_gauss_radau15(double *pos, double *vel, double *m_vec, double *r_vec,
                                size_t N, double G, double t, double t_end, double dt);

/* src/common.c:82-114 calculate_accelerations: the kernel boundary.  Host AoS
 * in, host AoS out; gravity on the GPU, plus ext-acc when enabled.  C > 0
 * (post-Newtonian terms, src/additional_forces.c) returns ABIE_HIP_EUNSUP. */
size_t calculate_accelerations(const double *pos, const double *vel, size_t N, double G,
                               const double *masses, const double *radii, double *acc);

/* src/common.c:39-77 ode_n_body_second_order and src/gpuforce.cu:129-170
 * ode_n_body_second_order_gpu: gravity only.  Sinks with m < 0 keep the
 * caller's acc values; sources are the bodies with m > 0. */
size_t ode_n_body_second_order(const double *pos, size_t N, double G,
                               const double *masses, const double *radii, double *acc);
size_t ode_n_body_second_order_gpu(const double *pos, size_t N, double G,
                                   const double *masses, const double *radii, double *acc);

/* src/common.c:116-158 check_collisions_close_encounters on the GPU, against
 * the rings and counters of the global simulation. */
size_t check_collisions_close_encounters(const double *pos, const double *radii,
                                         size_t N, double t);

/* src/gpuforce.cu:88-119 gpu_init / :121-126 gpu_finalize. */
void gpu_init(int N, int deviceID);
void gpu_finalize(void);

/* Outside the accelerated path (src/integrator_runge_kutta.c,
 * src/integrator_wisdom_holman.c, src/common.c:3-37): present so that symbol
 * lookups succeed; they return ABIE_HIP_EUNSUP and change nothing. */
int integrator_rk(double t, double t_end, double dt);
int integrator_wh(double t, double t_end, double dt);
/* src/common.h:119,121: void in the reference as well; they only record
 * ABIE_HIP_EUNSUP as the last error (abie_hip_last_error). */
void integrator_runge_kutta(double *pos, double *vel, double *m_vec, double *r_vec, size_t N, double G,
                            double t, double t_end, double dt);
void integrator_wisdom_holman(double *pos, double *vel, double *m_vec, double *r_vec, size_t N, double G,
                              double t, double t_end, double dt);
size_t ode_n_body_first_order(double *pos, size_t N, double G, const double *masses, double *dxdt);

/* ======================= part 2: device entry points ==================== */

/* Force evaluation modes. */
#define ABIE_FORCE_FP64       0  /* rsqrt seed + Taylor correction, FMA accumulate */
#define ABIE_FORCE_FP64_EXACT 1  /* sqrt + divide, reference operation order:
                                    bit-identical to src/common.c:59-71          */
#define ABIE_FORCE_MIXED      2  /* massive->massive FP64; massive->test at
                                    single-precision accuracy (hardware rsqrt
                                    without correction, ~3*2^-24 relative);
                                    FP64 coordinates and accumulation            */
#define ABIE_FORCE_FP32       3  /* massive->test in FP32 arithmetic: rsqrt +
                                    one Newton step, FP64 accumulation (force
                                    entry points abie_hip_accel_soa and
                                    abie_hip_time_accel_soa only)               */

/* Last error message of this thread's most recent failing call. */
const char *abie_hip_last_error(void);

/* Select the device for the calling thread's part-2 calls (hipSetDevice). */
int abie_hip_set_device(int device);

/* Force mode of the part-1 (reference ABI) calls; the default is
 * ABIE_FORCE_FP64, or the ABIE_HIP_FORCE_MODE environment variable
 * (fp64 | exact | mixed) read by initialize_code. */
int abie_hip_set_force_mode(int mode);

/* Accelerations of n_sink sinks from n_src sources, all device pointers.
 * Sources: SoA sx/sy/sz and gm = G*m (m > 0), ascending in the order their
 * contributions are summed.  Sinks: SoA px/py/pz; sinks [0, n_self) are the
 * sources [0, n_self) themselves (their self-pair is skipped), sinks beyond
 * are pure sinks.  Writes ax/ay/az[n_sink].  Launched on `stream`
 * (hipStream_t, 0 = default); returns without synchronising.  When the sinks
 * alone cannot fill the GPU the source range is split and the partial sums
 * go to a stream-ordered allocation (hipMallocAsync on `stream`, freed on it
 * after the launch), so calls on different streams may overlap.
 * Replaces the kernel of src/gpuforce.cu:60-84 (gpuforce_v2). */
int abie_hip_accel_soa(const double *sx, const double *sy, const double *sz,
                       const double *gm, int n_src,
                       const double *px, const double *py, const double *pz,
                       int n_sink, int n_self, double *ax, double *ay, double *az,
                       int mode, void *stream);

/* abie_hip_accel_soa for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  Same arguments, kernels and
 * bits; it calls no hipMalloc, hipFree or allocating library routine.  Its
 * source-split scratch is a 16 MiB device array of the library, one per GPU,
 * held until the process ends (a plan needing more returns ABIE_HIP_ETOOBIG;
 * on a 256-CU MI355X no plan needs more than 12 MiB).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (concurrent streams
 * or threads) race and must be serialised by the caller. */
int abie_hip_accel_soa_noalloc(const double *sx, const double *sy, const double *sz,
                               const double *gm, int n_src,
                               const double *px, const double *py, const double *pz,
                               int n_sink, int n_self, double *ax, double *ay, double *az,
                               int mode, void *stream);

/* Times `reps` back-to-back launches of abie_hip_accel_soa on `stream` with
 * HIP events recorded on that stream; *ms_per_launch receives the mean.
 * When ms_each is not NULL a second pass brackets each launch with its own
 * event pair and *ms_each receives the mean single-launch duration, and
 * *ms_each_median (when not NULL) the median.  When shader_ghz is not NULL,
 * `reps` back-to-back launches of an instance of the same kernel that reads
 * each wave's shader cycle counter and wall clock at its start and end
 * follow, and *shader_ghz receives the median over the last min(reps, 8)
 * launches of the median over waves of cycles per wall second: the clock
 * the kernel settles at under that run (plans of the scalar-source kernel
 * with one sink per thread and no source split; 0 for other plans). */
int abie_hip_time_accel_soa(const double *sx, const double *sy, const double *sz,
                            const double *gm, int n_src,
                            const double *px, const double *py, const double *pz,
                            int n_sink, int n_self, double *ax, double *ay, double *az,
                            int mode, void *stream, int reps, double *ms_per_launch,
                            double *ms_each, double *ms_each_median, double *shader_ghz);

/* Measurement aid, no counterpart in the reference: the shader clock and
 * the SIMD cycles per wave instruction of FP64 VALU work on the current
 * device.  A kernel keeps 8 waves per SIMD issuing independent chains of
 * kind 0 (v_fma_f64), 1 (v_rsq_f64) or 2 (one v_rsq_f64 per 16 v_fma_f64)
 * for about `milliseconds` (0 < ms <= 10000).  *shader_ghz receives the
 * median over waves of shader cycles (s_memtime) per second of the
 * constant-rate wall clock (s_memrealtime); *cycles_per_instruction the
 * kernel's event-timed duration at that clock over the wave instructions
 * each SIMD issued.  Allocates and frees its own small buffers; runs on and
 * synchronises the null stream. */
int abie_hip_issue_rate(int kind, double milliseconds, double *shader_ghz, double *cycles_per_instruction);

/* Check aid, no counterpart in the reference: the seventh root that the
 * massive-only step controllers take of estim / epsb (the reference's
 * pow(x, 1./7), integrator_gauss_radau15.c:316, correctly rounded, stays
 * with the plain controller), evaluated on the host with the same code as
 * on the device.  x >= 0; returns x for 0, +inf and NaN. */
double abie_hip_step_root7(double x);

/* ---- batched GR15 ensemble: many independent systems, one launch ---- */
typedef struct abie_ensemble abie_ensemble;

/* Controllers of the step size. */
#define ABIE_CTRL_PLAIN    0  /* the reference's controller over all components */
#define ABIE_CTRL_MASSIVE  1  /* PC test and error estimate over massive bodies
                                 only (FP64 components; used with MIXED forces) */
#define ABIE_CTRL_LEARNED  2  /* MASSIVE + learned step proposer, gated by the
                                 GR15 error test with the plain rule as fallback */

/* Create an ensemble of n_sys systems.  Bodies of all systems are
 * concatenated: system k owns bodies [offsets[k], offsets[k+1]) of the host
 * AoS arrays pos/vel [3*total], m/r [total].  G is shared. */
int abie_ensemble_create(abie_ensemble **out, int n_sys, const int *offsets,
                         const double *pos, const double *vel, const double *m,
                         const double *r, double G, int force_mode, int controller);
void abie_ensemble_destroy(abie_ensemble *e);

/* Integrate every system to t_end (each ends on its first accepted step past
 * t_end, as the reference does).  State persists across calls: a second call
 * continues the same step sequence (no restart).  max_seconds bounds one call
 * (0 = no bound); *done receives 1 when every system has reached t_end. */
int abie_ensemble_run(abie_ensemble *e, double t_end, double max_seconds, int *done);

/* Copy state back (host AoS, same layout as create), per-system model time
 * t[n_sys] and status[n_sys] (0, 1, 2 or negative error); any may be NULL. */
int abie_ensemble_get_state(abie_ensemble *e, double *pos, double *vel, double *t, int *status);

/* Per-system counters: [n_sys][6] = evals, attempts, accepted, pc_iters,
 * proposer_used, proposer_rejected. */
int abie_ensemble_get_counters(abie_ensemble *e, long long *counters);

/* Per-system finish times of the last abie_ensemble_run, [n_sys]: seconds of
 * the device's constant-rate clock from the first start of any unit to the
 * moment the system's last unit stopped (-1 while a unit is still running). */
int abie_ensemble_get_finish_times(abie_ensemble *e, double *seconds);

/* Energies (common.c:340-357 per system) of the current state, [n_sys]. */
int abie_ensemble_energy(abie_ensemble *e, double *energy);

/* Attach / detach a step log: every attempt of every system (its primary
 * slice) appends one record of ABIE_STEPLOG_FIELDS doubles while capacity
 * lasts: system, t, h, estimate, err, dtreq, accepted, PC iterations,
 * rejections before, previous h / err / dtreq / PC iterations, massive
 * bodies, test particles, next h, and for accepted attempts the largest
 * |v_ij|^2 / r_ij^2 over massive pairs of the new state, and the system's
 * mass ratio (sum of the other massive bodies) / (largest mass); then
 * (fields 18..30) the learned proposer's 12 inputs and its raw output as the
 * kernel evaluated them, NaN when the attempt made no proposal. */
#define ABIE_STEPLOG_FIELDS 31
int abie_ensemble_set_steplog(abie_ensemble *e, long long capacity);
int abie_ensemble_get_steplog(abie_ensemble *e, double *records, long long *n_records);

/* Load the learned step proposer's weights (FP32, layout of abie_mlp_*). */
int abie_ensemble_set_proposer(abie_ensemble *e, const float *weights, int n_weights);

/* Engine configurations (threads per group, bodies per thread) and the
 * rules that map a system size to one: a system of n bodies runs with the
 * first rule whose max_bodies >= n and whose configuration holds it.  The
 * rules come from the autotuner (abie_amd/autotune.py). */
int abie_gr15_num_configs(void);
int abie_gr15_config(int cfg, int *threads, int *bodies_per_thread);
int abie_gr15_set_tuning(int n_rules, const int *max_bodies, const int *cfg);

/* ---- learned step proposer: MLP 12 -> 64 -> 64 -> 1, FP32 ---- */
#define ABIE_MLP_IN 12
#define ABIE_MLP_HIDDEN 64
#define ABIE_MLP_NPARAM (ABIE_MLP_IN * ABIE_MLP_HIDDEN + ABIE_MLP_HIDDEN + \
                         ABIE_MLP_HIDDEN * ABIE_MLP_HIDDEN + ABIE_MLP_HIDDEN + ABIE_MLP_HIDDEN + 1)

/* Forward pass on device pointers: x[n][12] -> y[n]. */
int abie_mlp_forward(const float *w, const float *x, float *y, int n, void *stream);

/* Loss and gradient: mean squared error of y(x) against target[n] with
 * per-sample weights sw[n] (may be NULL); grad[ABIE_MLP_NPARAM], *loss.  Its
 * workspace is a stream-ordered allocation on `stream`. */
int abie_mlp_loss_grad(const float *w, const float *x, const float *target, const float *sw,
                       int n, float *grad, float *loss, void *stream);

/* One Adam step on device weights w with gradient g; m, v moment buffers. */
int abie_mlp_adam(float *w, const float *g, float *m, float *v, int n_param, int step,
                  float lr, float beta1, float beta2, float eps, void *stream);

#ifdef __cplusplus
}
#endif
#endif
