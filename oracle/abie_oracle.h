/*
 * abie_oracle.h -- CPU restatement of ABIE's all-pairs gravity path and the
 * Gauss-Radau15 driver that calls it.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and bench.py's CPU
 * baseline leg may load this library, and only as the checker (or, for the
 * baseline, as the thing timed on the host).  The product (libabie_hip.so)
 * never links or calls it.
 *
 * Every function names the reference file:line it restates.  Operation order
 * follows the reference exactly (same products, same left-to-right sums, no
 * FMA contraction), so results are bit-identical to the reference C compiled
 * with gcc -O3 -std=c99 on x86-64.  Deliberate, documented differences:
 *   - state lives in an explicit orc_sim instead of the global `sim`;
 *   - GR15 work arrays are heap-allocated and zero-initialised (the reference
 *     uses stack VLAs and, for deleted bodies (m < 0), reads uninitialised
 *     malloc memory; here their acceleration is 0);
 *   - instrumentation counters and an optional per-attempt log.
 */
#ifndef ABIE_ORACLE_H
#define ABIE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One GR15 attempt, as logged by orc_gr15 when a log buffer is attached. */
typedef struct orc_attempt {
    double t;         /* model time at the start of the attempt            */
    double h;         /* step size attempted                               */
    double estim_b6;  /* max|b7| / max|a| (integrator_gauss_radau15.c:315) */
    double err;       /* (estim_b6 / epsb)^(1/7)                           */
    double dtreq;     /* h / err                                           */
    int accepted;     /* 1 if the step was accepted                        */
    int pc_iters;     /* predictor-corrector iterations used (1..12)       */
} orc_attempt;

typedef struct orc_sim {
    size_t n;                 /* number of bodies                           */
    double G, C;              /* gravitational constant, speed of light     */
    double *pos, *vel;        /* AoS [3n], caller-owned, updated in place   */
    double *m, *r;            /* masses, radii [n]                          */
    const double *ext_acc;    /* [3n] or NULL                               */
    int ext_acc_on;
    double t;                 /* model time (sim.t_global)                  */
    double ce_dist;           /* close-encounter distance                   */
    size_t max_ce, max_coll;  /* ring sizes (events)                        */
    size_t n_ce, n_coll;      /* event counters                             */
    double *buf_ce, *buf_coll;/* [4*max] rings, or NULL                     */
    /* instrumentation */
    long long n_evals, n_attempts, n_accepted, n_pc_iters, n_hmin;
    orc_attempt *log;         /* optional attempt log                       */
    size_t log_cap, log_len;
    int skip_first_node;      /* 1: reuse a(y0) for node 0 instead of a new
                                 evaluation (bit-identical; saves 1 of 8)   */
    double (*pow_fn)(double, double); /* NULL: libm pow (as the reference);
                                 else used for the step-size exponents      */
} orc_sim;

/* common.c:39-77  ode_n_body_second_order */
void orc_accel(const double *pos, size_t n, double G, const double *m, double *acc);
/* orc_accel for the sinks [j0, j1) (host-thread sharding of the CPU baseline). */
void orc_accel_range(const double *pos, size_t n, double G, const double *m, double *acc, size_t j0, size_t j1);

/* common.c:82-114 calculate_accelerations (CPU branch; C>0 is rejected) */
int orc_calculate_accelerations(orc_sim *s, const double *pos, const double *vel, double *acc);

/* common.c:116-158 check_collisions_close_encounters */
int orc_check_events(orc_sim *s, const double *pos, double t);

/* common.c:340-357 calculate_energy */
double orc_energy(const orc_sim *s);

/* common.c:192-198 vector_max_abs */
double orc_vector_max_abs(const double *v, size_t n);

/* integrator_gauss_radau15.c:9-54 initial_time_step */
double orc_initial_time_step(orc_sim *s, const double *y0, const double *dy0);

/* integrator_gauss_radau15.c:221-380 integrator_gauss_radau15, driven the way
 * common.c:400-403 integrator_gr drives it.  Returns 0, 1 (close encounters)
 * or 2 (collisions); -1 on bad input. */
int orc_gr15(orc_sim *s, double t, double t_end, double dt);

/* Correctly rounded x / c for a constant c with rc = RN(1/c), by Markstein's
 * residual correction; exposed so the test-suite can check it exhaustively. */
double orc_div_const(double x, double c, double rc);

/* Tables, exposed for the constant checks. */
const double *orc_hs(void);
const double *orc_rs(void);
const double *orc_cs(void);

#ifdef __cplusplus
}
#endif
#endif
