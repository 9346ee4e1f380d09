/*
 * abie_oracle.c -- CPU restatement of ABIE's force path and GR15 driver.
 *
 * TEST INFRASTRUCTURE ONLY (see abie_oracle.h).  Written from the reference's
 * behaviour; each function cites the file:line it follows.  The operation
 * order is the reference's, so that results are bit-identical to the
 * reference C built with gcc -O3 -std=c99 (no FMA contraction).
 */
#include "abie_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "radau_tables.h"

/* integrator_gauss_radau15.h:17,58-63 */
static const double ORC_H_MIN = 1.e-13;
static const double ORC_EPSB = 1.e-4;
static const double ORC_FAC = 0.25;
static const double ORC_TOLPC = 1.e-16;
#define ORC_NH 8

const double *orc_hs(void) { return ORC_HS; }
const double *orc_rs(void) { return &ORC_RS[0][0]; }
const double *orc_cs(void) { return &ORC_CS[0][0]; }

double orc_div_const(double x, double c, double rc)
{
    double q = x * rc;
    double r = fma(q, c, -x);   /* exact residual, negated */
    return fma(-r, rc, q);
}

/* common.c:39-77.  Sink j (m_j >= 0) sums over sources k != j with m_k > 0,
 * ascending k, subtracting G m_k dx / (sqrt(r2) r2).  Deleted sinks (m < 0)
 * are left untouched. */
void orc_accel(const double *pos, size_t n, double G, const double *m, double *acc)
{
    orc_accel_range(pos, n, G, m, acc, 0, n);
}

/* The same loop for the sinks [j0, j1) only (every source still counts), so
 * that sink ranges can run on separate host threads; per sink the arithmetic
 * is identical to orc_accel. */
void orc_accel_range(const double *pos, size_t n, double G, const double *m, double *acc, size_t j0, size_t j1)
{
    for (size_t j = j0; j < j1 && j < n; j++) {
        if (m[j] < 0.0) continue;
        const double xj = pos[3 * j], yj = pos[3 * j + 1], zj = pos[3 * j + 2];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (size_t k = 0; k < n; k++) {
            if (k == j || m[k] <= 0.0) continue;
            const double gm = G * m[k];
            const double dx = xj - pos[3 * k];
            const double dy = yj - pos[3 * k + 1];
            const double dz = zj - pos[3 * k + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double r3 = sqrt(r2) * r2;
            ax -= gm * dx / r3;
            ay -= gm * dy / r3;
            az -= gm * dz / r3;
        }
        acc[3 * j] = ax;
        acc[3 * j + 1] = ay;
        acc[3 * j + 2] = az;
    }
}

/* common.c:82-114, CPU branch.  Post-Newtonian terms (C > 0,
 * additional_forces.c:8-10) are outside the restated path: rejected. */
int orc_calculate_accelerations(orc_sim *s, const double *pos, const double *vel, double *acc)
{
    (void)vel;
    if (s->C > 0.0) return -1;
    orc_accel(pos, s->n, s->G, s->m, acc);
    s->n_evals++;
    if (s->ext_acc_on && s->ext_acc != NULL)
        for (size_t i = 0; i < 3 * s->n; i++) acc[i] += s->ext_acc[i];
    return 0;
}

/* common.c:116-158.  All ordered pairs (j, k), j != k, test particles
 * included; events are recorded for j < k into 4-double rings. */
int orc_check_events(orc_sim *s, const double *pos, double t)
{
    const size_t n = s->n;
    for (size_t j = 0; j < n; j++) {
        const double x = pos[3 * j], y = pos[3 * j + 1], z = pos[3 * j + 2];
        for (size_t k = 0; k < n; k++) {
            if (j == k) continue;
            const double dx = x - pos[3 * k];
            const double dy = y - pos[3 * k + 1];
            const double dz = z - pos[3 * k + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double d = sqrt(r2);
            const double rr = s->r[j] + s->r[k];
            if (d <= s->ce_dist && j < k) {
                if (s->buf_ce != NULL) {
                    const size_t w = 4 * s->max_ce, o = 4 * s->n_ce;
                    s->buf_ce[o % w] = t;
                    s->buf_ce[(o + 1) % w] = (double)j;
                    s->buf_ce[(o + 2) % w] = (double)k;
                    s->buf_ce[(o + 3) % w] = d;
                }
                s->n_ce += 1;
            }
            if (rr > 0 && d <= rr && j < k) {
                if (s->buf_coll != NULL) {
                    const size_t w = 4 * s->max_coll, o = 4 * s->n_coll;
                    s->buf_coll[o % w] = t;
                    s->buf_coll[(o + 1) % w] = (double)j;
                    s->buf_coll[(o + 2) % w] = (double)k;
                    s->buf_coll[(o + 3) % w] = d;
                }
                s->n_coll += 1;
            }
        }
    }
    if (s->max_ce > 0 && s->n_ce >= s->max_ce) return 1;
    if (s->max_coll > 0 && s->n_coll >= s->max_coll) return 2;
    return 0;
}

/* common.c:192-198: the running max starts at -DBL_MAX; NaNs never win. */
double orc_vector_max_abs(const double *v, size_t n)
{
    double mx = -DBL_MAX;
    for (size_t i = 0; i < n; i++)
        if (mx < fabs(v[i])) mx = fabs(v[i]);
    return mx;
}

/* common.c:204-208 */
static double orc_norm3(const double *v)
{
    double sum = 0.0;
    for (size_t i = 0; i < 3; i++) sum += (v[i] * v[i]);
    return sqrt(sum);
}

/* common.c:340-357.  Bodies with m == 0 are skipped on both sides; the sum is
 * one running total, kinetic term first, then the pair terms of body i. */
double orc_energy(const orc_sim *s)
{
    double e = 0.0, d[3];
    for (size_t i = 0; i < s->n; i++) {
        if (s->m[i] == 0) continue;
        e += (0.5 * s->m[i] * pow(orc_norm3(&s->vel[3 * i]), 2.0));
        for (size_t j = 0; j < s->n; j++) {
            if (i == j || s->m[j] == 0) continue;
            d[0] = s->pos[3 * i] - s->pos[3 * j];
            d[1] = s->pos[3 * i + 1] - s->pos[3 * j + 1];
            d[2] = s->pos[3 * i + 2] - s->pos[3 * j + 2];
            e -= (0.5 * s->G * s->m[i] * s->m[j] / orc_norm3(d));
        }
    }
    return e;
}

static double orc_pow(const orc_sim *s, double x, double y)
{
    return s->pow_fn ? s->pow_fn(x, y) : pow(x, y);
}

/* calculate_accelerations as GR15 calls it.  A deleted body (m < 0) is no
 * sink: ode_n_body_second_order leaves its entries unwritten
 * (common.c:48-52), and inside the reference's GR15 those entries are
 * uninitialised stack arrays (integrator_gauss_radau15.c:234-238), to which
 * the external acceleration is then added (common.c:107-111).  Here, as on
 * the device, they are 0 before the external term: a deleted body feels
 * only the external acceleration. */
static void gr15_accel(orc_sim *s, const double *pos, const double *vel, double *acc)
{
    for (size_t j = 0; j < s->n; j++)
        if (s->m[j] < 0.0) acc[3 * j] = acc[3 * j + 1] = acc[3 * j + 2] = 0.0;
    orc_calculate_accelerations(s, pos, vel, acc);
}

/* integrator_gauss_radau15.c:9-54.  Note the two max-abs calls over only the
 * first n (not 3n) components (:23-24), kept as in the reference. */
double orc_initial_time_step(orc_sim *s, const double *y0, const double *dy0)
{
    const size_t n = s->n, dim = 3 * n;
    const int p = 15;
    double *f0 = calloc(dim ? dim : 1, sizeof(double));
    double *f1 = calloc(dim ? dim : 1, sizeof(double));
    double *y1 = calloc(dim ? dim : 1, sizeof(double));
    double *dy1 = calloc(dim ? dim : 1, sizeof(double));
    double d0, d1, d2, dt0, dt1;

    gr15_accel(s, y0, dy0, f0);
    d0 = orc_vector_max_abs(y0, n);
    d1 = orc_vector_max_abs(f0, n);
    if (d0 < 1e-5 || d1 < 1e-5) dt0 = 1e-6;
    else dt0 = 0.01 * (d0 / d1);

    for (size_t i = 0; i < dim; i++) {
        y1[i] = y0[i] + dt0 * dy0[i];
        dy1[i] = dy0[i] + dt0 * f0[i];
    }
    gr15_accel(s, y1, dy1, f1);
    d2 = -DBL_MAX;
    for (size_t i = 0; i < dim; i++)
        if (d2 < fabs(f1[i] - f0[i])) d2 = fabs(f1[i] - f0[i]);
    d2 = d2 / dt0;

    if (fmax(d1, d2) <= 1e-15) dt1 = fmax(1e-6, dt0 * 1e-3);
    else dt1 = orc_pow(s, 0.01 / fmax(d1, d2), (1.0 / (p + 1)));

    free(f0); free(f1); free(y1); free(dy1);
    return fmin(100 * dt0, dt1);
}

/* integrator_gauss_radau15.c:56-61.  s = node, dt = step size. */
static void approx_pos(const double *y1, const double *dy1, const double *F1, double s,
                       const double *b, size_t dim, double dt, double *y)
{
    const double *b0 = b, *b1 = b + dim, *b2 = b + 2 * dim, *b3 = b + 3 * dim;
    const double *b4 = b + 4 * dim, *b5 = b + 5 * dim, *b6 = b + 6 * dim;
    for (size_t i = 0; i < dim; i++) {
        double p = b5[i] / 0.28e2 + s * b6[i] / 0.36e2;
        p = b4[i] / 0.21e2 + s * p;
        p = b3[i] / 0.15e2 + s * p;
        p = b2[i] / 0.10e2 + s * p;
        p = b1[i] / 0.6e1 + s * p;
        p = b0[i] / 0.3e1 + s * p;
        p = F1[i] + s * p;
        y[i] = y1[i] + dt * s * (dy1[i] + dt * s * p / 0.2e1);
    }
}

/* integrator_gauss_radau15.c:63-68 */
static void approx_vel(const double *dy1, const double *F1, double s, const double *b,
                       size_t dim, double dt, double *dy)
{
    const double *b0 = b, *b1 = b + dim, *b2 = b + 2 * dim, *b3 = b + 3 * dim;
    const double *b4 = b + 4 * dim, *b5 = b + 5 * dim, *b6 = b + 6 * dim;
    for (size_t i = 0; i < dim; i++) {
        double p = b5[i] / 0.7e1 + s * b6[i] / 0.8e1;
        p = b4[i] / 0.6e1 + s * p;
        p = b3[i] / 0.5e1 + s * p;
        p = b2[i] / 0.4e1 + s * p;
        p = b1[i] / 0.3e1 + s * p;
        p = b0[i] / 0.2e1 + s * p;
        p = F1[i] + s * p;
        dy[i] = dy1[i] + dt * s * p;
    }
}

/* integrator_gauss_radau15.c:70-121: at node ih (1..7) every g[0..ih-1] is
 * rebuilt from the stored node accelerations by the divided-difference
 * recurrence g_k = (((F_{k+1}-F_0) r_{k+1,0} - g_0) r_{k+1,1} - ... ) r_{k+1,k}. */
static void compute_gs(const double *ddys, int ih, size_t dim, double *g)
{
    const double *F0 = ddys;
    for (size_t i = 0; i < dim; i++) {
        for (int k = 0; k < ih; k++) {
            double v = (ddys[(size_t)(k + 1) * dim + i] - F0[i]) * ORC_RS[k + 1][0];
            for (int l = 1; l <= k; l++)
                v = (v - g[(size_t)(l - 1) * dim + i]) * ORC_RS[k + 1][l];
            g[(size_t)k * dim + i] = v;
        }
    }
}

/* integrator_gauss_radau15.c:123-175: b_k = sum_{j=k..6} c_{j,k} g_j,
 * left to right, for k < ih. */
static void compute_bs_from_gs(const double *g, int ih, size_t dim, double *b)
{
    for (size_t i = 0; i < dim; i++) {
        for (int k = 0; k < ih; k++) {
            double v = ORC_CS[k][k] * g[(size_t)k * dim + i];
            for (int j = k + 1; j < 7; j++)
                v = v + ORC_CS[j][k] * g[(size_t)j * dim + i];
            b[(size_t)k * dim + i] = v;
        }
    }
}

/* integrator_gauss_radau15.c:178-218.  The `bd` correction is never applied
 * in the reference (its `inited` flag is never set), so b <- E + 0.0. */
static void refine_bs(double *b, double q, double *E, size_t dim)
{
    const double q2 = q * q, q3 = q2 * q, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q2 * q5;
    double *b0 = b, *b1 = b + dim, *b2 = b + 2 * dim, *b3 = b + 3 * dim;
    double *b4 = b + 4 * dim, *b5 = b + 5 * dim, *b6 = b + 6 * dim;
    for (size_t i = 0; i < dim; i++) {
        E[i] = q * (b6[i] * 7.0 + b5[i] * 6.0 + b4[i] * 5.0 + b3[i] * 4.0 + b2[i] * 3.0 + b1[i] * 2.0 + b0[i]);
        E[dim + i] = q2 * (b6[i] * 21.0 + b5[i] * 15.0 + b4[i] * 10.0 + b3[i] * 6.0 + b2[i] * 3.0 + b1[i]);
        E[2 * dim + i] = q3 * (b6[i] * 35.0 + b5[i] * 20.0 + b4[i] * 10.0 + b3[i] * 4.0 + b2[i]);
        E[3 * dim + i] = q4 * (b6[i] * 35.0 + b5[i] * 15.0 + b4[i] * 5.0 + b3[i]);
        E[4 * dim + i] = q5 * (b6[i] * 21.0 + b5[i] * 6.0 + b4[i]);
        E[5 * dim + i] = q6 * (b6[i] * 7.0 + b5[i]);
        E[6 * dim + i] = q7 * b6[i];
    }
    for (size_t i = 0; i < 7 * dim; i++) b[i] = E[i] + 0.0;
}

/* integrator_gauss_radau15.c:221-380 (driven as common.c:400-403). */
int orc_gr15(orc_sim *s, double t_start, double t_end, double dt_user)
{
    (void)dt_user;  /* ignored by the reference: h comes from initial_time_step */
    const size_t n = s->n, dim = 3 * n;
    if (n == 0) return -1;
    if (s->C > 0.0) return -1;
    double *y0 = malloc(dim * sizeof(double)), *dy0 = malloc(dim * sizeof(double));
    double *ddy0 = calloc(dim, sizeof(double)), *y = calloc(dim, sizeof(double));
    double *dy = calloc(dim, sizeof(double)), *db6 = calloc(dim, sizeof(double));
    double *ddy = calloc(dim, sizeof(double));
    double *bs0 = calloc(7 * dim, sizeof(double)), *bs = calloc(7 * dim, sizeof(double));
    double *g = calloc(7 * dim, sizeof(double)), *E = calloc(7 * dim, sizeof(double));
    double *ddys = calloc(8 * dim, sizeof(double));
    double h, t = t_start;
    int flag = 0;

    memcpy(y0, s->pos, dim * sizeof(double));
    memcpy(dy0, s->vel, dim * sizeof(double));

    gr15_accel(s, y0, dy0, ddy0);
    h = orc_initial_time_step(s, y0, dy0);

    int advance = 1, rejects = 0, warned = 0;
    while (advance) {
        int pc_used = 0;
        for (int ipc = 0; ipc < 12; ipc++) {
            pc_used = ipc + 1;
            memset(ddys, 0, 8 * dim * sizeof(double));
            for (int ih = 0; ih < ORC_NH; ih++) {
                approx_pos(y0, dy0, ddy0, ORC_HS[ih], bs, dim, h, y);
                approx_vel(dy0, ddy0, ORC_HS[ih], bs, dim, h, dy);
                if (ih == 0 && s->skip_first_node)
                    memcpy(ddys, ddy0, dim * sizeof(double));
                else
                    gr15_accel(s, y, dy, ddys + (size_t)ih * dim);
                compute_gs(ddys, ih, dim, g);
                compute_bs_from_gs(g, ih, dim, bs);
            }
            for (size_t i = 0; i < dim; i++) db6[i] = bs[6 * dim + i] - bs0[6 * dim + i];
            if (orc_vector_max_abs(db6, dim) / orc_vector_max_abs(ddys + 7 * dim, dim) < ORC_TOLPC)
                break;
            memcpy(bs0, bs, 7 * dim * sizeof(double));
        }
        s->n_pc_iters += pc_used;
        s->n_attempts++;

        approx_pos(y0, dy0, ddy0, 1., bs, dim, h, y);
        approx_vel(dy0, ddy0, 1., bs, dim, h, dy);
        gr15_accel(s, y, dy, ddy);

        const double estim_b6 = orc_vector_max_abs(bs + 6 * dim, dim) / orc_vector_max_abs(ddy, dim);
        const double err = orc_pow(s, estim_b6 / ORC_EPSB, 1. / 7);
        const double dtreq = h / err;
        const double t_attempt = t, h_attempt = h;
        int accepted = 0;

        if (err <= 1 || rejects > 100) {
            accepted = 1;
            t += h;
            s->t = t;
            rejects = 0;
            if (t > t_end) {
                h = t_end - t;
                advance = 0;
            }
            memcpy(y0, y, dim * sizeof(double));
            memcpy(dy0, dy, dim * sizeof(double));
            memcpy(ddy0, ddy, dim * sizeof(double));
            memcpy(bs0, bs, 7 * dim * sizeof(double));
            refine_bs(bs, dtreq / h, E, dim);
            s->n_accepted++;
            flag = orc_check_events(s, y, t);
        } else {
            rejects += 1;
        }
        if (s->log != NULL && s->log_len < s->log_cap) {
            orc_attempt *a = &s->log[s->log_len++];
            a->t = t_attempt; a->h = h_attempt; a->estim_b6 = estim_b6; a->err = err;
            a->dtreq = dtreq; a->accepted = accepted; a->pc_iters = pc_used;
        }
        if (accepted && flag > 0) break;

        if (dtreq / h > 1.0 / ORC_FAC) h /= ORC_FAC;
        else if (dtreq < 1.e-12) h *= ORC_FAC;
        else h = dtreq;
        if (h < ORC_H_MIN) {
            if (!warned) {
                fprintf(stderr, "oracle: time step %g below h_min, clamped (t = %g)\n", h, t);
                warned = 1;
            }
            s->n_hmin++;
            h = ORC_H_MIN;
        }
    }

    memcpy(s->pos, y0, dim * sizeof(double));
    memcpy(s->vel, dy, dim * sizeof(double));
    free(y0); free(dy0); free(ddy0); free(y); free(dy); free(db6); free(ddy);
    free(bs0); free(bs); free(g); free(E); free(ddys);
    return flag;
}
