/*
 * hfv_oracle.c -- CPU oracle for the MUSCL-Hancock finite-volume path (and
 * the first-order Godunov and random-choice schemes that share its parts).
 *
 * TEST INFRASTRUCTURE ONLY (see hfv_oracle.h).  Plain C restatement of the
 * reference algorithm; each function cites the reference file:line whose
 * arithmetic order it keeps.  Compile with -ffp-contract=off.
 *
 * Reference files (hydroFV, src/):
 *   slope_limiter.c, utils.c, riemann_solver.c, riemann_solver_hllc.c,
 *   riemann_solver_exact.c, system.c, boundary.c, source_term.c,
 *   integrator_muscl_hancock.c, integrator_godunov_first_order.c,
 *   integrator_random_choice_1d.c
 */
#include "hfv_oracle.h"

#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ======================================================================== */
/* Slope limiters: slope_limiter.c:49-91                                    */
/* ======================================================================== */
double hfo_limit(int limiter, double a, double b)
{
    if (a * b <= 0.0) {
        return 0.0;
    }
    switch (limiter) {
    case HFO_MINMOD: /* slope_limiter.c:49-63 -- ties return b */
        return (fabs(a) < fabs(b)) ? a : b;
    case HFO_VAN_LEER: /* slope_limiter.c:65-75 */
        return 2.0 * a * b / (a + b + 1e-10);
    case HFO_MC: /* slope_limiter.c:77-91 -- sign taken from a > 0 */
        if (a > 0) {
            return fmin(0.5 * (a + b), 2.0 * fmin(fabs(a), fabs(b)));
        }
        return -fmin(-0.5 * (a + b), 2.0 * fmin(fabs(a), fabs(b)));
    default:
        return NAN;
    }
}

void hfo_limit_batch(int limiter, int64_t n, const double *a, const double *b, double *out)
{
    for (int64_t k = 0; k < n; k++) {
        out[k] = hfo_limit(limiter, a[k], b[k]);
    }
}

/* utils.c:23-26 */
double hfo_sound_speed(double gamma, double rho, double p)
{
    return sqrt(gamma * p / rho);
}

/* ======================================================================== */
/* Pressure guess: riemann_solver.c:144-224                                 */
/* ======================================================================== */
static double coeff_A(double gamma, double rho) { return 2.0 / ((gamma + 1.0) * rho); }   /* :144-147 */
static double coeff_B(double gamma, double p) { return p * ((gamma - 1.0) / (gamma + 1.0)); } /* :149-152 */

double hfo_guess_p(double gamma, double rho_l, double u_l, double p_l, double a_l,
                   double rho_r, double u_r, double p_r, double a_r, double tol)
{
    const double p_lo = (p_l > p_r) ? p_r : p_l;
    const double p_hi = (p_l > p_r) ? p_l : p_r;
    const double ppv = 0.5 * (p_l + p_r) - 0.125 * (u_r - u_l) * (rho_l + rho_r) * (a_l + a_r);
    double guess;
    if (p_hi / p_lo <= 2.0 && (p_lo <= ppv && ppv <= p_hi)) {
        guess = ppv; /* PVRS */
    } else if (ppv < p_lo) {
        /* two-rarefaction */
        const double h = 0.5 * (gamma - 1.0);
        guess = (a_l + a_r - h * (u_r - u_l)) /
                (a_l / (pow(p_l, h / gamma)) + a_r / (pow(p_r, h / gamma)));
        guess = pow(guess, gamma / h);
    } else {
        /* two-shock with the PVRS estimate */
        const double gl = sqrt(coeff_A(gamma, rho_l) / (ppv + coeff_B(gamma, p_l)));
        const double gr = sqrt(coeff_A(gamma, rho_r) / (ppv + coeff_B(gamma, p_r)));
        guess = (gl * p_l + gr * p_r - (u_r - u_l)) / (gl + gr);
    }
    if (guess < tol) {
        guess = tol;
    }
    return guess;
}

/* ======================================================================== */
/* Vacuum sampling: riemann_solver.c:239-733                                */
/* The 2D variants carry the transverse velocity (index 2 of a 4-state).    */
/* ======================================================================== */
/* riemann_solver.c:239-286 / 483-535: rarefaction into vacuum on the right */
static void vac_right(double gamma, const double *w, int nf, double a, double speed, double *sol)
{
    const double rho = w[0], u = w[1], p = w[nf - 1];
    if (speed <= (u - a)) {
        for (int k = 0; k < nf; k++) sol[k] = w[k];
        return;
    }
    const double front = u + 2.0 * a / (gamma - 1.0);
    if (speed < front) {
        double t = (2.0 / (gamma + 1.0) + (u - speed) * (gamma - 1.0) / ((gamma + 1.0) * a));
        t = pow(t, 2.0 / (gamma - 1.0));
        sol[0] = rho * t;
        sol[1] = (a + 0.5 * (gamma - 1.0) * u + speed) * 2.0 / (gamma + 1.0);
        if (nf == 4) sol[2] = w[2];
        sol[nf - 1] = p * pow(t, gamma);
    } else {
        for (int k = 0; k < nf; k++) sol[k] = 0.0;
    }
}

/* riemann_solver.c:301-348 / 550-602: rarefaction into vacuum on the left */
static void vac_left(double gamma, const double *w, int nf, double a, double speed, double *sol)
{
    const double rho = w[0], u = w[1], p = w[nf - 1];
    if (speed >= (u + a)) {
        for (int k = 0; k < nf; k++) sol[k] = w[k];
        return;
    }
    const double front = u - 2.0 * a / (gamma - 1.0);
    if (speed > front) {
        double t = (2.0 / (gamma + 1.0) - (u - speed) * (gamma - 1.0) / ((gamma + 1.0) * a));
        t = pow(t, 2.0 / (gamma - 1.0));
        sol[0] = rho * t;
        sol[1] = (-a + 0.5 * (gamma - 1.0) * u + speed) * 2.0 / (gamma + 1.0);
        if (nf == 4) sol[2] = w[2];
        sol[nf - 1] = p * pow(t, gamma);
    } else {
        for (int k = 0; k < nf; k++) sol[k] = 0.0;
    }
}

/* riemann_solver.c:409-468 / 667-733.  The reference hands the vacuum side's
 * own state to the sampler when one side is vacuum (435-446), and measures
 * the vacuum-generation fronts without the velocities (384-385); both kept. */
static void vacuum(double gamma, const double *wl, double a_l, const double *wr, double a_r,
                   int nf, double speed, double *sol)
{
    const double rho_l = wl[0], rho_r = wr[0];
    if (rho_l <= 0.0 && rho_r <= 0.0) {
        for (int k = 0; k < nf; k++) sol[k] = 0.0;
    } else if (rho_l <= 0.0) {
        vac_right(gamma, wl, nf, a_l, speed, sol);
    } else if (rho_r <= 0.0) {
        vac_left(gamma, wr, nf, a_r, speed, sol);
    } else {
        const double front_l = 2.0 * a_l / (gamma - 1.0);
        const double front_r = 2.0 * a_r / (gamma - 1.0);
        if (speed <= front_l) {
            vac_right(gamma, wl, nf, a_l, speed, sol);
        } else if (front_r <= speed) {
            vac_left(gamma, wr, nf, a_r, speed, sol);
        } else {
            for (int k = 0; k < nf; k++) sol[k] = 0.0;
        }
    }
}

void hfo_vacuum_1d(double gamma, const double wl[3], double a_l, const double wr[3],
                   double a_r, double speed, double sol[3])
{
    vacuum(gamma, wl, a_l, wr, a_r, 3, speed, sol);
}

void hfo_vacuum_2d(double gamma, const double wl[4], double a_l, const double wr[4],
                   double a_r, double speed, double sol[4])
{
    vacuum(gamma, wl, a_l, wr, a_r, 4, speed, sol);
}

/* ======================================================================== */
/* HLLC: riemann_solver_hllc.c:30-376                                       */
/* ======================================================================== */
static double hllc_q(double gamma, double p_k, double p_star) /* hllc.c:30-40 */
{
    if (p_star <= p_k) {
        return 1.0;
    }
    return sqrt(1.0 + (0.5 + 0.5 / gamma) * (p_star / p_k - 1.0));
}

/* hllc.c:42-200 */
int hfo_flux_hllc_1d(double gamma, double tol, const double wl[3], const double wr[3], double f[3])
{
    const double rl = wl[0], ul = wl[1], pl = wl[2];
    const double rr = wr[0], ur = wr[1], pr = wr[2];
    const double al = (rl > 0.0) ? hfo_sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? hfo_sound_speed(gamma, rr, pr) : 0.0;

    if (rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul)) {
        double s[3];
        hfo_vacuum_1d(gamma, wl, al, wr, ar, 0.0, s);
        const double e = s[0] * (0.5 * s[1] * s[1] + s[2] / (s[0] * (gamma - 1.0)));
        f[0] = s[0] * s[1];
        f[1] = s[0] * s[1] * s[1] + s[2];
        f[2] = s[1] * (e + s[2]);
        return HFO_OK;
    }

    const double ps = hfo_guess_p(gamma, rl, ul, pl, al, rr, ur, pr, ar, tol);
    const double sl = ul - al * hllc_q(gamma, pl, ps);
    const double sr = ur + ar * hllc_q(gamma, pr, ps);
    const double sm = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) /
                      (rl * (sl - ul) - rr * (sr - ur));

    if (0.0 <= sl) {
        const double el = rl * (0.5 * ul * ul + pl / (rl * (gamma - 1.0)));
        f[0] = rl * ul;
        f[1] = rl * ul * ul + pl;
        f[2] = ul * (el + pl);
    } else if (sl <= 0.0 && 0.0 <= sm) {
        const double el = rl * (0.5 * ul * ul + pl / (rl * (gamma - 1.0)));
        const double fm = rl * ul;
        const double fu = fm * ul + pl;
        const double fe = ul * (el + pl);
        const double r_s = rl * (sl - ul) / (sl - sm);
        const double m_s = r_s * sm;
        const double e_s = r_s * (el / rl + (sm - ul) * (sm + pl / (rl * (sl - ul))));
        f[0] = fm + sl * (r_s - rl);
        f[1] = fu + sl * (m_s - rl * ul);
        f[2] = fe + sl * (e_s - el);
    } else if (sm <= 0.0 && 0.0 <= sr) {
        const double er = rr * (0.5 * ur * ur + pr / (rr * (gamma - 1.0)));
        const double fm = rr * ur;
        const double fu = fm * ur + pr;
        const double fe = ur * (er + pr);
        const double r_s = rr * (sr - ur) / (sr - sm);
        const double m_s = r_s * sm;
        const double e_s = r_s * (er / rr + (sm - ur) * (sm + pr / (rr * (sr - ur))));
        f[0] = fm + sr * (r_s - rr);
        f[1] = fu + sr * (m_s - rr * ur);
        f[2] = fe + sr * (e_s - er);
    } else if (sr <= 0.0) {
        const double er = rr * (0.5 * ur * ur + pr / (rr * (gamma - 1.0)));
        f[0] = rr * ur;
        f[1] = rr * ur * ur + pr;
        f[2] = ur * (er + pr);
    } else {
        return HFO_VALUE_ERROR; /* hllc.c:196-199 invalid wave speed */
    }
    return HFO_OK;
}

/* hllc.c:202-376; the state is (rho, u_normal, u_transverse, p). */
int hfo_flux_hllc_2d(double gamma, double tol, const double wl[4], const double wr[4], double f[4])
{
    const double rl = wl[0], ul = wl[1], vl = wl[2], pl = wl[3];
    const double rr = wr[0], ur = wr[1], vr = wr[2], pr = wr[3];
    const double al = (rl > 0.0) ? hfo_sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? hfo_sound_speed(gamma, rr, pr) : 0.0;

    if (rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul)) {
        double s[4];
        hfo_vacuum_2d(gamma, wl, al, wr, ar, 0.0, s);
        const double e = s[0] * (0.5 * (s[1] * s[1] + s[2] * s[2]) + s[3] / (s[0] * (gamma - 1.0)));
        f[0] = s[0] * s[1];
        f[1] = s[0] * s[1] * s[1] + s[3];
        f[2] = s[0] * s[1] * s[2];
        f[3] = s[1] * (e + s[3]);
        return HFO_OK;
    }

    const double ps = hfo_guess_p(gamma, rl, ul, pl, al, rr, ur, pr, ar, tol);
    const double sl = ul - al * hllc_q(gamma, pl, ps);
    const double sr = ur + ar * hllc_q(gamma, pr, ps);
    const double sm = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) /
                      (rl * (sl - ul) - rr * (sr - ur));

    if (0.0 <= sl) {
        const double el = rl * (0.5 * (ul * ul + vl * vl) + pl / (rl * (gamma - 1.0)));
        f[0] = rl * ul;
        f[1] = rl * ul * ul + pl;
        f[2] = rl * ul * vl;
        f[3] = ul * (el + pl);
    } else if (sl <= 0.0 && 0.0 <= sm) {
        const double el = rl * (0.5 * (ul * ul + vl * vl) + pl / (rl * (gamma - 1.0)));
        const double fm = rl * ul;
        const double fu = fm * ul + pl;
        const double fv = fm * vl;
        const double fe = ul * (el + pl);
        const double r_s = rl * (sl - ul) / (sl - sm);
        const double mu_s = r_s * sm;
        const double mv_s = r_s * vl;
        const double e_s = r_s * (el / rl + (sm - ul) * (sm + pl / (rl * (sl - ul))));
        f[0] = fm + sl * (r_s - rl);
        f[1] = fu + sl * (mu_s - rl * ul);
        f[2] = fv + sl * (mv_s - rl * vl);
        f[3] = fe + sl * (e_s - el);
    } else if (sm <= 0.0 && 0.0 <= sr) {
        const double er = rr * (0.5 * (ur * ur + vr * vr) + pr / (rr * (gamma - 1.0)));
        const double fm = rr * ur;
        const double fu = fm * ur + pr;
        const double fv = fm * vr;
        const double fe = ur * (er + pr);
        const double r_s = rr * (sr - ur) / (sr - sm);
        const double mu_s = r_s * sm;
        const double mv_s = r_s * vr;
        const double e_s = r_s * (er / rr + (sm - ur) * (sm + pr / (rr * (sr - ur))));
        f[0] = fm + sr * (r_s - rr);
        f[1] = fu + sr * (mu_s - rr * ur);
        f[2] = fv + sr * (mv_s - rr * vr);
        f[3] = fe + sr * (e_s - er);
    } else if (sr <= 0.0) {
        const double er = rr * (0.5 * (ur * ur + vr * vr) + pr / (rr * (gamma - 1.0)));
        f[0] = rr * ur;
        f[1] = rr * ur * ur + pr;
        f[2] = rr * ur * vr;
        f[3] = ur * (er + pr);
    } else {
        return HFO_VALUE_ERROR; /* hllc.c:372-375 */
    }
    return HFO_OK;
}

/* ======================================================================== */
/* Exact Riemann solver: riemann_solver_exact.c:332-1080                    */
/* ======================================================================== */
#define EXACT_NEWTON_MAX 100      /* riemann_solver_exact.c:26 */
#define EXACT_BRACKET_INTERVALS 100 /* :27 */
#define EXACT_BISECTION_MAX 5000  /* :28 */

/* exact.c:512-537 */
static double wave_f(double gamma, double rho, double p, double a, double ps)
{
    if (ps > p) {
        return (ps - p) * sqrt(coeff_A(gamma, rho) / (ps + coeff_B(gamma, p)));
    }
    return 2.0 * a * (pow((ps / p), (gamma - 1.0) / (2.0 * gamma)) - 1.0) / (gamma - 1.0);
}

/* exact.c:539-565 -- the shock branch's (B - p*) is kept as written (553). */
static double wave_df(double gamma, double rho, double p, double a, double ps)
{
    if (ps > p) {
        const double A = coeff_A(gamma, rho);
        const double B = coeff_B(gamma, p);
        return (1.0 - 0.5 * (ps - p) / (B - ps)) * sqrt(A / (B + ps));
    }
    return pow(ps / p, -0.5 * (gamma + 1.0) / gamma) / (p * a);
}

/* exact.c:567-585 */
static double pressure_fn(double gamma, const double *wl, double al, const double *wr, double ar, double ps)
{
    return wave_f(gamma, wl[0], wl[2], al, ps) + wave_f(gamma, wr[0], wr[2], ar, ps) + (wr[1] - wl[1]);
}

/* exact.c:604-763: Newton, then a bracket scan, then bisection */
static int solve_p_star(double gamma, const double *wl, double al, const double *wr, double ar,
                        double tol, double *p_star)
{
    const double pg = hfo_guess_p(gamma, wl[0], wl[1], wl[2], al, wr[0], wr[1], wr[2], ar, tol);
    bool bracketed = false;
    double p_hi = pg;
    double f_hi = pressure_fn(gamma, wl, al, wr, ar, p_hi);
    double p_lo = 0.0, f_lo = 0.0;

    double p0 = pg;
    for (int it = 0; it < EXACT_NEWTON_MAX; it++) {
        const double f = pressure_fn(gamma, wl, al, wr, ar, p0);
        const double df = wave_df(gamma, wl[0], wl[2], al, p0) + wave_df(gamma, wr[0], wr[2], ar, p0);
        if (!bracketed && f * f_hi < 0.0) {
            bracketed = true;
            f_lo = f;
            p_lo = p0;
        }
        const double p1 = p0 - f / df;
        if (p1 < 0.0) {
            break;
        }
        if (2.0 * fabs(p1 - p0) / (p1 + p0) < tol) {
            *p_star = p1;
            return HFO_OK;
        }
        p0 = p1;
    }

    if (!bracketed) {
        p_lo = 0.0;
        f_lo = pressure_fn(gamma, wl, al, wr, ar, p_lo);
        if (f_lo * f_hi >= 0.0) {
            p_hi *= 10.0;
            const double dp = (p_hi - p_lo) / EXACT_BRACKET_INTERVALS;
            for (int k = 0; k < EXACT_BRACKET_INTERVALS; k++) {
                const double a = p_lo + k * dp;
                const double b = p_lo + (k + 1) * dp;
                const double fa = pressure_fn(gamma, wl, al, wr, ar, a);
                const double fb = pressure_fn(gamma, wl, al, wr, ar, b);
                if (fa * fb < 0.0) {
                    p_lo = a;
                    f_lo = fa;
                    p_hi = b;
                    f_hi = fb;
                    break;
                }
            }
            if (f_lo * f_hi >= 0.0) {
                return HFO_FAILURE; /* exact.c:702-707 */
            }
        }
    }

    int count = 0;
    for (;;) {
        const double pm = 0.5 * (p_lo + p_hi);
        if (p_hi - p_lo < tol) {
            *p_star = pm;
            return HFO_OK;
        }
        const double fm = pressure_fn(gamma, wl, al, wr, ar, pm);
        if (fm * f_lo < 0.0) {
            p_hi = pm;
            f_hi = fm;
        } else {
            p_lo = pm;
            f_lo = fm;
        }
        count++;
        if (count > EXACT_BISECTION_MAX) {
            return HFO_FAILURE; /* exact.c:745-754 */
        }
    }
}

/* exact.c:765-815, 869-974 */
static void sample_left(double gamma, const double *w, double a, double us, double ps, double speed, double *sol)
{
    const double rho = w[0], u = w[1], p = w[2];
    if (ps > p) {
        const double ratio = ps / p;
        const double s = u - a * sqrt((gamma + 1.0) / (2.0 * gamma) * ratio + (gamma - 1.0) / (2.0 * gamma));
        if (speed < s) {
            sol[0] = rho; sol[1] = u; sol[2] = p;
        } else {
            const double gg = (gamma - 1.0) / (gamma + 1.0);
            sol[0] = (rho * (ratio + gg) / (gg * ratio + 1.0));
            sol[1] = us;
            sol[2] = ps;
        }
        return;
    }
    const double head = u - a;
    if (speed < head) {
        sol[0] = rho; sol[1] = u; sol[2] = p;
        return;
    }
    const double a_star = a * pow(ps / p, (gamma - 1.0) / (2.0 * gamma));
    const double tail = us - a_star;
    if (speed < tail) {
        const double c2 = 2.0 / (gamma + 1.0);
        const double common = pow((c2 + (u - speed) * ((gamma - 1.0) / (a * (gamma + 1.0)))), 2.0 / (gamma - 1.0));
        sol[0] = rho * common;
        sol[1] = c2 * (a + 0.5 * u * (gamma - 1.0) + speed);
        sol[2] = p * pow(common, gamma);
    } else {
        sol[0] = rho * pow(ps / p, 1.0 / gamma);
        sol[1] = us;
        sol[2] = ps;
    }
}

/* exact.c:817-867, 976-1080 (note the shock-speed product order at 993-995
 * and the fan's division order at 1062 differ from the left side) */
static void sample_right(double gamma, const double *w, double a, double us, double ps, double speed, double *sol)
{
    const double rho = w[0], u = w[1], p = w[2];
    if (ps > p) {
        const double ratio = ps / p;
        const double s = u + a * sqrt(ratio * (gamma + 1.0) / (2.0 * gamma) + (gamma - 1.0) / (2.0 * gamma));
        if (speed > s) {
            sol[0] = rho; sol[1] = u; sol[2] = p;
        } else {
            const double gg = (gamma - 1.0) / (gamma + 1.0);
            sol[0] = (rho * (ratio + gg) / (gg * ratio + 1.0));
            sol[1] = us;
            sol[2] = ps;
        }
        return;
    }
    const double head = u + a;
    if (speed > head) {
        sol[0] = rho; sol[1] = u; sol[2] = p;
        return;
    }
    const double a_star = a * pow(ps / p, (gamma - 1.0) / (2.0 * gamma));
    const double tail = us + a_star;
    if (speed > tail) {
        const double c2 = 2.0 / (gamma + 1.0);
        const double common = pow((c2 - (u - speed) * (gamma - 1.0) / (a * (gamma + 1.0))), 2.0 / (gamma - 1.0));
        sol[0] = rho * common;
        sol[1] = c2 * (-a + 0.5 * u * (gamma - 1.0) + speed);
        sol[2] = p * pow(common, gamma);
    } else {
        sol[0] = rho * pow(ps / p, 1.0 / gamma);
        sol[1] = us;
        sol[2] = ps;
    }
}

/* exact.c:332-469; also returns u* through *u_star_out when non-NULL
 * (NaN for the vacuum branch, which has no star state). */
static int exact_sample(double gamma, double tol, const double wl[3], const double wr[3],
                        double speed, double sol[3], double *u_star_out)
{
    const double al = (wl[0] > 0.0) ? hfo_sound_speed(gamma, wl[0], wl[2]) : 0.0;
    const double ar = (wr[0] > 0.0) ? hfo_sound_speed(gamma, wr[0], wr[2]) : 0.0;
    if (u_star_out) *u_star_out = NAN;
    if (wl[0] <= 0.0 || wr[0] <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (wr[1] - wl[1])) {
        hfo_vacuum_1d(gamma, wl, al, wr, ar, speed, sol);
        return HFO_OK;
    }
    double ps;
    const int rc = solve_p_star(gamma, wl, al, wr, ar, tol, &ps);
    if (rc != HFO_OK) {
        return rc;
    }
    const double us = 0.5 * (wl[1] + wr[1]) +
                      0.5 * (wave_f(gamma, wr[0], wr[2], ar, ps) - wave_f(gamma, wl[0], wl[2], al, ps));
    if (u_star_out) *u_star_out = us;
    if (speed < us) {
        sample_left(gamma, wl, al, us, ps, speed, sol);
    } else {
        sample_right(gamma, wr, ar, us, ps, speed, sol);
    }
    return HFO_OK;
}

int hfo_exact_1d(double gamma, double tol, const double wl[3], const double wr[3],
                 double speed, double sol[3])
{
    return exact_sample(gamma, tol, wl, wr, speed, sol, NULL);
}

/* exact.c:471-510 */
int hfo_flux_exact_1d(double gamma, double tol, const double wl[3], const double wr[3], double f[3])
{
    double s[3] = {NAN, NAN, NAN}; /* left unset by the reference on failure */
    const int rc = exact_sample(gamma, tol, wl, wr, 0.0, s, NULL);
    f[0] = s[0] * s[1];
    f[1] = s[0] * s[1] * s[1] + s[2];
    f[2] = (s[2] * (gamma / (gamma - 1.0)) + 0.5 * s[0] * s[1] * s[1]) * s[1];
    return rc;
}

/* No reference counterpart (solve_flux_2d rejects the exact solver,
 * riemann_solver.c:120-121).  Normal components: the 1D exact solution
 * sampled at x/t = 0, written exactly as exact.c:506-508 so that they equal
 * hfo_flux_exact_1d bitwise.  The transverse velocity is a passive scalar
 * carried by the contact: v_L when the sample lies left of u* (the same
 * "speed < u_star" test as exact.c:431), else v_R.  In the vacuum branch the
 * 2D vacuum sampler (riemann_solver.c:667-733) supplies v. */
int hfo_flux_exact_2d(double gamma, double tol, const double wl[4], const double wr[4], double f[4])
{
    const double l3[3] = {wl[0], wl[1], wl[3]};
    const double r3[3] = {wr[0], wr[1], wr[3]};
    double s[3] = {NAN, NAN, NAN}, us;
    const int rc = exact_sample(gamma, tol, l3, r3, 0.0, s, &us);
    double v;
    if (isnan(us)) {
        const double al = (wl[0] > 0.0) ? hfo_sound_speed(gamma, wl[0], wl[3]) : 0.0;
        const double ar = (wr[0] > 0.0) ? hfo_sound_speed(gamma, wr[0], wr[3]) : 0.0;
        double s4[4];
        hfo_vacuum_2d(gamma, wl, al, wr, ar, 0.0, s4);
        v = s4[2];
    } else {
        v = (0.0 < us) ? wl[2] : wr[2];
    }
    f[0] = s[0] * s[1];
    f[1] = s[0] * s[1] * s[1] + s[2];
    f[2] = f[0] * v;
    f[3] = (s[2] * (gamma / (gamma - 1.0)) + (0.5 * s[0] * s[1] * s[1] + 0.5 * s[0] * v * v)) * s[1];
    return rc;
}

void hfo_flux_batch(int dim, int solver, double gamma, double tol, int64_t n,
                    const double *wl, const double *wr, double *f, int *err)
{
    for (int64_t k = 0; k < n; k++) {
        int rc;
        if (dim == 1) {
            rc = (solver == HFO_HLLC) ? hfo_flux_hllc_1d(gamma, tol, wl + 3 * k, wr + 3 * k, f + 3 * k)
                                      : hfo_flux_exact_1d(gamma, tol, wl + 3 * k, wr + 3 * k, f + 3 * k);
        } else {
            rc = (solver == HFO_HLLC) ? hfo_flux_hllc_2d(gamma, tol, wl + 4 * k, wr + 4 * k, f + 4 * k)
                                      : hfo_flux_exact_2d(gamma, tol, wl + 4 * k, wr + 4 * k, f + 4 * k);
        }
        err[k] = rc;
    }
}

static int flux_1d(const hfo_params *p, const double wl[3], const double wr[3], double f[3])
{
    return (p->solver == HFO_HLLC) ? hfo_flux_hllc_1d(p->gamma, p->tol, wl, wr, f)
                                   : hfo_flux_exact_1d(p->gamma, p->tol, wl, wr, f);
}

static int flux_2d(const hfo_params *p, const double wl[4], const double wr[4], double f[4])
{
    return (p->solver == HFO_HLLC) ? hfo_flux_hllc_2d(p->gamma, p->tol, wl, wr, f)
                                   : hfo_flux_exact_2d(p->gamma, p->tol, wl, wr, f);
}

/* ======================================================================== */
/* Geometry and conversions: system.c:291-529, 735-917                      */
/* ======================================================================== */
void hfo_geometry(const hfo_params *p, double x_min, double x_max, double y_min, double y_max,
                  double *dx, double *dy, double *mid_x, double *mid_y,
                  double *area_x, double *area_y, double *volume)
{
    const int tx = p->nx + 2 * p->g;
    const double hx = (x_max - x_min) / (p->nx + 2 * p->g); /* system.c:302 */
    *dx = hx;
    for (int i = 0; i < tx; i++) mid_x[i] = x_min + (i + 0.5) * hx; /* :345 */
    if (p->coord == HFO_CART2D) {
        const int ty = p->ny + 2 * p->g;
        const double hy = (y_max - y_min) / (p->ny + 2 * p->g); /* :299 */
        *dy = hy;
        for (int j = 0; j < ty; j++) mid_y[j] = y_min + (j + 0.5) * hy;
        /* :400-415 (the last face entry is left unset by the reference and is
         * never read; it is given the same value here) */
        for (int i = 0; i <= tx; i++) area_x[i] = hy;
        for (int j = 0; j <= ty; j++) area_y[j] = hx;
        const double v = hx * hy; /* :490-503 */
        for (int k = 0; k < tx * ty; k++) volume[k] = v;
        return;
    }
    switch (p->coord) {
    case HFO_CART1D: /* :369-376, 456-464 */
        for (int i = 0; i <= tx; i++) area_x[i] = 1.0;
        for (int i = 0; i < tx; i++) volume[i] = hx;
        break;
    case HFO_CYL1D: /* :378-388 (entry 0 unset by the reference, never read), 466-477 */
        for (int i = 0; i <= tx; i++) area_x[i] = 2.0 * M_PI * (x_min + i * hx);
        for (int i = 0; i < tx; i++) {
            const double r0 = x_min + i * hx, r1 = x_min + (i + 1) * hx;
            volume[i] = M_PI * (r1 * r1 - r0 * r0);
        }
        break;
    case HFO_SPH1D: /* :389-399, 478-489 */
        for (int i = 0; i <= tx; i++) {
            const double r = x_min + i * hx;
            area_x[i] = 4.0 * M_PI * r * r;
        }
        for (int i = 0; i < tx; i++) {
            const double r0 = x_min + i * hx, r1 = x_min + (i + 1) * hx;
            volume[i] = (4.0 / 3.0) * M_PI * (r1 * r1 * r1 - r0 * r0 * r0);
        }
        break;
    default:
        break;
    }
}

static int64_t n_total(const hfo_params *p)
{
    const int64_t tx = p->nx + 2 * p->g;
    return (p->coord == HFO_CART2D) ? tx * (p->ny + 2 * p->g) : tx;
}

/* system.c:849-917 */
void hfo_p2c(const hfo_params *p, const double *volume, double *const *w, double *const *u)
{
    const double gm = p->gamma;
    const int64_t n = n_total(p);
    if (p->coord == HFO_CART2D) {
        for (int64_t k = 0; k < n; k++) {
            const double r = w[0][k], vx = w[1][k], vy = w[2][k], pr = w[3][k], vol = volume[k];
            u[0][k] = r * vol;
            u[1][k] = u[0][k] * vx;
            u[2][k] = u[0][k] * vy;
            u[3][k] = vol * (0.5 * r * (vx * vx + vy * vy) + pr / (gm - 1.0));
        }
    } else {
        for (int64_t k = 0; k < n; k++) {
            const double r = w[0][k], vx = w[1][k], pr = w[2][k], vol = volume[k];
            u[0][k] = r * vol;
            u[1][k] = u[0][k] * vx;
            u[2][k] = volume[k] * (0.5 * r * vx * vx + pr / (gm - 1.0));
        }
    }
}

/* system.c:735-802 */
void hfo_c2p(const hfo_params *p, const double *volume, double *const *u, double *const *w)
{
    const double gm = p->gamma;
    const int64_t n = n_total(p);
    if (p->coord == HFO_CART2D) {
        for (int64_t k = 0; k < n; k++) {
            const double m = u[0][k], mx = u[1][k], my = u[2][k], e = u[3][k], vol = volume[k];
            w[0][k] = m / vol;
            w[1][k] = mx / m;
            w[2][k] = my / m;
            w[3][k] = (gm - 1.0) * (e - 0.5 * m * (w[1][k] * w[1][k] + w[2][k] * w[2][k])) / vol;
        }
    } else {
        for (int64_t k = 0; k < n; k++) {
            const double m = u[0][k], mx = u[1][k], e = u[2][k], vol = volume[k];
            w[0][k] = m / vol;
            w[1][k] = mx / m;
            w[2][k] = (gm - 1.0) * (e - 0.5 * m * w[1][k] * w[1][k]) / vol;
        }
    }
}

/* ======================================================================== */
/* Boundary fill on primitives: boundary.c:411-777                          */
/* Ghost cells mirror (reflective/transmissive) or wrap (periodic); 2D      */
/* x-ghosts are filled on interior rows only and y-ghosts on interior       */
/* columns only, so the corner blocks are never written.                    */
/* ======================================================================== */
/* one side of a line of cells; cell k of the line lives at base + k*stride,
 * k in [0, n+2g).  A periodic pair is filled with the high side, both ghost
 * strips at once (boundary.c:454-457, 477-490). */
static void bc_side(int hi, int flag, double *const *w, int nf, int normal, int64_t base, int64_t stride, int n,
                    int g)
{
#define AT(f, k) w[f][base + (int64_t)(k) * stride]
    if (flag == HFO_BC_REFLECTIVE || flag == HFO_BC_TRANSMISSIVE) {
        for (int i = 0; i < g; i++) {
            for (int f = 0; f < nf; f++) {
                const double s = hi ? AT(f, g + n - 1 - i) : AT(f, g + i);
                const double v = (flag == HFO_BC_REFLECTIVE && f == normal) ? -s : s;
                if (hi) AT(f, g + n + i) = v;
                else AT(f, g - 1 - i) = v;
            }
        }
    } else if (hi && flag == HFO_BC_PERIODIC) {
        for (int i = 0; i < g; i++) {
            for (int f = 0; f < nf; f++) {
                AT(f, g + n + i) = AT(f, g + i);
                AT(f, g - 1 - i) = AT(f, g + n - 1 - i);
            }
        }
    }
#undef AT
}

/* boundary.c:411-777: x_min, x_max, then y_min, y_max, each side complete
 * before the next; a custom side calls its hook (boundary.h:17-33) */
void hfo_bc(const hfo_params *p, double *const *w)
{
    const int g = p->g, nx = p->nx;
    if (p->coord != HFO_CART2D) {
        for (int hi = 0; hi < 2; hi++) {
            if (p->bc[hi] == HFO_BC_CUSTOM) {
                if (p->hook[hi]) ((hfo_bc1d_fn)p->hook[hi])(w[0], w[1], w[2], g, nx);
            } else {
                bc_side(hi, p->bc[hi], w, 3, 1, 0, 1, nx, g);
            }
        }
        return;
    }
    const int ny = p->ny;
    const int64_t tx = nx + 2 * g;
    for (int side = 0; side < 4; side++) {
        const int hi = side & 1;
        if (p->bc[side] == HFO_BC_CUSTOM) {
            if (p->hook[side]) ((hfo_bc2d_fn)p->hook[side])(w[0], w[1], w[2], w[3], g, nx, ny);
        } else if (side < 2) {
            for (int j = g; j < g + ny; j++) bc_side(hi, p->bc[side], w, 4, 1, (int64_t)j * tx, 1, nx, g);
        } else {
            for (int i = g; i < g + nx; i++) bc_side(hi, p->bc[side], w, 4, 2, i, tx, ny, g);
        }
    }
}

/* ======================================================================== */
/* CFL time step: utils.c:28-95                                             */
/* ======================================================================== */
double hfo_dt_1d(const hfo_params *p, double *const *w, double cfl)
{
    double smax = -INFINITY;
    for (int i = p->g; i < p->nx + p->g; i++) {
        const double a = hfo_sound_speed(p->gamma, w[0][i], w[2][i]);
        const double s = fabs(w[1][i]);
        if (a + s > smax) smax = a + s;
    }
    return cfl * p->dx / smax;
}

double hfo_dt_2d(const hfo_params *p, double *const *w, double cfl)
{
    const int64_t tx = p->nx + 2 * p->g;
    double sx = -INFINITY, sy = -INFINITY;
    for (int i = p->g; i < p->nx + p->g; i++) {
        for (int j = p->g; j < p->ny + p->g; j++) {
            const int64_t k = (int64_t)j * tx + i;
            const double a = hfo_sound_speed(p->gamma, w[0][k], w[3][k]);
            const double ax = fabs(w[1][k]), ay = fabs(w[2][k]);
            if (a + ax > sx) sx = a + ax;
            if (a + ay > sy) sy = a + ay;
        }
    }
    return cfl / (sx / p->dx + sy / p->dy);
}

/* ======================================================================== */
/* Source terms: source_term.c:28-259                                       */
/* ======================================================================== */
void hfo_gravity_2d(const hfo_params *p, double dt, double *const *u)
{
    const double grav = p->gravity;
    if (grav == 0.0) return; /* :195-198 */
    const int64_t tx = p->nx + 2 * p->g;
    for (int i = p->g; i < p->nx + p->g; i++) {
        for (int j = p->g; j < p->ny + p->g; j++) {
            const int64_t k = (int64_t)j * tx + i;
            const double m = u[0][k], my = u[2][k];
            const double pre = -dt * grav;
            const double kmy = pre * m;
            const double k1 = pre * my;
            const double k2 = pre * (my + 0.5 * kmy);
            const double k3 = k2;
            const double k4 = pre * (my + kmy);
            u[2][k] += kmy;
            u[3][k] += (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0;
        }
    }
}

/* source_term.c:28-43 */
static void geo_rate(double pre, double gamma, double rho, double sm, double se, double out[3])
{
    const double pr = (gamma - 1) * (se - 0.5 * sm * sm / rho);
    out[0] = pre * sm;
    out[1] = pre * sm * sm / rho;
    out[2] = pre * sm * (se + pr) / rho;
}

/* source_term.c:45-164: classical RK4 per interior cell on (m, mx, E). */
void hfo_geometry_source_1d(const hfo_params *p, double dt, const double *mid_x,
                            const double *volume, double *const *w, double *const *u)
{
    const double alpha = (p->coord == HFO_CYL1D) ? 1.0 : 2.0;
    for (int i = p->g; i < p->nx + p->g; i++) {
        const double vol = volume[i];
        const double m = u[0][i], mx = u[1][i], e = u[2][i];
        const double pre = -alpha * vol / mid_x[i];
        double k1[3], k2[3], k3[3], k4[3], t[3];
        geo_rate(pre, p->gamma, w[0][i], mx / vol, e / vol, k1);
        t[0] = m + 0.5 * dt * k1[0];
        t[1] = mx + 0.5 * dt * k1[1];
        t[2] = e + 0.5 * dt * k1[2];
        geo_rate(pre, p->gamma, t[0] / vol, t[1] / vol, t[2] / vol, k2);
        t[0] = m + 0.5 * dt * k2[0];
        t[1] = mx + 0.5 * dt * k2[1];
        t[2] = e + 0.5 * dt * k2[2];
        geo_rate(pre, p->gamma, t[0] / vol, t[1] / vol, t[2] / vol, k3);
        t[0] = m + dt * k3[0];
        t[1] = mx + dt * k3[1];
        t[2] = e + dt * k3[2];
        geo_rate(pre, p->gamma, t[0] / vol, t[1] / vol, t[2] / vol, k4);
        for (int f = 0; f < 3; f++) {
            u[f][i] += dt * (k1[f] + 2.0 * k2[f] + 2.0 * k3[f] + k4[f]) / 6.0;
        }
    }
}

/* ======================================================================== */
/* MUSCL-Hancock half-step predictor                                        */
/* integrator_muscl_hancock.c:692-740 (x), 977-1025 (y), 202-232 (1D).      */
/* s = (rho, u_n, u_t, p) in sweep-local order, d = limited slopes,         */
/* c = dt*A/V.  Outputs the state at the low face (lo) and the high face    */
/* (hi) of the cell.  The y-sweep's density row multiplies the x-velocity   */
/* slope (977-982, 1002-1007), i.e. the transverse slope in sweep order;    */
/* 'quirk' selects that form.                                               */
/* ======================================================================== */
static void predict_x(const double *s, const double *d, double c, double a, double *lo, double *hi)
{
    /* 692-740 */
    const double up = 1.0 + c * s[1];
    const double um = 1.0 - c * s[1];
    const double dr_lo = -0.5 * (up * d[0] + c * s[0] * d[1] + 0.0 + 0.0);
    const double du_lo = -0.5 * (0.0 + up * d[1] + 0.0 + (c / s[0]) * d[3]);
    const double dv_lo = -0.5 * (0.0 + 0.0 + up * d[2] + 0.0);
    const double dp_lo = -0.5 * (0.0 + c * s[0] * a * a * d[1] + 0.0 + up * d[3]);
    const double dr_hi = 0.5 * (um * d[0] - c * s[0] * d[1] + 0.0 + 0.0);
    const double du_hi = 0.5 * (0.0 + um * d[1] + 0.0 - (c / s[0]) * d[3]);
    const double dv_hi = 0.5 * (0.0 + 0.0 + um * d[2] + 0.0);
    const double dp_hi = 0.5 * (0.0 - c * s[0] * a * a * d[1] + 0.0 + um * d[3]);
    lo[0] = s[0] + dr_lo; lo[1] = s[1] + du_lo; lo[2] = s[2] + dv_lo; lo[3] = s[3] + dp_lo;
    hi[0] = s[0] + dr_hi; hi[1] = s[1] + du_hi; hi[2] = s[2] + dv_hi; hi[3] = s[3] + dp_hi;
}

/* 977-1025: s and d are in (rho, u, v, p) order; the normal velocity is v. */
static void predict_y(const double *s, const double *d, double c, double a, double *lo, double *hi)
{
    const double vp = 1.0 + c * s[2];
    const double vm = 1.0 - c * s[2];
    const double dr_lo = -0.5 * (vp * d[0] + c * s[0] * d[1] + 0.0 + 0.0);
    const double du_lo = -0.5 * (0.0 + vp * d[1] + 0.0 + 0.0);
    const double dv_lo = -0.5 * (0.0 + 0.0 + vp * d[2] + (c / s[0]) * d[3]);
    const double dp_lo = -0.5 * (0.0 + 0.0 + c * s[0] * a * a * d[2] + vp * d[3]);
    const double dr_hi = 0.5 * (vm * d[0] - c * s[0] * d[1] + 0.0 + 0.0);
    const double du_hi = 0.5 * (0.0 + vm * d[1] + 0.0 + 0.0);
    const double dv_hi = 0.5 * (0.0 + 0.0 + vm * d[2] - (c / s[0]) * d[3]);
    const double dp_hi = 0.5 * (0.0 + 0.0 - c * s[0] * a * a * d[2] + vm * d[3]);
    lo[0] = s[0] + dr_lo; lo[1] = s[1] + du_lo; lo[2] = s[2] + dv_lo; lo[3] = s[3] + dp_lo;
    hi[0] = s[0] + dr_hi; hi[1] = s[1] + du_hi; hi[2] = s[2] + dv_hi; hi[3] = s[3] + dp_hi;
}

/* 202-232 */
static void predict_1d(const double *s, const double *d, double c, double a, double *lo, double *hi)
{
    const double dr_lo = -0.5 * ((1.0 + c * s[1]) * d[0] + c * s[0] * d[1] + 0.0);
    const double du_lo = -0.5 * (0.0 + (1.0 + c * s[1]) * d[1] + (c / s[0]) * d[2]);
    const double dp_lo = -0.5 * (0.0 + c * s[0] * a * a * d[1] + (1.0 + c * s[1]) * d[2]);
    const double dr_hi = 0.5 * ((1.0 - c * s[1]) * d[0] - c * s[0] * d[1] + 0.0);
    const double du_hi = 0.5 * (0.0 + (1.0 - c * s[1]) * d[1] - (c / s[0]) * d[2]);
    const double dp_hi = 0.5 * (0.0 + -c * s[0] * a * a * d[1] + (1.0 - c * s[1]) * d[2]);
    lo[0] = s[0] + dr_lo; lo[1] = s[1] + du_lo; lo[2] = s[2] + dp_lo;
    hi[0] = s[0] + dr_hi; hi[1] = s[1] + du_hi; hi[2] = s[2] + dp_hi;
}

/* ======================================================================== */
/* 2D step: integrator_muscl_hancock.c:552-1141                             */
/* ======================================================================== */
typedef struct {
    int code;
    int64_t cell;
} step_err;

static void record_err(step_err *e, int code, int64_t cell)
{
    if (code != HFO_OK && e->code == HFO_OK) {
        e->code = code;
        e->cell = cell;
    }
}

/* one x-row: reconstruct cells g-1..g+nx (625-756), then fluxes and updates
 * for interfaces 0..nx (759-815).  lo/hi hold face states per cell. */
static void sweep_x_row(const hfo_params *p, double dt, int j, const double *area_x, const double *volume,
                        double *const *w, double *const *u, double (*lo)[4], double (*hi)[4], step_err *e)
{
    const int g = p->g, nx = p->nx;
    const int64_t tx = nx + 2 * g;
    for (int i = 0; i < nx + 2; i++) {
        const int64_t k = (int64_t)j * tx + (i + g - 1);
        const double s[4] = {w[0][k], w[1][k], w[2][k], w[3][k]};
        const double c = dt * area_x[i + g - 1] / volume[k];
        const double a = hfo_sound_speed(p->gamma, s[0], s[3]);
        double d[4];
        for (int f = 0; f < 4; f++) {
            d[f] = hfo_limit(p->limiter, s[f] - w[f][k - 1], w[f][k + 1] - s[f]);
        }
        predict_x(s, d, c, a, lo[i], hi[i]);
    }
    for (int i = 0; i <= nx; i++) {
        double f[4];
        const int64_t kl = (int64_t)j * tx + (g + i - 1);
        const int64_t kr = kl + 1;
        record_err(e, flux_2d(p, hi[i], lo[i + 1], f), kl);
        const double dm = f[0] * dt, dmx = f[1] * dt, dmy = f[2] * dt, de = f[3] * dt;
        const double al = area_x[g + i - 1], ar = area_x[g + i];
        u[0][kl] -= dm * al; u[1][kl] -= dmx * al; u[2][kl] -= dmy * al; u[3][kl] -= de * al;
        u[0][kr] += dm * ar; u[1][kr] += dmx * ar; u[2][kr] += dmy * ar; u[3][kr] += de * ar;
    }
}

/* one y-column: 905-1101 */
static void sweep_y_col(const hfo_params *p, double dt, int i, const double *area_y, const double *volume,
                        double *const *w, double *const *u, double (*lo)[4], double (*hi)[4], step_err *e)
{
    const int g = p->g, ny = p->ny;
    const int64_t tx = p->nx + 2 * g;
    for (int j = 0; j < ny + 2; j++) {
        const int64_t k = (int64_t)(j + g - 1) * tx + i;
        const double s[4] = {w[0][k], w[1][k], w[2][k], w[3][k]};
        const double c = dt * area_y[j + g - 1] / volume[k];
        const double a = hfo_sound_speed(p->gamma, s[0], s[3]);
        double d[4];
        for (int f = 0; f < 4; f++) {
            d[f] = hfo_limit(p->limiter, s[f] - w[f][k - tx], w[f][k + tx] - s[f]);
        }
        predict_y(s, d, c, a, lo[j], hi[j]);
    }
    for (int j = 0; j <= ny; j++) {
        /* solver sees (rho, v, u, p) and returns (mass, my, mx, E): 1051-1067 */
        const double l[4] = {hi[j][0], hi[j][2], hi[j][1], hi[j][3]};
        const double r[4] = {lo[j + 1][0], lo[j + 1][2], lo[j + 1][1], lo[j + 1][3]};
        double f[4];
        const int64_t kb = (int64_t)(g + j - 1) * tx + i;
        const int64_t kt = kb + tx;
        record_err(e, flux_2d(p, l, r, f), kb);
        const double dm = f[0] * dt, dmx = f[2] * dt, dmy = f[1] * dt, de = f[3] * dt;
        const double ab = area_y[g + j - 1], at = area_y[g + j];
        u[0][kb] -= dm * ab; u[1][kb] -= dmx * ab; u[2][kb] -= dmy * ab; u[3][kb] -= de * ab;
        u[0][kt] += dm * at; u[1][kt] += dmx * at; u[2][kt] += dmy * at; u[3][kt] += de * at;
    }
}

int hfo_mh2d_step(const hfo_params *p, double dt, const double *area_x, const double *area_y,
                  const double *volume, double *const *w, double *const *u, int64_t *err_cell)
{
    step_err err = {HFO_OK, -1};
    const int nthreads = (p->nthreads > 0) ? p->nthreads : 1;

#pragma omp parallel num_threads(nthreads)
    {
        step_err mine = {HFO_OK, -1};
        double (*lo)[4] = malloc((size_t)(p->nx + 2) * sizeof *lo);
        double (*hi)[4] = malloc((size_t)(p->nx + 2) * sizeof *hi);
#pragma omp for schedule(static)
        for (int j = p->g; j < p->g + p->ny; j++) {
            sweep_x_row(p, dt, j, area_x, volume, w, u, lo, hi, &mine);
        }
        free(lo);
        free(hi);
#pragma omp critical
        if (mine.code != HFO_OK && (err.code == HFO_OK || mine.cell < err.cell)) err = mine;
    }
    if (err.code == HFO_OK) {
#pragma omp parallel num_threads(nthreads)
        {
            step_err mine = {HFO_OK, -1};
            double (*lo)[4] = malloc((size_t)(p->ny + 2) * sizeof *lo);
            double (*hi)[4] = malloc((size_t)(p->ny + 2) * sizeof *hi);
#pragma omp for schedule(static)
            for (int i = p->g; i < p->g + p->nx; i++) {
                sweep_y_col(p, dt, i, area_y, volume, w, u, lo, hi, &mine);
            }
            free(lo);
            free(hi);
#pragma omp critical
            if (mine.code != HFO_OK && (err.code == HFO_OK || mine.cell < err.cell)) err = mine;
        }
    }
    if (err.code != HFO_OK) {
        if (err_cell) *err_cell = err.cell;
        return err.code;
    }
    /* 1121-1141: c2p over every cell, BC on primitives, gravity (+c2p, BC) */
    hfo_c2p(p, volume, u, w);
    hfo_bc(p, w);
    if (p->gravity != 0.0) {
        hfo_gravity_2d(p, dt, u);
        hfo_c2p(p, volume, u, w);
        hfo_bc(p, w);
    }
    return HFO_OK;
}

/* ======================================================================== */
/* 1D step: integrator_muscl_hancock.c:146-330 (serial flux loop)           */
/* ======================================================================== */
int hfo_mh1d_step(const hfo_params *p, double dt, const double *area_x, const double *mid_x,
                  const double *volume, double *const *w, double *const *u, int64_t *err_cell)
{
    const int g = p->g, n = p->nx;
    double (*lo)[3] = malloc((size_t)(n + 2) * sizeof *lo);
    double (*hi)[3] = malloc((size_t)(n + 2) * sizeof *hi);
    step_err err = {HFO_OK, -1};
    for (int i = 0; i < n + 2; i++) {
        const int k = g + i - 1;
        const double s[3] = {w[0][k], w[1][k], w[2][k]};
        const double c = dt * area_x[k] / volume[k];
        const double a = hfo_sound_speed(p->gamma, s[0], s[2]);
        double d[3];
        for (int f = 0; f < 3; f++) {
            d[f] = hfo_limit(p->limiter, s[f] - w[f][k - 1], w[f][k + 1] - s[f]);
        }
        predict_1d(s, d, c, a, lo[i], hi[i]);
    }
    for (int i = 0; i <= n; i++) {
        double f[3];
        const int k = g + i;
        record_err(&err, flux_1d(p, hi[i], lo[i + 1], f), k - 1);
        const double dm = f[0] * dt, dmx = f[1] * dt, de = f[2] * dt;
        u[0][k - 1] -= dm * area_x[k - 1]; u[1][k - 1] -= dmx * area_x[k - 1]; u[2][k - 1] -= de * area_x[k - 1];
        u[0][k] += dm * area_x[k]; u[1][k] += dmx * area_x[k]; u[2][k] += de * area_x[k];
    }
    free(lo);
    free(hi);
    if (err.code != HFO_OK) {
        if (err_cell) *err_cell = err.cell;
        return err.code;
    }
    hfo_c2p(p, volume, u, w);
    hfo_bc(p, w);
    if (p->coord == HFO_CYL1D || p->coord == HFO_SPH1D) {
        hfo_geometry_source_1d(p, dt, mid_x, volume, w, u);
        hfo_c2p(p, volume, u, w);
        hfo_bc(p, w);
    }
    return HFO_OK;
}

/* ======================================================================== */
/* First-order Godunov steps: integrator_godunov_first_order.c              */
/* The faces take the cells' own primitive states.                          */
/* ======================================================================== */
/* 2D: 385-505 (x sweep over interior rows, then y sweep over interior
 * columns, both from the same W), then 508-529 (c2p, BC, gravity). */
int hfo_god2d_step(const hfo_params *p, double dt, const double *area_x, const double *area_y,
                   const double *volume, double *const *w, double *const *u, int64_t *err_cell)
{
    const int g = p->g, nx = p->nx, ny = p->ny;
    const int64_t tx = nx + 2 * g;
    step_err err = {HFO_OK, -1};
    for (int j = g; j < g + ny; j++) {
        for (int i = g; i < g + nx + 1; i++) {
            const int64_t kl = (int64_t)j * tx + (i - 1);
            const int64_t kr = kl + 1;
            const double wl[4] = {w[0][kl], w[1][kl], w[2][kl], w[3][kl]};
            const double wr[4] = {w[0][kr], w[1][kr], w[2][kr], w[3][kr]};
            double f[4];
            /* the row loop runs on after a failure; the sweep then ends the step (441-444) */
            record_err(&err, flux_2d(p, wl, wr, f), kl);
            for (int q = 0; q < 4; q++) u[q][kl] -= dt * f[q] * area_x[i - 1];
            for (int q = 0; q < 4; q++) u[q][kr] += dt * f[q] * area_x[i];
        }
    }
    if (err.code == HFO_OK) {
        for (int i = g; i < g + nx; i++) {
            for (int j = g; j < g + ny + 1; j++) {
                const int64_t kb = (int64_t)(j - 1) * tx + i;
                const int64_t kt = kb + tx;
                /* the solver sees (rho, v, u, p) and returns (mass, my, mx, E): 461-476 */
                const double wl[4] = {w[0][kb], w[2][kb], w[1][kb], w[3][kb]};
                const double wr[4] = {w[0][kt], w[2][kt], w[1][kt], w[3][kt]};
                double f[4];
                record_err(&err, flux_2d(p, wl, wr, f), kb);
                const double fx[4] = {f[0], f[2], f[1], f[3]};
                for (int q = 0; q < 4; q++) u[q][kb] -= dt * fx[q] * area_y[j - 1];
                for (int q = 0; q < 4; q++) u[q][kt] += dt * fx[q] * area_y[j];
            }
        }
    }
    if (err.code != HFO_OK) {
        if (err_cell) *err_cell = err.cell;
        return err.code;
    }
    hfo_c2p(p, volume, u, w);
    hfo_bc(p, w);
    if (p->gravity != 0.0) {
        hfo_gravity_2d(p, dt, u);
        hfo_c2p(p, volume, u, w);
        hfo_bc(p, w);
    }
    return HFO_OK;
}

/* 1D: 150-199 (the serial build leaves the loop at the first failed flux,
 * before that interface's update: 174-179), then 202-220 (c2p, BC,
 * geometric source). */
int hfo_god1d_step(const hfo_params *p, double dt, const double *area_x, const double *mid_x,
                   const double *volume, double *const *w, double *const *u, int64_t *err_cell)
{
    const int g = p->g, n = p->nx;
    for (int i = g; i < g + n + 1; i++) {
        const double wl[3] = {w[0][i - 1], w[1][i - 1], w[2][i - 1]};
        const double wr[3] = {w[0][i], w[1][i], w[2][i]};
        double f[3];
        const int rc = flux_1d(p, wl, wr, f);
        if (rc != HFO_OK) {
            if (err_cell) *err_cell = i - 1;
            return rc;
        }
        const double dm = f[0] * dt, dmx = f[1] * dt, de = f[2] * dt;
        u[0][i - 1] -= dm * area_x[i - 1]; u[1][i - 1] -= dmx * area_x[i - 1]; u[2][i - 1] -= de * area_x[i - 1];
        u[0][i] += dm * area_x[i]; u[1][i] += dmx * area_x[i]; u[2][i] += de * area_x[i];
    }
    hfo_c2p(p, volume, u, w);
    hfo_bc(p, w);
    if (p->coord == HFO_CYL1D || p->coord == HFO_SPH1D) {
        hfo_geometry_source_1d(p, dt, mid_x, volume, w, u);
        hfo_c2p(p, volume, u, w);
        hfo_bc(p, w);
    }
    return HFO_OK;
}

/* ======================================================================== */
/* Random choice (Glimm) 1D: integrator_random_choice_1d.c                  */
/* ======================================================================== */
/* 28-43: the van der Corput sequence with k1 = 5, k2 = 3 (lines 25-26) */
double hfo_van_der_corput(int64_t n)
{
    double result = 0.0;
    const int base = 5;
    double bk = 1.0 / (double)base;
    while (n > 0) {
        const int a = (int)(n % base);
        result += ((3 * a) % base) * bk;
        n /= base;
        bk /= (double)base;
    }
    return result;
}

/* 160-242: every cell g..g+n takes the exact solution sampled at
 * theta*dx/dt of the problem with its left neighbour (theta <= 1/2) or at
 * (theta-1)*dx/dt with its right neighbour, from a copy of W; then the
 * boundary fill, p2c and the geometric source.  A failed solve does not stop
 * the loop; the last one is reported (213-221).  With one ghost layer the
 * reference reads one cell past its arrays when theta > 1/2; the read is
 * clamped to the last cell here. */
int hfo_rc1d_step(const hfo_params *p, double dt, int64_t step, const double *mid_x, const double *volume,
                  double *const *w, double *const *u, int64_t *err_cell)
{
    const int g = p->g, n = p->nx, tx = n + 2 * g;
    const double theta = hfo_van_der_corput(step + 1);
    double *tmp = malloc((size_t)3 * tx * sizeof(double));
    for (int f = 0; f < 3; f++) memcpy(tmp + (size_t)f * tx, w[f], (size_t)tx * sizeof(double));
    int err = HFO_OK;
    for (int i = g; i < g + n + 1; i++) {
        int a, b;
        double speed;
        if (theta <= 0.5) {
            a = i - 1;
            b = i;
            speed = theta * p->dx / dt;
        } else {
            a = i;
            b = (i + 1 < tx) ? i + 1 : tx - 1;
            speed = (theta - 1.0) * p->dx / dt;
        }
        const double wl[3] = {tmp[a], tmp[tx + a], tmp[2 * tx + a]};
        const double wr[3] = {tmp[b], tmp[tx + b], tmp[2 * tx + b]};
        double sol[3];
        const int rc = exact_sample(p->gamma, p->tol, wl, wr, speed, sol, NULL);
        w[0][i] = sol[0];
        w[1][i] = sol[1];
        w[2][i] = sol[2];
        if (rc != HFO_OK) {
            err = rc;
            if (err_cell) *err_cell = i;
        }
    }
    free(tmp);
    if (err != HFO_OK) return err;
    hfo_bc(p, w);
    hfo_p2c(p, volume, w, u);
    if (p->coord == HFO_CYL1D || p->coord == HFO_SPH1D) {
        hfo_geometry_source_1d(p, dt, mid_x, volume, w, u);
        hfo_c2p(p, volume, u, w);
        hfo_bc(p, w);
    }
    return HFO_OK;
}

/* ======================================================================== */
/* Time loops: integrator_muscl_hancock.c:124-349 (1D), 530-1160 (2D);      */
/* integrator_godunov_first_order.c:128-240 (1D), 363-546 (2D), and         */
/* integrator_random_choice_1d.c:138-254, the same loop.                    */
/* scheme: 0 MUSCL-Hancock, 1 first-order Godunov, 2 random choice (1D).    */
/* ======================================================================== */
static void run_loop(const hfo_params *p, hfo_run *r, int dim, int scheme, const double *area_x,
                     const double *area_y, const double *mid_x, const double *volume, double *const *w,
                     double *const *u)
{
    r->num_steps = 0;
    r->dt = 0.0;
    r->t = 0.0;
    r->error = HFO_OK;
    r->error_cell = -1;
    while (r->t < r->tf && (r->max_steps < 0 || r->num_steps < r->max_steps)) {
        if (r->fixed_dt > 0.0) {
            r->dt = r->fixed_dt;
        } else {
            const double cfl = (r->num_steps < r->shrink_steps) ? r->cfl * r->shrink_factor : r->cfl;
            r->dt = (dim == 2) ? hfo_dt_2d(p, w, cfl) : hfo_dt_1d(p, w, cfl);
        }
        if (r->t + r->dt > r->tf) {
            r->dt = r->tf - r->t;
        }
        if (r->dt == 0.0) {
            r->error = HFO_VALUE_ERROR; /* "dt is zero." */
            return;
        }
        int rc;
        if (dim == 2) {
            rc = scheme ? hfo_god2d_step(p, r->dt, area_x, area_y, volume, w, u, &r->error_cell)
                        : hfo_mh2d_step(p, r->dt, area_x, area_y, volume, w, u, &r->error_cell);
        } else if (scheme == 2) {
            rc = hfo_rc1d_step(p, r->dt, r->num_steps, mid_x, volume, w, u, &r->error_cell);
        } else {
            rc = scheme ? hfo_god1d_step(p, r->dt, area_x, mid_x, volume, w, u, &r->error_cell)
                        : hfo_mh1d_step(p, r->dt, area_x, mid_x, volume, w, u, &r->error_cell);
        }
        if (rc != HFO_OK) {
            r->error = rc;
            return;
        }
        r->t += r->dt;
        r->num_steps++;
    }
}

void hfo_mh2d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *area_y,
                  const double *volume, double *const *w, double *const *u)
{
    run_loop(p, r, 2, 0, area_x, area_y, NULL, volume, w, u);
}

void hfo_mh1d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *mid_x,
                  const double *volume, double *const *w, double *const *u)
{
    run_loop(p, r, 1, 0, area_x, NULL, mid_x, volume, w, u);
}

void hfo_god2d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *area_y,
                   const double *volume, double *const *w, double *const *u)
{
    run_loop(p, r, 2, 1, area_x, area_y, NULL, volume, w, u);
}

void hfo_god1d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *mid_x,
                   const double *volume, double *const *w, double *const *u)
{
    run_loop(p, r, 1, 1, area_x, NULL, mid_x, volume, w, u);
}

void hfo_rc1d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *mid_x,
                  const double *volume, double *const *w, double *const *u)
{
    run_loop(p, r, 1, 2, area_x, NULL, mid_x, volume, w, u);
}
