// hfv_math.hpp -- per-cell and per-interface physics for the MUSCL-Hancock
// update, as __device__ functions for gfx950.
//
// Every expression keeps the operand order of the reference C code it cites,
// and the library is compiled with -ffp-contract=off, so each result is the
// correctly rounded value of the same sequence of IEEE operations (divisions
// and square roots are correctly rounded on this target).  pow() is the only
// routine whose last-place rounding may differ from glibc's; it appears only
// in the two-rarefaction guess, the exact solver and the vacuum samplers.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace hfv {

// flags: the reference's numeric values (slope_limiter.h, riemann_solver.h, boundary.h)
enum Limiter : int { MINMOD = 1, VAN_LEER = 2, MC = 3 };
enum Solver : int { EXACT = 1, HLLC = 2 };
enum Bc : int { BC_NONE = 1, BC_REFLECTIVE = 2, BC_TRANSMISSIVE = 3, BC_PERIODIC = 4, BC_CUSTOM = 5 };

// device-side failure kinds (mapped to the reference's error messages by the host)
enum ErrKind : int {
    ERR_NONE = 0,
    ERR_DT_ZERO = 1,        // integrator_muscl_hancock.c:545-549
    ERR_HLLC_SPEED = 2,     // riemann_solver_hllc.c:372-375
    ERR_EXACT_BRACKET = 3,  // riemann_solver_exact.c:702-707
    ERR_EXACT_BISECT = 4,   // riemann_solver_exact.c:745-754
};

// ---------------------------------------------------------------------------
// slope limiters: slope_limiter.c:49-91
template <int LIM>
__device__ __forceinline__ double limit(double a, double b)
{
    if (a * b <= 0.0) return 0.0;
    if constexpr (LIM == MINMOD) {
        return (fabs(a) < fabs(b)) ? a : b;
    } else if constexpr (LIM == VAN_LEER) {
        return 2.0 * a * b / (a + b + 1e-10);
    } else {
        if (a > 0) return fmin(0.5 * (a + b), 2.0 * fmin(fabs(a), fabs(b)));
        return -fmin(-0.5 * (a + b), 2.0 * fmin(fabs(a), fabs(b)));
    }
}

__device__ __forceinline__ double limit_rt(int lim, double a, double b)
{
    switch (lim) {
    case MINMOD: return limit<MINMOD>(a, b);
    case VAN_LEER: return limit<VAN_LEER>(a, b);
    default: return limit<MC>(a, b);
    }
}

// utils.c:23-26
__device__ __forceinline__ double sound_speed(double gamma, double rho, double p)
{
    return sqrt(gamma * p / rho);
}

// riemann_solver.c:144-152
__device__ __forceinline__ double coeff_A(double gamma, double rho) { return 2.0 / ((gamma + 1.0) * rho); }
__device__ __forceinline__ double coeff_B(double gamma, double p) { return p * ((gamma - 1.0) / (gamma + 1.0)); }

// riemann_solver.c:154-224: PVRS / two-rarefaction / two-shock pressure guess
__device__ __forceinline__ double guess_p(double gamma, double rl, double ul, double pl, double al,
                                          double rr, double ur, double pr, double ar, double tol)
{
    const double plo = (pl > pr) ? pr : pl;
    const double phi = (pl > pr) ? pl : pr;
    const double ppv = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (al + ar);
    double guess;
    if (phi / plo <= 2.0 && (plo <= ppv && ppv <= phi)) {
        guess = ppv;
    } else if (ppv < plo) {
        const double h = 0.5 * (gamma - 1.0);
        guess = (al + ar - h * (ur - ul)) / (al / (pow(pl, h / gamma)) + ar / (pow(pr, h / gamma)));
        guess = pow(guess, gamma / h);
    } else {
        const double gl = sqrt(coeff_A(gamma, rl) / (ppv + coeff_B(gamma, pl)));
        const double gr = sqrt(coeff_A(gamma, rr) / (ppv + coeff_B(gamma, pr)));
        guess = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
    }
    if (guess < tol) guess = tol;
    return guess;
}

// ---------------------------------------------------------------------------
// Vacuum sampling: riemann_solver.c:239-733.  NF = 3 (rho,u,p) or 4
// (rho,u,v,p); the pressure is the last component.  The reference hands the
// vacuum side's own state to the sampler when one side is vacuum (435-446)
// and measures the vacuum-generation fronts without the velocities (384-385).
template <int NF>
__device__ __forceinline__ void vac_right(double gamma, const double* w, double a, double speed, double* sol)
{
    const double rho = w[0], u = w[1], p = w[NF - 1];
    if (speed <= (u - a)) {
        for (int k = 0; k < NF; k++) sol[k] = w[k];
        return;
    }
    const double front = u + 2.0 * a / (gamma - 1.0);
    if (speed < front) {
        double t = (2.0 / (gamma + 1.0) + (u - speed) * (gamma - 1.0) / ((gamma + 1.0) * a));
        t = pow(t, 2.0 / (gamma - 1.0));
        sol[0] = rho * t;
        sol[1] = (a + 0.5 * (gamma - 1.0) * u + speed) * 2.0 / (gamma + 1.0);
        if (NF == 4) sol[2] = w[2];
        sol[NF - 1] = p * pow(t, gamma);
    } else {
        for (int k = 0; k < NF; k++) sol[k] = 0.0;
    }
}

template <int NF>
__device__ __forceinline__ void vac_left(double gamma, const double* w, double a, double speed, double* sol)
{
    const double rho = w[0], u = w[1], p = w[NF - 1];
    if (speed >= (u + a)) {
        for (int k = 0; k < NF; k++) sol[k] = w[k];
        return;
    }
    const double front = u - 2.0 * a / (gamma - 1.0);
    if (speed > front) {
        double t = (2.0 / (gamma + 1.0) - (u - speed) * (gamma - 1.0) / ((gamma + 1.0) * a));
        t = pow(t, 2.0 / (gamma - 1.0));
        sol[0] = rho * t;
        sol[1] = (-a + 0.5 * (gamma - 1.0) * u + speed) * 2.0 / (gamma + 1.0);
        if (NF == 4) sol[2] = w[2];
        sol[NF - 1] = p * pow(t, gamma);
    } else {
        for (int k = 0; k < NF; k++) sol[k] = 0.0;
    }
}

template <int NF>
__device__ __noinline__ void vacuum(double gamma, const double* wl, double al, const double* wr, double ar,
                                    double speed, double* sol)
{
    const double rl = wl[0], rr = wr[0];
    if (rl <= 0.0 && rr <= 0.0) {
        for (int k = 0; k < NF; k++) sol[k] = 0.0;
    } else if (rl <= 0.0) {
        vac_right<NF>(gamma, wl, al, speed, sol);
    } else if (rr <= 0.0) {
        vac_left<NF>(gamma, wr, ar, speed, sol);
    } else {
        const double fl = 2.0 * al / (gamma - 1.0);
        const double fr = 2.0 * ar / (gamma - 1.0);
        if (speed <= fl) {
            vac_right<NF>(gamma, wl, al, speed, sol);
        } else if (fr <= speed) {
            vac_left<NF>(gamma, wr, ar, speed, sol);
        } else {
            for (int k = 0; k < NF; k++) sol[k] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// HLLC: riemann_solver_hllc.c:30-376.  State order (rho, u_n, u_t, p); the
// flux is (mass, normal momentum, transverse momentum, energy).  NF = 3 drops
// the transverse component (1D).  Returns ERR_NONE or ERR_HLLC_SPEED.
__device__ __forceinline__ double hllc_q(double gamma, double pk, double ps)
{
    if (ps <= pk) return 1.0;
    return sqrt(1.0 + (0.5 + 0.5 / gamma) * (ps / pk - 1.0));
}

template <int NF>
__device__ __forceinline__ int flux_hllc(double gamma, double tol, const double* wl, const double* wr, double* f)
{
    const double rl = wl[0], ul = wl[1], pl = wl[NF - 1];
    const double rr = wr[0], ur = wr[1], pr = wr[NF - 1];
    const double vl = (NF == 4) ? wl[2] : 0.0;
    const double vr = (NF == 4) ? wr[2] : 0.0;
    const double al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0;

    if (rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul)) {
        double s[NF];
        vacuum<NF>(gamma, wl, al, wr, ar, 0.0, s);
        if constexpr (NF == 4) {
            const double e = s[0] * (0.5 * (s[1] * s[1] + s[2] * s[2]) + s[3] / (s[0] * (gamma - 1.0)));
            f[0] = s[0] * s[1];
            f[1] = s[0] * s[1] * s[1] + s[3];
            f[2] = s[0] * s[1] * s[2];
            f[3] = s[1] * (e + s[3]);
        } else {
            const double e = s[0] * (0.5 * s[1] * s[1] + s[2] / (s[0] * (gamma - 1.0)));
            f[0] = s[0] * s[1];
            f[1] = s[0] * s[1] * s[1] + s[2];
            f[2] = s[1] * (e + s[2]);
        }
        return ERR_NONE;
    }

    const double ps = guess_p(gamma, rl, ul, pl, al, rr, ur, pr, ar, tol);
    const double sl = ul - al * hllc_q(gamma, pl, ps);
    const double sr = ur + ar * hllc_q(gamma, pr, ps);
    const double sm = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur));

    // Which state feeds the flux, and whether the star correction applies.
    // The four regions of hllc.c:302-371 are tested in the same order.
    int region;
    if (0.0 <= sl) region = 0;
    else if (sl <= 0.0 && 0.0 <= sm) region = 1;
    else if (sm <= 0.0 && 0.0 <= sr) region = 2;
    else if (sr <= 0.0) region = 3;
    else return ERR_HLLC_SPEED;

    const bool left = region <= 1;
    const double r = left ? rl : rr;
    const double u = left ? ul : ur;
    const double v = left ? vl : vr;
    const double p = left ? pl : pr;
    const double s = left ? sl : sr;
    double e;
    if constexpr (NF == 4) e = r * (0.5 * (u * u + v * v) + p / (r * (gamma - 1.0)));
    else e = r * (0.5 * u * u + p / (r * (gamma - 1.0)));

    const double fm = r * u;
    const double fu = fm * u + p;  // == rho*u*u + p in the one-sided regions
    const double fe = u * (e + p);
    if (region == 0 || region == 3) {
        f[0] = fm;
        f[1] = fu;
        if constexpr (NF == 4) { f[2] = fm * v; f[3] = fe; }
        else f[2] = fe;
        return ERR_NONE;
    }
    const double rs = r * (s - u) / (s - sm);
    const double mus = rs * sm;
    const double es = rs * (e / r + (sm - u) * (sm + p / (r * (s - u))));
    f[0] = fm + s * (rs - r);
    f[1] = fu + s * (mus - r * u);
    if constexpr (NF == 4) {
        const double mvs = rs * v;
        f[2] = fm * v + s * (mvs - r * v);
        f[3] = fe + s * (es - e);
    } else {
        f[2] = fe + s * (es - e);
    }
    return ERR_NONE;
}

// ---------------------------------------------------------------------------
// Exact Riemann solver: riemann_solver_exact.c:332-1080 (NEWTON 100 /
// BRACKET 100 / BISECTION 5000 iterations, lines 26-28).
__device__ __forceinline__ double wave_f(double gamma, double rho, double p, double a, double ps)
{
    if (ps > p) return (ps - p) * sqrt(coeff_A(gamma, rho) / (ps + coeff_B(gamma, p)));
    return 2.0 * a * (pow((ps / p), (gamma - 1.0) / (2.0 * gamma)) - 1.0) / (gamma - 1.0);
}

// exact.c:539-565 -- the shock branch's (B - p*) is kept as written (553)
__device__ __forceinline__ double wave_df(double gamma, double rho, double p, double a, double ps)
{
    if (ps > p) {
        const double A = coeff_A(gamma, rho);
        const double B = coeff_B(gamma, p);
        return (1.0 - 0.5 * (ps - p) / (B - ps)) * sqrt(A / (B + ps));
    }
    return pow(ps / p, -0.5 * (gamma + 1.0) / gamma) / (p * a);
}

__device__ __forceinline__ double pressure_fn(double gamma, const double* wl, double al, const double* wr,
                                              double ar, double ps)
{
    return wave_f(gamma, wl[0], wl[2], al, ps) + wave_f(gamma, wr[0], wr[2], ar, ps) + (wr[1] - wl[1]);
}

__device__ __noinline__ int solve_p_star(double gamma, const double* wl, double al, const double* wr, double ar,
                                         double tol, double* p_star)
{
    const double pg = guess_p(gamma, wl[0], wl[1], wl[2], al, wr[0], wr[1], wr[2], ar, tol);
    bool bracketed = false;
    double p_hi = pg;
    double f_hi = pressure_fn(gamma, wl, al, wr, ar, p_hi);
    double p_lo = 0.0, f_lo = 0.0;
    double p0 = pg;
    for (int it = 0; it < 100; it++) {
        const double f = pressure_fn(gamma, wl, al, wr, ar, p0);
        const double df = wave_df(gamma, wl[0], wl[2], al, p0) + wave_df(gamma, wr[0], wr[2], ar, p0);
        if (!bracketed && f * f_hi < 0.0) {
            bracketed = true;
            f_lo = f;
            p_lo = p0;
        }
        const double p1 = p0 - f / df;
        if (p1 < 0.0) break;
        if (2.0 * fabs(p1 - p0) / (p1 + p0) < tol) {
            *p_star = p1;
            return ERR_NONE;
        }
        p0 = p1;
    }
    if (!bracketed) {
        p_lo = 0.0;
        f_lo = pressure_fn(gamma, wl, al, wr, ar, p_lo);
        if (f_lo * f_hi >= 0.0) {
            p_hi *= 10.0;
            const double dp = (p_hi - p_lo) / 100;
            for (int k = 0; k < 100; k++) {
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
            if (f_lo * f_hi >= 0.0) return ERR_EXACT_BRACKET;
        }
    }
    int count = 0;
    for (;;) {
        const double pm = 0.5 * (p_lo + p_hi);
        if (p_hi - p_lo < tol) {
            *p_star = pm;
            return ERR_NONE;
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
        if (count > 5000) return ERR_EXACT_BISECT;
    }
}

// exact.c:765-974 (left), 817-1080 (right; note its own operand orders)
__device__ __forceinline__ void sample_left(double gamma, const double* w, double a, double us, double ps,
                                            double speed, double* sol)
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
    if (speed < u - a) {
        sol[0] = rho; sol[1] = u; sol[2] = p;
        return;
    }
    const double a_star = a * pow(ps / p, (gamma - 1.0) / (2.0 * gamma));
    if (speed < us - a_star) {
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

__device__ __forceinline__ void sample_right(double gamma, const double* w, double a, double us, double ps,
                                             double speed, double* sol)
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
    if (speed > u + a) {
        sol[0] = rho; sol[1] = u; sol[2] = p;
        return;
    }
    const double a_star = a * pow(ps / p, (gamma - 1.0) / (2.0 * gamma));
    if (speed > us + a_star) {
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

// exact.c:332-469 at a given x/t; *us_out = u* (NaN on the vacuum branch)
__device__ __noinline__ int exact_sample(double gamma, double tol, const double* wl, const double* wr,
                                         double speed, double* sol, double* us_out)
{
    const double al = (wl[0] > 0.0) ? sound_speed(gamma, wl[0], wl[2]) : 0.0;
    const double ar = (wr[0] > 0.0) ? sound_speed(gamma, wr[0], wr[2]) : 0.0;
    *us_out = __builtin_nan("");
    if (wl[0] <= 0.0 || wr[0] <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (wr[1] - wl[1])) {
        vacuum<3>(gamma, wl, al, wr, ar, speed, sol);
        return ERR_NONE;
    }
    double ps;
    const int rc = solve_p_star(gamma, wl, al, wr, ar, tol, &ps);
    if (rc != ERR_NONE) return rc;
    const double us = 0.5 * (wl[1] + wr[1]) + 0.5 * (wave_f(gamma, wr[0], wr[2], ar, ps) - wave_f(gamma, wl[0], wl[2], al, ps));
    *us_out = us;
    if (speed < us) sample_left(gamma, wl, al, us, ps, speed, sol);
    else sample_right(gamma, wr, ar, us, ps, speed, sol);
    return ERR_NONE;
}

// Exact Godunov flux at x/t = 0.  NF = 3: exact.c:471-510.  NF = 4: the
// normal components are computed exactly as in 1D; the transverse velocity is
// advected by the contact (v_L when 0 < u*, the "speed < u_star" side of
// exact.c:431, else v_R); on the vacuum branch the 2D vacuum sampler supplies
// it.  At v = 0 the 2D flux equals the 1D flux bitwise.
template <int NF>
__device__ __forceinline__ int flux_exact(double gamma, double tol, const double* wl, const double* wr, double* f)
{
    const double l3[3] = {wl[0], wl[1], wl[NF - 1]};
    const double r3[3] = {wr[0], wr[1], wr[NF - 1]};
    double s[3] = {__builtin_nan(""), __builtin_nan(""), __builtin_nan("")};
    double us;
    const int rc = exact_sample(gamma, tol, l3, r3, 0.0, s, &us);
    f[0] = s[0] * s[1];
    f[1] = s[0] * s[1] * s[1] + s[2];
    if constexpr (NF == 4) {
        double v;
        if (isnan(us)) {
            const double al = (wl[0] > 0.0) ? sound_speed(gamma, wl[0], wl[3]) : 0.0;
            const double ar = (wr[0] > 0.0) ? sound_speed(gamma, wr[0], wr[3]) : 0.0;
            double s4[4];
            vacuum<4>(gamma, wl, al, wr, ar, 0.0, s4);
            v = s4[2];
        } else {
            v = (0.0 < us) ? wl[2] : wr[2];
        }
        f[2] = f[0] * v;
        f[3] = (s[2] * (gamma / (gamma - 1.0)) + (0.5 * s[0] * s[1] * s[1] + 0.5 * s[0] * v * v)) * s[1];
    } else {
        f[2] = (s[2] * (gamma / (gamma - 1.0)) + 0.5 * s[0] * s[1] * s[1]) * s[1];
    }
    return rc;
}

template <int SOLVER, int NF>
__device__ __forceinline__ int flux(double gamma, double tol, const double* wl, const double* wr, double* f)
{
    if constexpr (SOLVER == HLLC) return flux_hllc<NF>(gamma, tol, wl, wr, f);
    else return flux_exact<NF>(gamma, tol, wl, wr, f);
}

// ---------------------------------------------------------------------------
// Half-step predictor (integrator_muscl_hancock.c:692-740 x-sweep,
// 977-1025 y-sweep, 202-232 1D).  s = (rho, u, v, p) of the cell, d = limited
// slopes, c = dt*A/V, a = sound speed.  lo = state at the cell's low face,
// hi = state at its high face.  The y-sweep's density row multiplies the
// x-velocity slope (977-982, 1002-1007); that form is kept.
__device__ __forceinline__ void predict_x(const double* s, const double* d, double c, double a, double* lo, double* hi)
{
    const double up = 1.0 + c * s[1];
    const double um = 1.0 - c * s[1];
    lo[0] = s[0] + -0.5 * (up * d[0] + c * s[0] * d[1] + 0.0 + 0.0);
    lo[1] = s[1] + -0.5 * (0.0 + up * d[1] + 0.0 + (c / s[0]) * d[3]);
    lo[2] = s[2] + -0.5 * (0.0 + 0.0 + up * d[2] + 0.0);
    lo[3] = s[3] + -0.5 * (0.0 + c * s[0] * a * a * d[1] + 0.0 + up * d[3]);
    hi[0] = s[0] + 0.5 * (um * d[0] - c * s[0] * d[1] + 0.0 + 0.0);
    hi[1] = s[1] + 0.5 * (0.0 + um * d[1] + 0.0 - (c / s[0]) * d[3]);
    hi[2] = s[2] + 0.5 * (0.0 + 0.0 + um * d[2] + 0.0);
    hi[3] = s[3] + 0.5 * (0.0 - c * s[0] * a * a * d[1] + 0.0 + um * d[3]);
}

__device__ __forceinline__ void predict_y(const double* s, const double* d, double c, double a, double* lo, double* hi)
{
    const double vp = 1.0 + c * s[2];
    const double vm = 1.0 - c * s[2];
    lo[0] = s[0] + -0.5 * (vp * d[0] + c * s[0] * d[1] + 0.0 + 0.0);
    lo[1] = s[1] + -0.5 * (0.0 + vp * d[1] + 0.0 + 0.0);
    lo[2] = s[2] + -0.5 * (0.0 + 0.0 + vp * d[2] + (c / s[0]) * d[3]);
    lo[3] = s[3] + -0.5 * (0.0 + 0.0 + c * s[0] * a * a * d[2] + vp * d[3]);
    hi[0] = s[0] + 0.5 * (vm * d[0] - c * s[0] * d[1] + 0.0 + 0.0);
    hi[1] = s[1] + 0.5 * (0.0 + vm * d[1] + 0.0 + 0.0);
    hi[2] = s[2] + 0.5 * (0.0 + 0.0 + vm * d[2] - (c / s[0]) * d[3]);
    hi[3] = s[3] + 0.5 * (0.0 + 0.0 - c * s[0] * a * a * d[2] + vm * d[3]);
}

__device__ __forceinline__ void predict_1d(const double* s, const double* d, double c, double a, double* lo, double* hi)
{
    lo[0] = s[0] + -0.5 * ((1.0 + c * s[1]) * d[0] + c * s[0] * d[1] + 0.0);
    lo[1] = s[1] + -0.5 * (0.0 + (1.0 + c * s[1]) * d[1] + (c / s[0]) * d[2]);
    lo[2] = s[2] + -0.5 * (0.0 + c * s[0] * a * a * d[1] + (1.0 + c * s[1]) * d[2]);
    hi[0] = s[0] + 0.5 * ((1.0 - c * s[1]) * d[0] - c * s[0] * d[1] + 0.0);
    hi[1] = s[1] + 0.5 * (0.0 + (1.0 - c * s[1]) * d[1] - (c / s[0]) * d[2]);
    hi[2] = s[2] + 0.5 * (0.0 + -c * s[0] * a * a * d[1] + (1.0 - c * s[1]) * d[2]);
}

// ---------------------------------------------------------------------------
// Conversions: system.c:763-802 (2D c2p), 735-761 (1D c2p).
__device__ __forceinline__ void c2p_2d(double gamma, double vol, double m, double mx, double my, double e, double* w)
{
    w[0] = m / vol;
    w[1] = mx / m;
    w[2] = my / m;
    w[3] = (gamma - 1.0) * (e - 0.5 * m * (w[1] * w[1] + w[2] * w[2])) / vol;
}

__device__ __forceinline__ void c2p_1d(double gamma, double vol, double m, double mx, double e, double* w)
{
    w[0] = m / vol;
    w[1] = mx / m;
    w[2] = (gamma - 1.0) * (e - 0.5 * m * w[1] * w[1]) / vol;
}

// CFL reduction key: the maximum of non-negative doubles as an unsigned
// maximum of (bit pattern + 1); key 0 stands for "no finite candidate",
// i.e. the reference's initial -INFINITY (utils.c:75-92).
__device__ __forceinline__ unsigned long long smax_key(double v)
{
    return (v >= 0.0) ? (unsigned long long)__double_as_longlong(v) + 1ull : 0ull;
}

__host__ __device__ __forceinline__ double smax_from_key(unsigned long long k)
{
    if (k == 0ull) return -INFINITY;
    union { unsigned long long u; double d; } c;
    c.u = k - 1ull;
    return c.d;
}

}  // namespace hfv
