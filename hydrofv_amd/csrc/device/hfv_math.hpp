// hfv_math.hpp -- per-cell and per-interface physics for the MUSCL-Hancock
// update, as functions for gfx950 kernels.  They are also compiled for the
// host (HFV_HD), so that CPU tests can run the very same source.
//
// Every expression keeps the operand order of the reference C code it cites,
// and the library is compiled with -ffp-contract=off, so each result is the
// correctly rounded value of the same sequence of IEEE operations (divisions
// and square roots are correctly rounded on this target).  Powers are the
// only results whose last place may differ from glibc's: they all come from
// pow_pos (hfv_pow.hpp), which differs from glibc's pow on about 1 argument
// in 1000 (the device library pow on about 1 in 5).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define HFV_HD __host__ __device__

#include "hfv_pow.hpp"

namespace hfv {

// flags: the reference's numeric values (slope_limiter.h, riemann_solver.h, boundary.h)
enum Limiter : int { MINMOD = 1, VAN_LEER = 2, MC = 3 };
// In place of a limiter: no reconstruction, the faces take the cells' own
// states (integrator_godunov_first_order.c).
constexpr int FIRST_ORDER = 0;
enum Solver : int { EXACT = 1, HLLC = 2 };
enum Bc : int { BC_NONE = 1, BC_REFLECTIVE = 2, BC_TRANSMISSIVE = 3, BC_PERIODIC = 4, BC_CUSTOM = 5 };

// device-side failure kinds (mapped to the reference's error messages by the host)
enum ErrKind : int {
    ERR_NONE = 0,
    ERR_DT_ZERO = 1,        // integrator_muscl_hancock.c:545-549
    ERR_HLLC_SPEED = 2,     // riemann_solver_hllc.c:372-375
    ERR_EXACT_BRACKET = 3,  // riemann_solver_exact.c:702-707
    ERR_EXACT_BISECT = 4,   // riemann_solver_exact.c:745-754
    ERR_PENDING = -1,       // internal: the face needs the iterative exact solve (DIRECT_ONLY)
};

// ---------------------------------------------------------------------------
// slope limiters: slope_limiter.c:49-91
template <int LIM>
HFV_HD __forceinline__ double limit(double a, double b)
{
    if constexpr (LIM == MINMOD) {
        if (a * b <= 0.0) return 0.0;
        return (fabs(a) < fabs(b)) ? a : b;
    } else if constexpr (LIM == VAN_LEER) {
        if (a * b <= 0.0) return 0.0;
        return 2.0 * a * b / (a + b + 1e-10);
    } else {
        // Where a b > 0, a + b has a's sign and -0.5 (a + b) = 0.5 |a + b|
        // exactly when a < 0, so both branches of slope_limiter.c:83-90 are
        // a's sign on fmin(0.5 |a + b|, 2 fmin(|a|, |b|)): computed once, and
        // selected (no branch) with the a b <= 0 case's zero (:79-82)
        const double m = fmin(0.5 * fabs(a + b), 2.0 * fmin(fabs(a), fabs(b)));
        const double r = (a > 0) ? m : -m;
        return (a * b <= 0.0) ? 0.0 : r;
    }
}

HFV_HD __forceinline__ double limit_rt(int lim, double a, double b)
{
    switch (lim) {
    case MINMOD: return limit<MINMOD>(a, b);
    case VAN_LEER: return limit<VAN_LEER>(a, b);
    default: return limit<MC>(a, b);
    }
}

// ---------------------------------------------------------------------------
// Square roots of groups of arguments.  sqrt() of a double compiles for
// gfx950 to a reciprocal square root estimate refined by two Goldschmidt
// steps (ten instructions), wrapped in a scaling by 2^256 of arguments below
// 2^-767 and a fix-up that passes 0 and +inf through (about ten more).  For an
// argument in [2^-767, +inf) the wrapping leaves the refined value unchanged,
// so sqrt_core, the refinement alone in the same operation order, returns the
// bits of sqrt() there: the correctly rounded root.  sqrt_rn takes the core
// for a group of arguments when every argument on every lane of the
// wavefront lies in that range, and sqrt() for the whole group otherwise.
HFV_HD __forceinline__ bool sqrt_core_ok(double x) { return (x >= 0x1p-767) & (x < INFINITY); }

HFV_HD __forceinline__ double sqrt_core(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    const double s0 = x * y;
    const double h0 = y * 0.5;
    const double r0 = fma(-h0, s0, 0.5);
    const double h1 = fma(h0, r0, h0);
    const double s1 = fma(s0, r0, s0);
    const double d0 = fma(-s1, s1, x);
    const double s2 = fma(d0, h1, s1);
    const double d1 = fma(-s2, s2, x);
    return fma(d1, h1, s2);
#else
    return sqrt(x);  // host builds (CPU tests): correctly rounded as well
#endif
}

// ok on every active lane of the wavefront (host builds: ok)
HFV_HD __forceinline__ bool wave_all(bool ok)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(!ok) == 0;
#else
    return ok;
#endif
}

// x[k] = sqrt(x[k]) for k < K, bit for bit
template <int K>
HFV_HD __forceinline__ void sqrt_rn(double (&x)[K])
{
    bool ok = true;
#pragma unroll
    for (int k = 0; k < K; k++) ok = ok & sqrt_core_ok(x[k]);
    if (wave_all(ok)) {
#pragma unroll
        for (int k = 0; k < K; k++) x[k] = sqrt_core(x[k]);
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) x[k] = sqrt(x[k]);
    }
}

// ---------------------------------------------------------------------------
// Divisions that share a reciprocal.  a / b of doubles compiles for gfx950
// to: d = v_div_scale(b), n = v_div_scale(a), r = v_rcp(d) refined by two
// Newton steps r = fma(r, fma(-d, r, 1), r), q = n r, the residual
// fma(-d, q, n), then v_div_fmas (fma(residual, r, q), rescaled when the
// scalings set their flag) and v_div_fixup (special operands, overflow,
// underflow).  For finite a, b with |a|, |b| in [2^-300, 2^300] the scalings
// return their operands with the flag clear, and the fix-up returns the
// quotient as it is (none of its cases arises, and a nonzero finite
// quotient already has the right sign): rcp_core(b) followed by
// div_core(a, b, r), the same refinement in the same order, gives the bits
// of a / b there, and one reciprocal serves every division by b.
// div_core_ok(x): x lies in that range (false for 0, infinities and NaN).
// Host builds divide.
HFV_HD __forceinline__ bool div_core_ok(double x)
{
    const double ax = fabs(x);
    return (ax >= 0x1p-300) & (ax <= 0x1p300);
}

HFV_HD __forceinline__ double rcp_core(double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    return r;
#else
    return b;  // unused: div_core divides
#endif
}

HFV_HD __forceinline__ double div_core(double a, double b, double r)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
#else
    (void)r;
    return a / b;
#endif
}

// utils.c:23-26
HFV_HD __forceinline__ double sound_speed(double gamma, double rho, double p)
{
    return sqrt(gamma * p / rho);
}

// sound_speed of a lane's cell, with the root taken by sqrt_rn
HFV_HD __forceinline__ double sound_speed_rn(double gamma, double rho, double p)
{
    double x[1] = {gamma * p / rho};
    sqrt_rn<1>(x);
    return x[0];
}

// riemann_solver.c:144-152
HFV_HD __forceinline__ double coeff_A(double gamma, double rho) { return 2.0 / ((gamma + 1.0) * rho); }
HFV_HD __forceinline__ double coeff_B(double gamma, double p) { return p * ((gamma - 1.0) / (gamma + 1.0)); }

// riemann_solver.c:154-224: PVRS / two-rarefaction / two-shock pressure guess,
// in two parts.  guess_select returns true on the two-rarefaction branch
// (188-199) with *g = its numerator al + ar - h*(ur - ul); otherwise *g is the
// PVRS or two-shock guess before the clamp to tol.
HFV_HD __forceinline__ bool guess_select(double gamma, double rl, double ul, double pl, double al,
                                             double rr, double ur, double pr, double ar, double* g)
{
    const double plo = (pl > pr) ? pr : pl;
    const double phi = (pl > pr) ? pl : pr;
    const double ppv = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (al + ar);
    if (phi / plo <= 2.0 && (plo <= ppv && ppv <= phi)) {
        *g = ppv;
        return false;
    }
    if (ppv < plo) {
        const double h = 0.5 * (gamma - 1.0);
        *g = al + ar - h * (ur - ul);
        return true;
    }
    const double gl = sqrt(coeff_A(gamma, rl) / (ppv + coeff_B(gamma, pl)));
    const double gr = sqrt(coeff_A(gamma, rr) / (ppv + coeff_B(gamma, pr)));
    *g = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
    return false;
}

// Two-rarefaction guess from its numerator and pow(pl, h/gamma), pow(pr, h/gamma);
// the caller raises the result to gamma/h (riemann_solver.c:191-199).
HFV_HD __forceinline__ double trrs_base(double num, double al, double ar, double pow_l, double pow_r)
{
    return num / (al / (pow_l) + ar / (pow_r));
}

HFV_HD __forceinline__ double guess_p(double gamma, double rl, double ul, double pl, double al,
                                      double rr, double ur, double pr, double ar, double tol)
{
    double guess;
    if (guess_select(gamma, rl, ul, pl, al, rr, ur, pr, ar, &guess)) {
        const double h = 0.5 * (gamma - 1.0);
        guess = trrs_base(guess, al, ar, pow_pos(pl, h / gamma), pow_pos(pr, h / gamma));
        guess = pow_pos(guess, gamma / h);
    }
    if (guess < tol) guess = tol;
    return guess;
}

// ---------------------------------------------------------------------------
// Vacuum sampling: riemann_solver.c:239-733.  NF = 3 (rho,u,p) or 4
// (rho,u,v,p); with NF = 3 the transverse velocity v is carried as 0 and
// ignored.  The reference hands the vacuum side's own state to the sampler
// when one side is vacuum (435-446) and measures the vacuum-generation fronts
// without the velocities (384-385).
struct Prim {
    double r, u, v, p;
};

// The two rarefaction-into-vacuum samplers as one path (riemann_solver.c:
// 239-286 for a vacuum on the right, 288-335 on the left).  sigma = +1 for
// the right-vacuum case; the left case's expressions are the right case's
// with a, the front offset and the (u - speed) term negated, and negation is
// exact, so both give the bits of the reference's separate branches.
HFV_HD __forceinline__ Prim vac_fan(double gamma, Prim w, double a, double speed, bool right)
{
    const bool outside = right ? (speed <= (w.u - a)) : (speed >= (w.u + a));
    if (outside) return w;
    const double off = 2.0 * a / (gamma - 1.0);
    const double front = right ? w.u + off : w.u - off;
    const bool fan = right ? (speed < front) : (speed > front);
    if (!fan) return Prim{0.0, 0.0, 0.0, 0.0};
    const double x = (w.u - speed) * (gamma - 1.0) / ((gamma + 1.0) * a);
    double t = (2.0 / (gamma + 1.0) + (right ? x : -x));
    t = pow_pos(t, 2.0 / (gamma - 1.0));
    Prim o;
    o.r = w.r * t;
    o.u = ((right ? a : -a) + 0.5 * (gamma - 1.0) * w.u + speed) * 2.0 / (gamma + 1.0);
    o.v = w.v;
    o.p = w.p * pow_pos(t, gamma);
    return o;
}

HFV_HD __forceinline__ Prim vacuum(double gamma, double rl, double ul, double vl, double pl, double al, double rr,
                                       double ur, double vr, double pr, double ar, double speed)
{
    // riemann_solver.c:377-481: which fan (if any) x/t = speed falls in
    bool zero = false, right = true;
    if (rl <= 0.0 && rr <= 0.0) {
        zero = true;
    } else if (rl <= 0.0) {
        right = true;
    } else if (rr <= 0.0) {
        right = false;
    } else {
        const double fl = 2.0 * al / (gamma - 1.0);
        const double fr = 2.0 * ar / (gamma - 1.0);
        if (speed <= fl) right = true;
        else if (fr <= speed) right = false;
        else zero = true;
    }
    if (zero) return Prim{0.0, 0.0, 0.0, 0.0};
    const Prim w = right ? Prim{rl, ul, vl, pl} : Prim{rr, ur, vr, pr};
    return vac_fan(gamma, w, right ? al : ar, speed, right);
}

// vacuum() on NF-component state arrays
template <int NF>
HFV_HD __forceinline__ Prim vacuum_of(double gamma, const double* wl, double al, const double* wr, double ar,
                                          double speed)
{
    const double vl = (NF == 4) ? wl[2] : 0.0;
    const double vr = (NF == 4) ? wr[2] : 0.0;
    return vacuum(gamma, wl[0], wl[1], vl, wl[NF - 1], al, wr[0], wr[1], vr, wr[NF - 1], ar, speed);
}

// ---------------------------------------------------------------------------
// HLLC: riemann_solver_hllc.c:30-376.  State order (rho, u_n, u_t, p); the
// flux is (mass, normal momentum, transverse momentum, energy).  NF = 3 drops
// the transverse component (1D).  Returns ERR_NONE or ERR_HLLC_SPEED.
HFV_HD __forceinline__ double hllc_q(double gamma, double pk, double ps)
{
    if (ps <= pk) return 1.0;
    return sqrt(1.0 + (0.5 + 0.5 / gamma) * (ps / pk - 1.0));
}

// Flux of the vacuum branch (hllc.c:42-66 via riemann_solver.c:377-481).
template <int NF>
HFV_HD __forceinline__ void hllc_vacuum_flux(double gamma, const double* wl, double al, const double* wr,
                                                 double ar, double* f)
{
    const Prim v = vacuum_of<NF>(gamma, wl, al, wr, ar, 0.0);
    if constexpr (NF == 4) {
        const double e = v.r * (0.5 * (v.u * v.u + v.v * v.v) + v.p / (v.r * (gamma - 1.0)));
        f[0] = v.r * v.u;
        f[1] = v.r * v.u * v.u + v.p;
        f[2] = v.r * v.u * v.v;
        f[3] = v.u * (e + v.p);
    } else {
        const double e = v.r * (0.5 * v.u * v.u + v.p / (v.r * (gamma - 1.0)));
        f[0] = v.r * v.u;
        f[1] = v.r * v.u * v.u + v.p;
        f[2] = v.u * (e + v.p);
    }
}

// Wave speeds and the flux of the region x/t = 0 falls in, given the
// pressure guess ps (hllc.c:68-371).
template <int NF>
HFV_HD __forceinline__ int hllc_from_guess(double gamma, const double* wl, const double* wr, double al,
                                               double ar, double ps, double* f)
{
    const double rl = wl[0], ul = wl[1], pl = wl[NF - 1];
    const double rr = wr[0], ur = wr[1], pr = wr[NF - 1];
    const double vl = (NF == 4) ? wl[2] : 0.0;
    const double vr = (NF == 4) ? wr[2] : 0.0;
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

HFV_HD __forceinline__ bool hllc_is_vacuum(double gamma, double rl, double ul, double al, double rr, double ur,
                                               double ar)
{
    return rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul);
}

template <int NF>
HFV_HD __forceinline__ int flux_hllc(double gamma, double tol, const double* wl, const double* wr, double* f)
{
    const double rl = wl[0], ul = wl[1], pl = wl[NF - 1];
    const double rr = wr[0], ur = wr[1], pr = wr[NF - 1];
    const double al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0;
    if (hllc_is_vacuum(gamma, rl, ul, al, rr, ur, ar)) {
        hllc_vacuum_flux<NF>(gamma, wl, al, wr, ar, f);
        return ERR_NONE;
    }
    const double ps = guess_p(gamma, rl, ul, pl, al, rr, ur, pr, ar, tol);
    return hllc_from_guess<NF>(gamma, wl, wr, al, ar, ps, f);
}

// The 2D HLLC flux in two halves around the three pow() calls of the
// two-rarefaction guess (the only powers HLLC takes), so that a wavefront
// can evaluate the powers of all its lanes' faces together instead of every
// lane waiting through each face's calls (trrs_powers in mh2d_strip.hpp).
// hllc_begin, the powers and hllc_end perform the operations of
// flux_hllc<4> in the same order.
enum HllcKind : int { HLLC_READY = 0, HLLC_TRRS = 1, HLLC_VACUUM = 2 };

struct HllcFace {
    double wl[4], wr[4];  // (rho, u_n, u_t, p) left and right
    double al, ar;
    double ps;            // pressure guess; HLLC_TRRS: the guess's numerator until the powers are taken
    int kind;
};

// hllc_begin_faces and hllc_end_regular_faces evaluate every candidate of a
// branch (pressure-guess kind, q_K, wave region, star state) and select, for
// N faces at once, so that the compiler can interleave the independent
// evaluations of a lane's faces; each selected value comes from the same
// expression as in flux_hllc, so the bits are the same.  F[i].wl and F[i].wr
// are set by the caller.
// The guess of hllc_begin_faces in three phases, so that a caller can take
// the two-shock candidates of only the faces that use them (strip_step packs
// them across the wavefront): hllc_guess_begin (sound speeds, vacuum test,
// PVRS value and which guess applies), tsrs_guess for each face that needs
// it, hllc_guess_end (the choice).
template <int N>
struct HllcGuess {
    double ppv[N], num[N], tsrs[N];
    bool vac[N], pvrs[N], trrs[N];
};

template <int N>
HFV_HD __forceinline__ void hllc_guess_begin(double gamma, HllcFace (&F)[N], HllcGuess<N>& G)
{
    const double h = 0.5 * (gamma - 1.0);
    // the faces' sound speeds (utils.c:23-26), roots taken together
    double a2[2 * N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        a2[2 * i] = gamma * F[i].wl[3] / F[i].wl[0];
        a2[2 * i + 1] = gamma * F[i].wr[3] / F[i].wr[0];
    }
    sqrt_rn<2 * N>(a2);
    // hllc_is_vacuum's (a_L + a_R) 2 / (gamma + 1), its divisions sharing the
    // reciprocal of gamma + 1 when every lane's operands allow (div_core),
    // and evaluated for every face: the test below takes no branch
    double cv[N];
    bool cv_ok = div_core_ok(gamma + 1.0);
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double al = (F[i].wl[0] > 0.0) ? a2[2 * i] : 0.0;
        const double ar = (F[i].wr[0] > 0.0) ? a2[2 * i + 1] : 0.0;
        cv[i] = (al + ar) * 2.0;
        cv_ok = cv_ok & div_core_ok(cv[i]);
    }
    if (wave_all(cv_ok)) {
        const double rg = rcp_core(gamma + 1.0);
#pragma unroll
        for (int i = 0; i < N; i++) cv[i] = div_core(cv[i], gamma + 1.0, rg);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) cv[i] = cv[i] / (gamma + 1.0);
    }
    // The guess's test p_max / p_min <= 2 (riemann_solver.c:185).  The
    // quotient rounds to at most 2 exactly when p_max <= (2 + 2^-52) p_min
    // (2 + 2^-52, the midpoint to the next double, rounds to the even 2), and
    // for 2^-970 <= p_min, p_max <= 2^1020 the test p_max - 2 p_min <= 2^-52
    // p_min is that comparison without rounding: 2 p_min and 2^-52 p_min are
    // exact, and so is the difference where p_min <= p_max <= 4 p_min
    // (Sterbenz), which holds whenever it can be small enough.  A wavefront
    // with any other pair divides.
    bool ratio2[N];
    bool ratio_ok = true;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double pl = F[i].wl[3], pr = F[i].wr[3];
        const double plo = (pl > pr) ? pr : pl;
        const double phi = (pl > pr) ? pl : pr;
        ratio_ok = ratio_ok & (plo >= 0x1p-970) & (phi <= 0x1p1020);
    }
    if (wave_all(ratio_ok)) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            const double pl = F[i].wl[3], pr = F[i].wr[3];
            const double plo = (pl > pr) ? pr : pl;
            const double phi = (pl > pr) ? pl : pr;
            ratio2[i] = (phi - 2.0 * plo) <= 0x1p-52 * plo;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) {
            const double pl = F[i].wl[3], pr = F[i].wr[3];
            const double plo = (pl > pr) ? pr : pl;
            const double phi = (pl > pr) ? pl : pr;
            ratio2[i] = phi / plo <= 2.0;
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double rl = F[i].wl[0], ul = F[i].wl[1], pl = F[i].wl[3];
        const double rr = F[i].wr[0], ur = F[i].wr[1], pr = F[i].wr[3];
        const double al = (rl > 0.0) ? a2[2 * i] : 0.0;
        const double ar = (rr > 0.0) ? a2[2 * i + 1] : 0.0;
        F[i].al = al;
        F[i].ar = ar;
        G.vac[i] = (rl <= 0.0) | (rr <= 0.0) | (cv[i] <= (ur - ul));  // hllc_is_vacuum
        // riemann_solver.c:154-224 (as guess_select)
        const double plo = (pl > pr) ? pr : pl;
        const double phi = (pl > pr) ? pl : pr;
        G.ppv[i] = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (al + ar);
        G.pvrs[i] = ratio2[i] & (plo <= G.ppv[i]) & (G.ppv[i] <= phi);
        G.trrs[i] = !G.pvrs[i] & (G.ppv[i] < plo);
        G.num[i] = al + ar - h * (ur - ul);
        G.tsrs[i] = 0.0;
    }
}

// The two-shock guess (riemann_solver.c:201-215) from the face's pressures
// and densities, du = ur - ul and the PVRS value.
HFV_HD __forceinline__ double tsrs_guess(double gamma, double rl, double pl, double rr, double pr, double du,
                                         double ppv)
{
    double g2[2] = {coeff_A(gamma, rl) / (ppv + coeff_B(gamma, pl)), coeff_A(gamma, rr) / (ppv + coeff_B(gamma, pr))};
    sqrt_rn<2>(g2);
    const double gl = g2[0], gr = g2[1];
    return (gl * pl + gr * pr - du) / (gl + gr);
}

// faces that take the two-shock guess (not vacuum, not two-rarefaction, not PVRS)
template <int N>
HFV_HD __forceinline__ bool needs_tsrs(const HllcGuess<N>& G, int i)
{
    return !G.vac[i] & !G.trrs[i] & !G.pvrs[i];
}

template <int N>
HFV_HD __forceinline__ void hllc_guess_end(double tol, HllcFace (&F)[N], const HllcGuess<N>& G)
{
#pragma unroll
    for (int i = 0; i < N; i++) {
        double g = G.pvrs[i] ? G.ppv[i] : G.tsrs[i];
        g = (g < tol) ? tol : g;
        F[i].ps = G.vac[i] ? 0.0 : (G.trrs[i] ? G.num[i] : g);
        F[i].kind = G.vac[i] ? HLLC_VACUUM : (G.trrs[i] ? HLLC_TRRS : HLLC_READY);
    }
}

template <int N>
HFV_HD __forceinline__ void hllc_begin_faces(double gamma, double tol, HllcFace (&F)[N])
{
    HllcGuess<N> G;
    hllc_guess_begin<N>(gamma, F, G);
    // the two-shock candidates as a phase of their own
#pragma unroll
    for (int i = 0; i < N; i++)
        G.tsrs[i] = tsrs_guess(gamma, F[i].wl[0], F[i].wl[3], F[i].wr[0], F[i].wr[3], F[i].wr[1] - F[i].wl[1],
                               G.ppv[i]);
    hllc_guess_end<N>(tol, F, G);
}

HFV_HD __forceinline__ void hllc_begin(double gamma, double tol, const double* wl, const double* wr, HllcFace& F)
{
    HllcFace G[1];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        G[0].wl[k] = wl[k];
        G[0].wr[k] = wr[k];
    }
    hllc_begin_faces<1>(gamma, tol, G);
    F = G[0];
}

// The fluxes of hllc_from_guess<4> from ready faces, whatever their kind (a
// vacuum face's result is replaced by the caller): hllc.c:68-371.
template <int N>
HFV_HD __forceinline__ void hllc_end_regular_faces(double gamma, const HllcFace (&F)[N], double (&f)[N][4],
                                                   int (&err)[N])
{
    // hllc_q (hllc.c:30-40) on both sides
    const double kq = 0.5 + 0.5 / gamma;
    double ql[N], qr[N], q2[2 * N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double ps = F[i].ps, pl = F[i].wl[3], pr = F[i].wr[3];
        q2[2 * i] = 1.0 + kq * (ps / pl - 1.0);
        q2[2 * i + 1] = 1.0 + kq * (ps / pr - 1.0);
    }
    sqrt_rn<2 * N>(q2);
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double ps = F[i].ps, pl = F[i].wl[3], pr = F[i].wr[3];
        ql[i] = (ps <= pl) ? 1.0 : q2[2 * i];
        qr[i] = (ps <= pr) ? 1.0 : q2[2 * i + 1];
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double rl = F[i].wl[0], ul = F[i].wl[1], vl = F[i].wl[2], pl = F[i].wl[3];
        const double rr = F[i].wr[0], ur = F[i].wr[1], vr = F[i].wr[2], pr = F[i].wr[3];
        const double al = F[i].al, ar = F[i].ar;
        const double sl = ul - al * ql[i];
        const double sr = ur + ar * qr[i];
        const double sm = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur));
        // the regions of hllc.c:302-371, tested in the same order
        const bool r0 = 0.0 <= sl;
        const bool r1 = !r0 & (sl <= 0.0) & (0.0 <= sm);
        const bool r2 = !r0 & !r1 & (sm <= 0.0) & (0.0 <= sr);
        const bool r3 = !r0 & !r1 & !r2 & (sr <= 0.0);
        const bool left = r0 | r1;
        const bool star = r1 | r2;
        const double r = left ? rl : rr;
        const double u = left ? ul : ur;
        const double v = left ? vl : vr;
        const double p = left ? pl : pr;
        const double s = left ? sl : sr;
        const double e = r * (0.5 * (u * u + v * v) + p / (r * (gamma - 1.0)));
        const double fm = r * u;
        const double fu = fm * u + p;
        const double fe = u * (e + p);
        const double rs = r * (s - u) / (s - sm);
        const double mus = rs * sm;
        const double es = rs * (e / r + (sm - u) * (sm + p / (r * (s - u))));
        const double mvs = rs * v;
        f[i][0] = star ? fm + s * (rs - r) : fm;
        f[i][1] = star ? fu + s * (mus - r * u) : fu;
        f[i][2] = star ? fm * v + s * (mvs - r * v) : fm * v;
        f[i][3] = star ? fe + s * (es - e) : fe;
        err[i] = (r0 | r1 | r2 | r3) ? ERR_NONE : ERR_HLLC_SPEED;
    }
}

HFV_HD __forceinline__ int hllc_end_regular(double gamma, const HllcFace& F, double* f)
{
    const HllcFace G[1] = {F};
    double o[1][4];
    int e[1];
    hllc_end_regular_faces<1>(gamma, G, o, e);
#pragma unroll
    for (int k = 0; k < 4; k++) f[k] = o[0][k];
    return e[0];
}

// The same face set-up and flux with branches in place of the selects
// (hllc.c:68-371 and riemann_solver.c:154-224 as written): cheaper than the
// selects where the lanes of a wavefront agree on the branches, as in
// uniform regions; the bits are the same.
HFV_HD __forceinline__ void hllc_begin_branchy(double gamma, double tol, HllcFace& F)
{
    const double rl = F.wl[0], ul = F.wl[1], pl = F.wl[3];
    const double rr = F.wr[0], ur = F.wr[1], pr = F.wr[3];
    F.al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    F.ar = (rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0;
    F.ps = 0.0;
    if (hllc_is_vacuum(gamma, rl, ul, F.al, rr, ur, F.ar)) {
        F.kind = HLLC_VACUUM;
        return;
    }
    double g;
    if (guess_select(gamma, rl, ul, pl, F.al, rr, ur, pr, F.ar, &g)) {
        F.kind = HLLC_TRRS;
        F.ps = g;
        return;
    }
    F.kind = HLLC_READY;
    F.ps = (g < tol) ? tol : g;
}

HFV_HD __forceinline__ int hllc_end_branchy(double gamma, const HllcFace& F, double* f)
{
    if (F.kind == HLLC_VACUUM) {
        hllc_vacuum_flux<4>(gamma, F.wl, F.al, F.wr, F.ar, f);
        return ERR_NONE;
    }
    return hllc_from_guess<4>(gamma, F.wl, F.wr, F.al, F.ar, F.ps, f);
}

// After the first two powers of a two-rarefaction face (pzl = pow(pl, h/gamma),
// pzr = pow(pr, h/gamma)): the guess is p* = pow(X, gamma/h), clamped to tol,
// and HLLC uses it only through q_K (hllc_q), which is 1 whenever
// p* <= p_K.  When X >= 0 lies below min(pzl, pzr) by a relative 2^-40 and
// min(pl, pr) >= tol, p* (after the clamp) is below both pressures: the
// margin covers the last-place errors of the powers and of X, and
// (h/gamma)(gamma/h) != 1 in rounding, by orders of magnitude.  Then the
// last power cannot change the flux, the face is settled with ps =
// min(pl, pr) (q_L = q_R = 1, as with the reference's p*) and true is
// returned; otherwise *X receives the base of the last power.
HFV_HD __forceinline__ bool trrs_settle(double tol, HllcFace& F, double pzl, double pzr, double* X)
{
    const double x = trrs_base(F.ps, F.al, F.ar, pzl, pzr);
    const double pz_lo = (pzl < pzr) ? pzl : pzr;
    const double p_lo = (F.wl[3] > F.wr[3]) ? F.wr[3] : F.wl[3];
    if (x >= 0.0 && x < pz_lo * (1.0 - 0x1p-40) && p_lo >= tol) {
        F.ps = p_lo;
        F.kind = HLLC_READY;
        return true;
    }
    *X = x;
    return false;
}

// trrs_settle's decision without the powers, for a two-rarefaction face
// whose pressures are within 12% of each other.  With z = h/gamma,
// X = num / (al pl^-z + ar pr^-z) lies below plo^z exactly when
// num < a_lo + a_hi (plo/phi)^z, where a_lo is the sound speed on the side
// of the lower pressure plo and a_hi on the side of phi.  (plo/phi)^z =
// exp(z log(plo/phi)) is evaluated with log(plo/phi) = 2 atanh(s),
// s = (plo - phi)/(plo + phi) in [-1/16, 0] (plo - phi is exact there), by
// their Taylor series to s^13 and y^9 (y = z log(plo/phi), |y| < 0.07 for
// any gamma > 1): truncation below 2^-50 and rounding a few ulps, so the
// right-hand side is within 2^-47 of its value.  The test with the margin
// 1 - 2^-39 then implies X < (1 - 2^-40) plo^z in exact arithmetic.  The
// reference's guess (riemann_solver.c:191-199) computes X from num, which
// it forms by the same operations, with relative error below 4 ulps, and
// raises it to 1/z with pow's error below 1 ulp; as (1 - 2^-40)(1 + 4u) < 1,
// its p* is below plo for every gamma > 1, and with plo >= tol the clamp
// to tol keeps it there.  q_L = q_R = 1 then, as with ps = plo, and p*
// enters the flux nowhere else: the face is settled exactly as by
// trrs_settle, and no power needs to be taken.  F.ps holds num on entry.
HFV_HD __forceinline__ bool trrs_settle_direct(double gamma, double tol, HllcFace& F)
{
    const double pl = F.wl[3], pr = F.wr[3];
    const bool left_lo = !(pl > pr);
    const double plo = left_lo ? pl : pr, phi = left_lo ? pr : pl;
    const double alo = left_lo ? F.al : F.ar, ahi = left_lo ? F.ar : F.al;
    const double s = (plo - phi) / (plo + phi);
    const double t = s * s;
    double q = 1.0 / 13.0;
    q = fma(q, t, 1.0 / 11.0);
    q = fma(q, t, 1.0 / 9.0);
    q = fma(q, t, 1.0 / 7.0);
    q = fma(q, t, 1.0 / 5.0);
    q = fma(q, t, 1.0 / 3.0);
    q = fma(q, t, 1.0);
    const double y = ((0.5 * (gamma - 1.0)) / gamma) * (2.0 * s * q);
    double e = 1.0 / 362880.0;
    e = fma(e, y, 1.0 / 40320.0);
    e = fma(e, y, 1.0 / 5040.0);
    e = fma(e, y, 1.0 / 720.0);
    e = fma(e, y, 1.0 / 120.0);
    e = fma(e, y, 1.0 / 24.0);
    e = fma(e, y, 1.0 / 6.0);
    e = fma(e, y, 0.5);
    e = fma(e, y, 1.0);
    const double pz = fma(e, y, 1.0);  // (plo/phi)^z
    const double rhs = fma(ahi, pz, alo);
    const double num = F.ps;
    if ((s >= -0.0625) & (plo >= tol) & (num >= 0.0) & (alo >= 0.0) & (num < (1.0 - 0x1p-39) * rhs)) {
        F.ps = plo;
        F.kind = HLLC_READY;
        return true;
    }
    return false;
}

HFV_HD __forceinline__ int hllc_end(double gamma, const HllcFace& F, double* f)
{
    int e = hllc_end_regular(gamma, F, f);
    if (F.kind == HLLC_VACUUM) {
        hllc_vacuum_flux<4>(gamma, F.wl, F.al, F.wr, F.ar, f);
        e = ERR_NONE;
    }
    return e;
}

// ---------------------------------------------------------------------------
// Exact Riemann solver: riemann_solver_exact.c:332-1080 (NEWTON 100 /
// BRACKET 100 / BISECTION 5000 iterations, lines 26-28).
HFV_HD __forceinline__ double wave_f(double gamma, double rho, double p, double a, double ps)
{
    if (ps > p) return (ps - p) * sqrt(coeff_A(gamma, rho) / (ps + coeff_B(gamma, p)));
    return 2.0 * a * (pow_pos((ps / p), (gamma - 1.0) / (2.0 * gamma)) - 1.0) / (gamma - 1.0);
}

// The side states are (rho, u, p) in the exact solver (3-component order).
struct Side {
    double r, u, p, a;
};

// One side of the pressure function with its coefficients that do not
// depend on p*: A and B (exact.c:516-519, 543-546), 2a and p a
// (exact.c:527, 561): the reference's expressions, evaluated once per face
// instead of once per iteration, so the same bits.
struct WaveSide {
    double p, A, B, a2, pa;
};

HFV_HD __forceinline__ WaveSide wave_side(double gamma, const Side& S)
{
    return WaveSide{S.p, coeff_A(gamma, S.r), coeff_B(gamma, S.p), 2.0 * S.a, S.p * S.a};
}

// The pressure function f(p*) = f_L(p*) + f_R(p*) + (u_R - u_L)
// (exact.c:567-585) and its derivative, set up once per face for the
// iteration.  The sides are held with the lower pressure first: IEEE
// addition commutes, so f_L + f_R and f_R + f_L are the same number, and in
// this order every lane whose iterate lies between the two pressures takes
// the shock branch on s0 and the rarefaction branch on s1, whichever of L
// and R each is, so that a wavefront's lanes diverge less.
// lt, et: the power tables pow_pos reads (the global ones, or copies).
struct PressureFn {
    WaveSide s0, s1;
    double du;   // u_R - u_L
    double zf;   // (gamma - 1) / (2 gamma): f's rarefaction exponent
    double zd;   // -0.5 (gamma + 1) / gamma: f''s
    double gm1;  // gamma - 1
    const double* lt;
    const double* et;
};

HFV_HD __forceinline__ PressureFn pressure_fn_setup(double gamma, const Side& L, const Side& R,
                                                    const double* lt = pow_log_tab(),
                                                    const double* et = pow_exp_tab())
{
    const bool swap = R.p < L.p;
    PressureFn F;
    F.s0 = wave_side(gamma, swap ? R : L);
    F.s1 = wave_side(gamma, swap ? L : R);
    F.du = R.u - L.u;
    F.zf = (gamma - 1.0) / (2.0 * gamma);
    F.zd = -0.5 * (gamma + 1.0) / gamma;
    F.gm1 = gamma - 1.0;
    F.lt = lt;
    F.et = et;
    return F;
}

// wave_f of one side (exact.c:512-537), by branch: shock (p* > p) or
// rarefaction
HFV_HD __forceinline__ double shock_f(const WaveSide& S, double ps)
{
    return (ps - S.p) * sqrt(S.A / (ps + S.B));
}

HFV_HD __forceinline__ double rarefaction_f(const PressureFn& F, const WaveSide& S, double ps)
{
    return S.a2 * (pow_pos(ps / S.p, F.zf, F.lt, F.et) - 1.0) / F.gm1;
}

// wave_f and its derivative (exact.c:539-565) at one pressure, for the
// Newton step, by branch.  The derivative's shock branch keeps the (B - p*)
// of exact.c:553 as written.  The shock branch's square root serves both
// (ps + B and B + ps are the same sum), and so does the rarefaction
// branch's log(ps / p) (pow_log); each result has the bits of evaluating
// the reference's two functions separately.
HFV_HD __forceinline__ void shock_f_df(const WaveSide& S, double ps, double* f, double* df)
{
    const double s = sqrt(S.A / (ps + S.B));
    *f = (ps - S.p) * s;
    *df = (1.0 - 0.5 * (ps - S.p) / (S.B - ps)) * s;
}

HFV_HD __forceinline__ void rarefaction_f_df(const PressureFn& F, const WaveSide& S, double ps, double* f, double* df)
{
    const double x = ps / S.p;
    const DD l = pow_log(x, F.lt);
    *f = S.a2 * (pow_with_log(x, F.zf, l, F.et) - 1.0) / F.gm1;
    *df = pow_with_log(x, F.zd, l, F.et) / S.pa;
}

// f(p*) and f'(p*).  The branch each side takes is selected from values
// computed as follows: the lower-pressure side's shock value and the
// higher-pressure side's rarefaction value unconditionally (the pair a p*
// between the two pressures needs, in one stretch of code whose two chains
// overlap), and the other two under a branch, taken where some lane needs
// them.  The selected values are the reference's expressions, so the bits
// are the reference's.
HFV_HD __forceinline__ double pressure_fn(const PressureFn& F, double ps)
{
    double f0 = shock_f(F.s0, ps);
    double f1 = rarefaction_f(F, F.s1, ps);
    if (!(ps > F.s0.p)) f0 = rarefaction_f(F, F.s0, ps);
    if (ps > F.s1.p) f1 = shock_f(F.s1, ps);
    return f0 + f1 + F.du;
}

HFV_HD __forceinline__ void pressure_fn_df(const PressureFn& F, double ps, double* f, double* df)
{
    double f0, d0, f1, d1;
    shock_f_df(F.s0, ps, &f0, &d0);
    rarefaction_f_df(F, F.s1, ps, &f1, &d1);
    if (!(ps > F.s0.p)) rarefaction_f_df(F, F.s0, ps, &f0, &d0);
    if (ps > F.s1.p) shock_f_df(F.s1, ps, &f1, &d1);
    *f = f0 + f1 + F.du;
    *df = d0 + d1;
}

struct PStar {
    double p;
    int rc;
};

// The pressure iteration of exact.c:604-763 in resumable pieces: the
// Newton loop can stop after any iteration and continue later from the
// state it leaves, with the same operations on the same values as one
// uninterrupted loop.
struct NewtonState {
    double p0;         // current iterate
    double p_hi, f_hi; // the guess and f there (the loop's reference sign)
    double df0;        // f' at the guess (the first iteration's)
    double p_lo, f_lo; // the first iterate where f changed sign
    int it;            // next iteration
    int bracketed;
};

HFV_HD __forceinline__ NewtonState newton_begin(double gamma, Side L, Side R, double tol, const PressureFn& F)
{
    const double pg = guess_p(gamma, L.r, L.u, L.p, L.a, R.r, R.u, R.p, R.a, tol);
    NewtonState s;
    s.p0 = pg;
    s.p_hi = pg;
    pressure_fn_df(F, pg, &s.f_hi, &s.df0);
    s.p_lo = 0.0;
    s.f_lo = 0.0;
    s.it = 0;
    s.bracketed = 0;
    return s;
}

// Newton iterations s.it .. it_end - 1 (it_end <= 100).  Returns 1 with
// *ps when the iterate converged, -1 when the loop has ended without
// converging (a negative iterate, or 100 iterations), 0 when it stopped at
// it_end.
HFV_HD __forceinline__ int newton_run(const PressureFn& F, double tol, NewtonState& s, int it_end, double* ps)
{
    for (; s.it < it_end; s.it++) {
        double f, df;
        if (s.it == 0) {
            f = s.f_hi;
            df = s.df0;
        } else {
            pressure_fn_df(F, s.p0, &f, &df);
        }
        if (!s.bracketed && f * s.f_hi < 0.0) {
            s.bracketed = 1;
            s.f_lo = f;
            s.p_lo = s.p0;
        }
        const double p1 = s.p0 - f / df;
        if (p1 < 0.0) return -1;
        if (2.0 * fabs(p1 - s.p0) / (p1 + s.p0) < tol) {
            *ps = p1;
            return 1;
        }
        s.p0 = p1;
    }
    return (s.it >= 100) ? -1 : 0;
}

// After a Newton loop that did not converge: bracketing and bisection.
// The scan's f at the upper end of an interval is f at the lower end of the
// next (the reference evaluates it twice, exact.c:684-685; one evaluation
// gives the same bits).
// *bisections (when given): the bisection iterations taken.
HFV_HD __forceinline__ PStar newton_finish(const PressureFn& F, double tol, const NewtonState& s,
                                           int* bisections = nullptr)
{
    double p_lo = s.p_lo, f_lo = s.f_lo, p_hi = s.p_hi, f_hi = s.f_hi;
    if (!s.bracketed) {
        p_lo = 0.0;
        f_lo = pressure_fn(F, p_lo);
        if (f_lo * f_hi >= 0.0) {
            p_hi *= 10.0;
            const double dp = (p_hi - p_lo) / 100;
            double fb = pressure_fn(F, p_lo + 0 * dp);
            for (int k = 0; k < 100; k++) {
                const double a = p_lo + k * dp;
                const double b = p_lo + (k + 1) * dp;
                const double fa = fb;
                fb = pressure_fn(F, b);
                if (fa * fb < 0.0) {
                    p_lo = a;
                    f_lo = fa;
                    p_hi = b;
                    f_hi = fb;
                    break;
                }
            }
            if (f_lo * f_hi >= 0.0) return PStar{0.0, ERR_EXACT_BRACKET};
        }
    }
    int count = 0;
    for (;;) {
        const double pm = 0.5 * (p_lo + p_hi);
        if (p_hi - p_lo < tol) {
            if (bisections) *bisections = count;
            return PStar{pm, ERR_NONE};
        }
        const double fm = pressure_fn(F, pm);
        if (fm * f_lo < 0.0) {
            p_hi = pm;
            f_hi = fm;
        } else {
            p_lo = pm;
            f_lo = fm;
        }
        count++;
        if (count > 5000) {
            if (bisections) *bisections = count;
            return PStar{0.0, ERR_EXACT_BISECT};
        }
    }
}

HFV_HD __forceinline__ PStar solve_p_star(double gamma, Side L, Side R, double tol, int* bisections = nullptr)
{
    const PressureFn F = pressure_fn_setup(gamma, L, R);
    NewtonState s = newton_begin(gamma, L, R, tol, F);
    double ps;
    if (newton_run(F, tol, s, 100, &ps) > 0) return PStar{ps, ERR_NONE};
    return newton_finish(F, tol, s, bisections);
}

// exact.c:765-974 (left), 817-1080 (right; note its own operand orders).
// Returns (rho, u, p) in Prim's r, u, p.
HFV_HD __forceinline__ Prim sample_left(double gamma, Side W, double us, double ps, double speed)
{
    const double rho = W.r, u = W.u, p = W.p, a = W.a;
    Prim o{rho, u, 0.0, p};
    if (ps > p) {
        const double ratio = ps / p;
        const double s = u - a * sqrt((gamma + 1.0) / (2.0 * gamma) * ratio + (gamma - 1.0) / (2.0 * gamma));
        if (!(speed < s)) {
            const double gg = (gamma - 1.0) / (gamma + 1.0);
            o.r = (rho * (ratio + gg) / (gg * ratio + 1.0));
            o.u = us;
            o.p = ps;
        }
        return o;
    }
    if (speed < u - a) return o;
    const double a_star = a * pow_pos(ps / p, (gamma - 1.0) / (2.0 * gamma));
    if (speed < us - a_star) {
        const double c2 = 2.0 / (gamma + 1.0);
        const double common = pow_pos((c2 + (u - speed) * ((gamma - 1.0) / (a * (gamma + 1.0)))), 2.0 / (gamma - 1.0));
        o.r = rho * common;
        o.u = c2 * (a + 0.5 * u * (gamma - 1.0) + speed);
        o.p = p * pow_pos(common, gamma);
    } else {
        o.r = rho * pow_pos(ps / p, 1.0 / gamma);
        o.u = us;
        o.p = ps;
    }
    return o;
}

HFV_HD __forceinline__ Prim sample_right(double gamma, Side W, double us, double ps, double speed)
{
    const double rho = W.r, u = W.u, p = W.p, a = W.a;
    Prim o{rho, u, 0.0, p};
    if (ps > p) {
        const double ratio = ps / p;
        const double s = u + a * sqrt(ratio * (gamma + 1.0) / (2.0 * gamma) + (gamma - 1.0) / (2.0 * gamma));
        if (!(speed > s)) {
            const double gg = (gamma - 1.0) / (gamma + 1.0);
            o.r = (rho * (ratio + gg) / (gg * ratio + 1.0));
            o.u = us;
            o.p = ps;
        }
        return o;
    }
    if (speed > u + a) return o;
    const double a_star = a * pow_pos(ps / p, (gamma - 1.0) / (2.0 * gamma));
    if (speed > us + a_star) {
        const double c2 = 2.0 / (gamma + 1.0);
        const double common = pow_pos((c2 - (u - speed) * (gamma - 1.0) / (a * (gamma + 1.0))), 2.0 / (gamma - 1.0));
        o.r = rho * common;
        o.u = c2 * (-a + 0.5 * u * (gamma - 1.0) + speed);
        o.p = p * pow_pos(common, gamma);
    } else {
        o.r = rho * pow_pos(ps / p, 1.0 / gamma);
        o.u = us;
        o.p = ps;
    }
    return o;
}

// exact.c:332-469 at a given x/t.  w = (rho, u, v, p) of the sampled state:
// v is the transverse velocity carried by the contact (vl where x/t < u*,
// else vr) or by the vacuum sampler; us = u* (NaN on the vacuum branch).
// With vl = vr = 0 this is the 1D sampler.
struct Sample {
    Prim w;
    double us;
    int rc;
};

// exact.c:332-469 after the pressure iteration: u*, then the sample on the
// side of the contact that x/t falls on (or NaN and the failure's kind).
HFV_HD __forceinline__ Sample exact_sample_star(double gamma, Side L, double vl, Side R, double vr, PStar P,
                                                double speed)
{
    Sample S;
    S.us = __builtin_nan("");
    S.rc = ERR_NONE;
    if (P.rc != ERR_NONE) {
        const double nan = __builtin_nan("");
        S.w = Prim{nan, nan, nan, nan};
        S.rc = P.rc;
        return S;
    }
    const double ps = P.p;
    const double us = 0.5 * (L.u + R.u) + 0.5 * (wave_f(gamma, R.r, R.p, R.a, ps) - wave_f(gamma, L.r, L.p, L.a, ps));
    S.us = us;
    if (speed < us) {
        S.w = sample_left(gamma, L, us, ps, speed);
        S.w.v = vl;
    } else {
        S.w = sample_right(gamma, R, us, ps, speed);
        S.w.v = vr;
    }
    return S;
}

// SHORTCUT: a problem with equal left and right states (rho, u, p) whose
// solve runs clear of every special case (no vacuum, guess above tol, no
// overflow) takes the solve's result directly: the PVRS guess is p, the
// pressure function is +0 there, Newton returns p* = p at its first
// iterate, u* = (ul + ur)/2 + 0, and the sampler then returns the side's
// own state, or (rho, u*, p) on the star side of its rarefaction, whose fan
// has no width.  The bits equal the full solve's (checked on the host build
// against the full path, tests/test_adjoint_math.py).
// DIRECT_ONLY: return rc = ERR_PENDING for every problem the shortcut
// does not answer (the shortcut's answer is the same instructions as in the
// full solve).
// *bisections (when given): bisection iterations of the pressure solve (0
// when it took none).
template <bool SHORTCUT = true, bool DIRECT_ONLY = false>
HFV_HD __forceinline__ Sample exact_sample(double gamma, double tol, double rl, double ul, double vl, double pl,
                                           double rr, double ur, double vr, double pr, double speed,
                                           int* bisections = nullptr)
{
    Sample S;
    S.us = __builtin_nan("");
    S.rc = ERR_NONE;
    if constexpr (DIRECT_ONLY) {
        // only the shortcut can answer: test the states first; when they are
        // equal with rl > 0 and pl >= tol > 0 they have the same bits, so
        // the right sound speed is the left one
        if (!((rl == rr) & (ul == ur) & (pl == pr))) {
            S.rc = ERR_PENDING;
            return S;
        }
    }
    const double al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    const double ar = DIRECT_ONLY ? al : ((rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0);
    if constexpr (SHORTCUT) {
        const bool same = (rl == rr) & (ul == ur) & (pl == pr) & (rl > 0.0) & (pl >= tol) & (al > 0.0);
        const bool finite = (fabs(rl + rr) < INFINITY) & (fabs(pl + pr) < INFINITY) & (fabs(al + ar) < INFINITY) &
                            (fabs(pl * al) < INFINITY) & (fabs(ul + ur) < INFINITY);
        if (same && finite) {
            const double us = 0.5 * (ul + ur) + 0.5 * (0.0 - 0.0);
            S.us = us;
            if (speed < us) {
                S.w = Prim{rl, (speed < ul - al) ? ul : us, vl, pl};
            } else {
                S.w = Prim{rr, (speed > ur + ar) ? ur : us, vr, pr};
            }
            return S;
        }
    }
    if constexpr (DIRECT_ONLY) {
        S.rc = ERR_PENDING;
        return S;
    }
    if (rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul)) {
        S.w = vacuum(gamma, rl, ul, vl, pl, al, rr, ur, vr, pr, ar, speed);
        return S;
    }
    const Side L{rl, ul, pl, al}, R{rr, ur, pr, ar};
    return exact_sample_star(gamma, L, vl, R, vr, solve_p_star(gamma, L, R, tol, bisections), speed);
}

// The exact solve of a face in two parts, for a solver that sets aside the
// faces still iterating after newton_cap Newton iterations: exact_sample's
// answer, or rc = ERR_PAUSED with the iteration's state in *st; then
// exact_sample_resume finishes such a face.  Together they perform the
// operations of exact_sample (its shortcut included) on the same values.
constexpr int ERR_PAUSED = -2;

HFV_HD __forceinline__ Sample exact_sample_capped(double gamma, double tol, double rl, double ul, double vl,
                                                  double pl, double rr, double ur, double vr, double pr,
                                                  double speed, int newton_cap, NewtonState* st)
{
    const Sample D = exact_sample<true, true>(gamma, tol, rl, ul, vl, pl, rr, ur, vr, pr, speed);
    if (D.rc != ERR_PENDING) return D;
    const double al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0;
    if (rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul)) {
        Sample S;
        S.us = __builtin_nan("");
        S.rc = ERR_NONE;
        S.w = vacuum(gamma, rl, ul, vl, pl, al, rr, ur, vr, pr, ar, speed);
        return S;
    }
    const Side L{rl, ul, pl, al}, R{rr, ur, pr, ar};
    const PressureFn F = pressure_fn_setup(gamma, L, R);
    *st = newton_begin(gamma, L, R, tol, F);
    double ps;
    const int r = newton_run(F, tol, *st, newton_cap, &ps);
    if (r == 0) {
        Sample S;
        S.rc = ERR_PAUSED;
        return S;
    }
    return exact_sample_star(gamma, L, vl, R, vr, (r > 0) ? PStar{ps, ERR_NONE} : newton_finish(F, tol, *st),
                             speed);
}

// lt, et: the power tables the iteration reads (copies in LDS, say)
HFV_HD __forceinline__ Sample exact_sample_resume(double gamma, double tol, double rl, double ul, double vl,
                                                  double pl, double rr, double ur, double vr, double pr,
                                                  double speed, NewtonState st, const double* lt = pow_log_tab(),
                                                  const double* et = pow_exp_tab())
{
    const double al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0;
    const Side L{rl, ul, pl, al}, R{rr, ur, pr, ar};
    const PressureFn F = pressure_fn_setup(gamma, L, R, lt, et);
    double ps;
    const int r = newton_run(F, tol, st, 100, &ps);
    return exact_sample_star(gamma, L, vl, R, vr, (r > 0) ? PStar{ps, ERR_NONE} : newton_finish(F, tol, st), speed);
}

// The Newton loop of a set-aside face from its saved state st, without the
// finish that follows an unconverged loop: returns 1 with *P when the loop
// converged, or when it bracketed a root it had not bracketed when it was
// set aside (then *P is newton_finish's result); 0 when it ended without
// converging and with the bracket as it was, so that newton_finish of the
// saved state gives the result: newton_finish reads only the guess, f
// there and the bracket, which the loop changes only by bracketing a root.
HFV_HD __forceinline__ int newton_resume(const PressureFn& F, double tol, NewtonState st, PStar* P)
{
    const int br0 = st.bracketed;
    double ps;
    const int r = newton_run(F, tol, st, 100, &ps);
    if (r > 0) {
        *P = PStar{ps, ERR_NONE};
        return 1;
    }
    if (st.bracketed == br0) return 0;
    *P = newton_finish(F, tol, st);
    return 1;
}

// The flux of the state sampled at x/t = 0 (exact.c:471-510; NF = 4: the
// transverse velocity's terms added).  Returns the sample's failure kind.
template <int NF>
HFV_HD __forceinline__ int flux_from_sample(double gamma, const Sample& S, double* f)
{
    const double r = S.w.r, u = S.w.u, p = S.w.p;
    f[0] = r * u;
    f[1] = r * u * u + p;
    if constexpr (NF == 4) {
        const double v = S.w.v;
        f[2] = f[0] * v;
        f[3] = (p * (gamma / (gamma - 1.0)) + (0.5 * r * u * u + 0.5 * r * v * v)) * u;
    } else {
        f[2] = (p * (gamma / (gamma - 1.0)) + 0.5 * r * u * u) * u;
    }
    return S.rc;
}

// Exact Godunov flux at x/t = 0.  NF = 3: exact.c:471-510.  NF = 4: the
// normal components are computed exactly as in 1D; the transverse velocity is
// advected by the contact (v_L when 0 < u*, the "speed < u_star" side of
// exact.c:431, else v_R); on the vacuum branch the vacuum sampler carries it.
// At v = 0 the 2D flux equals the 1D flux bitwise.
// DIRECT_ONLY: ERR_PENDING (f untouched) for a face whose states differ.
template <int NF, bool DIRECT_ONLY = false>
HFV_HD __forceinline__ int flux_exact(double gamma, double tol, const double* wl, const double* wr, double* f)
{
    const double vl = (NF == 4) ? wl[2] : 0.0;
    const double vr = (NF == 4) ? wr[2] : 0.0;
    const Sample S = exact_sample<true, DIRECT_ONLY>(gamma, tol, wl[0], wl[1], vl, wl[NF - 1], wr[0], wr[1], vr,
                                                     wr[NF - 1], 0.0);
    if (DIRECT_ONLY && S.rc == ERR_PENDING) return ERR_PENDING;
    return flux_from_sample<NF>(gamma, S, f);
}

template <int SOLVER, int NF>
HFV_HD __forceinline__ int flux(double gamma, double tol, const double* wl, const double* wr, double* f)
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
// cr: the value of c / s[0], which the predictor takes as the caller has
// it (the strip step's grouped divisions, divide_by_rho)
HFV_HD __forceinline__ void predict_x_cr(const double* s, const double* d, double c, double cr, double a, double* lo,
                                         double* hi)
{
    const double up = 1.0 + c * s[1];
    const double um = 1.0 - c * s[1];
    lo[0] = s[0] + -0.5 * (up * d[0] + c * s[0] * d[1] + 0.0 + 0.0);
    lo[1] = s[1] + -0.5 * (0.0 + up * d[1] + 0.0 + cr * d[3]);
    lo[2] = s[2] + -0.5 * (0.0 + 0.0 + up * d[2] + 0.0);
    lo[3] = s[3] + -0.5 * (0.0 + c * s[0] * a * a * d[1] + 0.0 + up * d[3]);
    hi[0] = s[0] + 0.5 * (um * d[0] - c * s[0] * d[1] + 0.0 + 0.0);
    hi[1] = s[1] + 0.5 * (0.0 + um * d[1] + 0.0 - cr * d[3]);
    hi[2] = s[2] + 0.5 * (0.0 + 0.0 + um * d[2] + 0.0);
    hi[3] = s[3] + 0.5 * (0.0 - c * s[0] * a * a * d[1] + 0.0 + um * d[3]);
}

HFV_HD __forceinline__ void predict_x(const double* s, const double* d, double c, double a, double* lo, double* hi)
{
    predict_x_cr(s, d, c, c / s[0], a, lo, hi);
}

HFV_HD __forceinline__ void predict_y_cr(const double* s, const double* d, double c, double cr, double a, double* lo,
                                         double* hi)
{
    const double vp = 1.0 + c * s[2];
    const double vm = 1.0 - c * s[2];
    lo[0] = s[0] + -0.5 * (vp * d[0] + c * s[0] * d[1] + 0.0 + 0.0);
    lo[1] = s[1] + -0.5 * (0.0 + vp * d[1] + 0.0 + 0.0);
    lo[2] = s[2] + -0.5 * (0.0 + 0.0 + vp * d[2] + cr * d[3]);
    lo[3] = s[3] + -0.5 * (0.0 + 0.0 + c * s[0] * a * a * d[2] + vp * d[3]);
    hi[0] = s[0] + 0.5 * (vm * d[0] - c * s[0] * d[1] + 0.0 + 0.0);
    hi[1] = s[1] + 0.5 * (0.0 + vm * d[1] + 0.0 + 0.0);
    hi[2] = s[2] + 0.5 * (0.0 + 0.0 + vm * d[2] - cr * d[3]);
    hi[3] = s[3] + 0.5 * (0.0 + 0.0 - c * s[0] * a * a * d[2] + vm * d[3]);
}

HFV_HD __forceinline__ void predict_y(const double* s, const double* d, double c, double a, double* lo, double* hi)
{
    predict_y_cr(s, d, c, c / s[0], a, lo, hi);
}

HFV_HD __forceinline__ void predict_1d(const double* s, const double* d, double c, double a, double* lo, double* hi)
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
HFV_HD __forceinline__ void c2p_2d(double gamma, double vol, double m, double mx, double my, double e, double* w)
{
    w[0] = m / vol;
    w[1] = mx / m;
    w[2] = my / m;
    w[3] = (gamma - 1.0) * (e - 0.5 * m * (w[1] * w[1] + w[2] * w[2])) / vol;
}

// c2p_2d with the divisions by m sharing one reciprocal and those by the
// (uniform) cell volume sharing rvol = rcp_core(vol), vol_ok =
// div_core_ok(vol), when every lane's operands allow; the bits of c2p_2d.
HFV_HD __forceinline__ void c2p_2d_rn(double gamma, double vol, double rvol, bool vol_ok, double m, double mx,
                                      double my, double e, double* w)
{
    if (wave_all(vol_ok & div_core_ok(m) & div_core_ok(mx) & div_core_ok(my))) {
        const double rm = rcp_core(m);
        w[0] = div_core(m, vol, rvol);
        w[1] = div_core(mx, m, rm);
        w[2] = div_core(my, m, rm);
    } else {
        w[0] = m / vol;
        w[1] = mx / m;
        w[2] = my / m;
    }
    const double num = (gamma - 1.0) * (e - 0.5 * m * (w[1] * w[1] + w[2] * w[2]));
    if (wave_all(vol_ok & div_core_ok(num))) w[3] = div_core(num, vol, rvol);
    else w[3] = num / vol;
}

HFV_HD __forceinline__ void c2p_1d(double gamma, double vol, double m, double mx, double e, double* w)
{
    w[0] = m / vol;
    w[1] = mx / m;
    w[2] = (gamma - 1.0) * (e - 0.5 * m * w[1] * w[1]) / vol;
}

// ---------------------------------------------------------------------------
// Source terms (source_term.c).  Rates of the geometric source of one cell
// (:28-43): pre = -alpha V / r_mid, rho the density, sm and se the specific
// momentum and energy; o = (mass, momentum, energy) rates.
HFV_HD __forceinline__ void geo_rate(double pre, double gamma, double rho, double sm, double se, double* o)
{
    const double pr = (gamma - 1) * (se - 0.5 * sm * sm / rho);
    o[0] = pre * sm;
    o[1] = pre * sm * sm / rho;
    o[2] = pre * sm * (se + pr) / rho;
}

// One cell's classical RK4 step of the geometric source (:95-175; alpha 1
// cylindrical, 2 spherical): u = (m, mx, E) updated in place, rho the
// cell's density before the step.
HFV_HD __forceinline__ void geometry_source_rk4(double alpha, double gamma, double vol, double mid, double rho,
                                                double dt, double* u)
{
    const double m = u[0], mx = u[1], e = u[2];
    const double pre = -alpha * vol / mid;
    double k1[3], k2[3], k3[3], k4[3];
    geo_rate(pre, gamma, rho, mx / vol, e / vol, k1);
    double tm = m + 0.5 * dt * k1[0], tu = mx + 0.5 * dt * k1[1], te = e + 0.5 * dt * k1[2];
    geo_rate(pre, gamma, tm / vol, tu / vol, te / vol, k2);
    tm = m + 0.5 * dt * k2[0];
    tu = mx + 0.5 * dt * k2[1];
    te = e + 0.5 * dt * k2[2];
    geo_rate(pre, gamma, tm / vol, tu / vol, te / vol, k3);
    tm = m + dt * k3[0];
    tu = mx + dt * k3[1];
    te = e + dt * k3[2];
    geo_rate(pre, gamma, tm / vol, tu / vol, te / vol, k4);
    u[0] += dt * (k1[0] + 2.0 * k2[0] + 2.0 * k3[0] + k4[0]) / 6.0;
    u[1] += dt * (k1[1] + 2.0 * k2[1] + 2.0 * k3[1] + k4[1]) / 6.0;
    u[2] += dt * (k1[2] + 2.0 * k2[2] + 2.0 * k3[2] + k4[2]) / 6.0;
}

// Gravity on one cell (:213-240): the y momentum gains -dt g m, the energy
// its RK4 integral; m is unchanged.
HFV_HD __forceinline__ void gravity_source(double dt, double gravity, double m, double* my, double* e)
{
    const double pre = -dt * gravity;
    const double kmy = pre * m;
    const double k1 = pre * *my;
    const double k2 = pre * (*my + 0.5 * kmy);
    const double k3 = k2;
    const double k4 = pre * (*my + kmy);
    *my += kmy;
    *e += (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0;
}

// CFL reduction key: the maximum of non-negative doubles as an unsigned
// maximum of (bit pattern + 1); key 0 stands for "no finite candidate",
// i.e. the reference's initial -INFINITY (utils.c:75-92).
HFV_HD __forceinline__ unsigned long long smax_key(double v)
{
    union { double d; unsigned long long u; } c;
    c.d = v;
    return (v >= 0.0) ? c.u + 1ull : 0ull;
}

__host__ HFV_HD __forceinline__ double smax_from_key(unsigned long long k)
{
    if (k == 0ull) return -INFINITY;
    union { unsigned long long u; double d; } c;
    c.u = k - 1ull;
    return c.d;
}

}  // namespace hfv
