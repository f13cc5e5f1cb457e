// hfv_adjoint.hpp -- reverse-mode derivatives (vector-Jacobian products) of
// the pieces of the MUSCL-Hancock step, for the discrete adjoint.
//
// Each *_vjp function recomputes the forward quantities it needs and returns
// the cotangents of its inputs given the cotangent of its outputs.  The
// derivatives are those of the real-valued functions that the forward code
// evaluates (hfv_math.hpp): one-sided where the forward branches (limiter
// kinks, HLLC wave regions, pressure-guess choice, q = 1 when p* <= p_K, the
// tol clamp of the guess), with the branch taken by the forward evaluation.
// The vacuum branch of HLLC has no derivative here: its occurrence is
// reported and its contribution is zero.
// The forward quantities that choose a branch are evaluated with the
// forward's own expressions, so the branch is the forward's.  Quantities
// that enter only the derivatives divide by a denominator through its
// reciprocal, computed once: the derivatives are then within a few ulps of
// the exact quotients' (cheaper than one correctly rounded division each).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "hfv_math.hpp"

namespace hfv {

enum AdjFlag : int { ADJ_OK = 0, ADJ_VACUUM = 1, ADJ_INVALID = 2 };

// ---------------------------------------------------------------------------
// HLLC (2D, sweep-local state (rho, u_n, u_t, p)).  mf = cotangent of the
// flux (mass, normal momentum, transverse momentum, energy); ml, mr receive
// the cotangents of the left and right states (overwritten).
HFV_HD __forceinline__ int hllc_vjp(double gamma, double tol, const double* wl, const double* wr, const double* mf,
                                     double* ml, double* mr)
{
    for (int k = 0; k < 4; k++) ml[k] = mr[k] = 0.0;
    const double rl = wl[0], ul = wl[1], vl = wl[2], pl = wl[3];
    const double rr = wr[0], ur = wr[1], vr = wr[2], pr = wr[3];
    const double al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0;
    if (rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul)) return ADJ_VACUUM;

    // ---- forward: pressure guess (riemann_solver.c:154-224) -------------
    const double plo = (pl > pr) ? pr : pl;
    const double phi = (pl > pr) ? pl : pr;
    const double ppv = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (al + ar);
    int gb;  // 0 PVRS, 1 two-rarefaction, 2 two-shock
    double ps, h = 0.0, z = 0.0, X = 0.0, tnum = 0.0, tden = 0.0, pzl = 0.0, pzr = 0.0;
    double gl = 0.0, gr = 0.0, Al = 0.0, Ar = 0.0, Hl = 0.0, Hr = 0.0;
    if (phi / plo <= 2.0 && (plo <= ppv && ppv <= phi)) {
        gb = 0;
        ps = ppv;
    } else if (ppv < plo) {
        gb = 1;
        h = 0.5 * (gamma - 1.0);
        z = h / gamma;
        tnum = al + ar - h * (ur - ul);
        HllcFace F;
        F.wl[3] = pl;
        F.wr[3] = pr;
        F.al = al;
        F.ar = ar;
        F.ps = tnum;
        if (trrs_settle_direct(gamma, tol, F)) {
            // p* lies below both pressures (as the forward establishes it):
            // q_L = q_R = 1, so the guess gets no cotangent and its powers
            // are not needed; finite stand-ins keep the zero terms zero
            ps = F.ps;
            pzl = pzr = tden = X = 1.0;
        } else {
            pzl = pow_pos(pl, z);
            pzr = pow_pos(pr, z);
            tden = al / pzl + ar / pzr;
            X = tnum / tden;
            ps = pow_pos(X, gamma / h);
        }
    } else {
        gb = 2;
        Al = coeff_A(gamma, rl);
        Ar = coeff_A(gamma, rr);
        Hl = ppv + coeff_B(gamma, pl);
        Hr = ppv + coeff_B(gamma, pr);
        gl = sqrt(Al / Hl);
        gr = sqrt(Ar / Hr);
        ps = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
    }
    const bool clamped = ps < tol;
    if (clamped) ps = tol;
    const double kq = 0.5 + 0.5 / gamma;
    const bool shl = ps > pl, shr = ps > pr;
    const double ql = shl ? sqrt(1.0 + kq * (ps / pl - 1.0)) : 1.0;
    const double qr = shr ? sqrt(1.0 + kq * (ps / pr - 1.0)) : 1.0;
    const double sl = ul - al * ql;
    const double sr = ur + ar * qr;
    const double num = pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur);
    const double den = rl * (sl - ul) - rr * (sr - ur);
    const double sm = num / den;
    int region;
    if (0.0 <= sl) region = 0;
    else if (sl <= 0.0 && 0.0 <= sm) region = 1;
    else if (sm <= 0.0 && 0.0 <= sr) region = 2;
    else if (sr <= 0.0) region = 3;
    else return ADJ_INVALID;

    const bool left = region <= 1;
    const double r = left ? rl : rr, u = left ? ul : ur, v = left ? vl : vr, p = left ? pl : pr;
    const double e = 0.5 * r * (u * u + v * v) + p / (gamma - 1.0);
    const double fm = r * u;
    double m_r = 0.0, m_u = 0.0, m_v = 0.0, m_p = 0.0, m_s = 0.0, m_sm = 0.0, m_fm = 0.0, m_e = 0.0;
    const double m0 = mf[0], m1 = mf[1], m2 = mf[2], m3 = mf[3];
    if (region == 0 || region == 3) {
        // F = (fm, fm u + p, fm v, u (e + p))
        m_fm = m0 + m1 * u + m2 * v;
        m_u = m1 * fm + m3 * (e + p);
        m_v = m2 * fm;
        m_p = m1 + m3 * u;
        m_e = m3 * u;
    } else {
        const double s = left ? sl : sr;
        const double w = s - u;
        const double D = s - sm;
        const double ir = 1.0 / r, iD = 1.0 / D, irw = 1.0 / (r * w);
        const double iw = r * irw;
        const double rs = r * w * iD;
        const double T = sm + p * irw;
        const double Q = e * ir + (sm - u) * T;
        const double es = rs * Q;
        double m_rs = 0.0, m_es = 0.0, m_w = 0.0, m_D = 0.0;
        // F0 = fm + s (rs - r)
        m_fm += m0;
        m_s += m0 * (rs - r);
        m_rs += m0 * s;
        m_r -= m0 * s;
        // F1 = fm u + p + s (rs sm - r u)
        m_fm += m1 * u;
        m_u += m1 * fm;
        m_p += m1;
        m_s += m1 * (rs * sm - r * u);
        m_rs += m1 * s * sm;
        m_sm += m1 * s * rs;
        m_r -= m1 * s * u;
        m_u -= m1 * s * r;
        // F2 = fm v + s (rs v - r v)
        m_fm += m2 * v;
        m_v += m2 * (fm + s * (rs - r));
        m_s += m2 * (rs * v - r * v);
        m_rs += m2 * s * v;
        m_r -= m2 * s * v;
        // F3 = u (e + p) + s (es - e)
        m_u += m3 * (e + p);
        m_e += m3 * (u - s);
        m_p += m3 * u;
        m_s += m3 * (es - e);
        m_es += m3 * s;
        // es = rs Q, Q = e/r + (sm - u) T, T = sm + p/(r w)
        m_rs += m_es * Q;
        const double m_Q = m_es * rs;
        m_e += m_Q * ir;
        m_r -= m_Q * e * ir * ir;
        m_sm += m_Q * T;
        m_u -= m_Q * T;
        const double m_T = m_Q * (sm - u);
        m_sm += m_T;
        m_p += m_T * irw;
        m_r -= m_T * p * irw * ir;
        m_w -= m_T * p * irw * iw;
        // rs = r w / D
        m_r += m_rs * w * iD;
        m_w += m_rs * r * iD;
        m_D -= m_rs * rs * iD;
        // w = s - u, D = s - sm
        m_s += m_w + m_D;
        m_u -= m_w;
        m_sm -= m_D;
    }
    // fm = r u, e = r (u^2 + v^2)/2 + p/(gamma-1)
    m_r += m_fm * u + m_e * 0.5 * (u * u + v * v);
    m_u += m_fm * r + m_e * r * u;
    m_v += m_e * r * v;
    m_p += m_e * (1.0 / (gamma - 1.0));
    double* mk = left ? ml : mr;
    mk[0] += m_r;
    mk[1] += m_u;
    mk[2] += m_v;
    mk[3] += m_p;
    if (region == 0 || region == 3) return ADJ_OK;

    // ---- s (= sl or sr) and sm back to the inputs ------------------------
    double m_sl = left ? m_s : 0.0, m_sr = left ? 0.0 : m_s;
    const double iden = 1.0 / den;
    const double m_num = m_sm * iden;
    const double m_den = -m_sm * sm * iden;
    ml[3] -= m_num;
    mr[3] += m_num;
    ml[0] += m_num * ul * (sl - ul) + m_den * (sl - ul);
    ml[1] += m_num * (rl * (sl - ul) - rl * ul) - m_den * rl;
    m_sl += m_num * rl * ul + m_den * rl;
    mr[0] -= m_num * ur * (sr - ur) + m_den * (sr - ur);
    mr[1] += m_num * (rr * ur - rr * (sr - ur)) + m_den * rr;
    m_sr -= m_num * rr * ur + m_den * rr;
    // sl = ul - al ql, sr = ur + ar qr
    ml[1] += m_sl;
    mr[1] += m_sr;
    double m_al = -m_sl * ql, m_ar = m_sr * qr;
    const double m_ql = -m_sl * al, m_qr = m_sr * ar;
    double m_ps = 0.0;
    const double ipl = 1.0 / pl, ipr = 1.0 / pr;
    if (shl) {
        const double t = m_ql * kq / (2.0 * ql) * ipl;
        m_ps += t;
        ml[3] -= t * ps * ipl;
    }
    if (shr) {
        const double t = m_qr * kq / (2.0 * qr) * ipr;
        m_ps += t;
        mr[3] -= t * ps * ipr;
    }
    // ---- pressure guess back to the inputs ------------------------------
    double m_ppv = 0.0;
    if (!clamped) {
        if (gb == 0) {
            m_ppv = m_ps;
        } else if (gb == 1) {
            const double m_X = m_ps * ps / (z * X);
            const double itd = 1.0 / tden, ipzl = 1.0 / pzl, ipzr = 1.0 / pzr;
            const double m_tn = m_X * itd;
            const double m_td = -m_X * X * itd;
            m_al += m_tn + m_td * ipzl;
            m_ar += m_tn + m_td * ipzr;
            mr[1] -= m_tn * h;
            ml[1] += m_tn * h;
            ml[3] -= m_td * al * z * ipzl * ipl;
            mr[3] -= m_td * ar * z * ipzr * ipr;
        } else {
            const double iS = 1.0 / (gl + gr);
            const double m_gl = m_ps * (pl - ps) * iS;
            const double m_gr = m_ps * (pr - ps) * iS;
            ml[3] += m_ps * gl * iS;
            mr[3] += m_ps * gr * iS;
            mr[1] -= m_ps * iS;
            ml[1] += m_ps * iS;
            const double iHl = 1.0 / Hl, iHr = 1.0 / Hr;
            const double m_Al = m_gl / (2.0 * gl) * iHl, m_Hl = -m_gl * gl * 0.5 * iHl;
            const double m_Ar = m_gr / (2.0 * gr) * iHr, m_Hr = -m_gr * gr * 0.5 * iHr;
            ml[0] -= m_Al * Al / rl;
            mr[0] -= m_Ar * Ar / rr;
            const double gg = (gamma - 1.0) / (gamma + 1.0);
            ml[3] += m_Hl * gg;
            mr[3] += m_Hr * gg;
            m_ppv = m_Hl + m_Hr;
        }
    }
    if (m_ppv != 0.0) {
        // ppv = (pl + pr)/2 - (ur - ul)(rl + rr)(al + ar)/8
        ml[3] += 0.5 * m_ppv;
        mr[3] += 0.5 * m_ppv;
        const double A8 = 0.125 * (rl + rr) * (al + ar);
        mr[1] -= m_ppv * A8;
        ml[1] += m_ppv * A8;
        const double B8 = 0.125 * (ur - ul) * (al + ar);
        ml[0] -= m_ppv * B8;
        mr[0] -= m_ppv * B8;
        const double C8 = 0.125 * (ur - ul) * (rl + rr);
        m_al -= m_ppv * C8;
        m_ar -= m_ppv * C8;
    }
    // a = sqrt(gamma p / rho)
    const double hal = 0.5 * m_al * al, har = 0.5 * m_ar * ar;
    ml[3] += hal * ipl;
    ml[0] -= hal / rl;
    mr[3] += har * ipr;
    mr[0] -= har / rr;
    return ADJ_OK;
}

// ---------------------------------------------------------------------------
// Predictor in sweep-local order s = (rho, n, t, p), slopes d likewise;
// quirk = the density row uses the transverse slope (the y-sweep form).
// Given cotangents mlo, mhi of the face states, add the cotangents of s
// (ms) and of the slopes (md).  rho a^2 = gamma p is used for the pressure row.
HFV_HD __forceinline__ void predict_vjp(const double* s, const double* d, double c, double gamma, bool quirk,
                                            const double* mlo, const double* mhi, double* ms, double* md)
{
    const double rho = s[0], n = s[1], p = s[3];
    const double irho = 1.0 / rho;
    const double np_ = 1.0 + c * n, nm = 1.0 - c * n;
    const double dq = quirk ? d[2] : d[1];
    const double m0 = mlo[0] + mhi[0], mn = mlo[1] + mhi[1], mt = mlo[2] + mhi[2], m3 = mlo[3] + mhi[3];
    ms[0] += m0 * (1.0 - 0.5 * c * dq) + mn * 0.5 * c * d[3] * irho * irho;
    ms[1] += -m0 * 0.5 * c * d[0] + mn * (1.0 - 0.5 * c * d[1]) - mt * 0.5 * c * d[2] - m3 * 0.5 * c * d[3];
    ms[2] += mt;
    ms[3] += m3 * (1.0 - 0.5 * c * gamma * d[1]);
    md[0] = -0.5 * np_ * mlo[0] + 0.5 * nm * mhi[0];
    md[1] = -0.5 * np_ * mlo[1] + 0.5 * nm * mhi[1] - 0.5 * c * gamma * p * m3;
    md[2] = -0.5 * np_ * mlo[2] + 0.5 * nm * mhi[2];
    md[3] = -0.5 * (c * irho) * mn - 0.5 * np_ * mlo[3] + 0.5 * nm * mhi[3];
    if (quirk) md[2] += -0.5 * c * rho * m0;
    else md[1] += -0.5 * c * rho * m0;
}

// Partial derivatives (da, db) of limit<LIM>(a, b) (slope_limiter.c:49-91),
// one-sided at the kinks with the branch the forward evaluation takes.
template <int LIM>
HFV_HD __forceinline__ void limit_partials(double a, double b, double* da, double* db)
{
    *da = 0.0;
    *db = 0.0;
    if (a * b <= 0.0) return;
    if constexpr (LIM == MINMOD) {
        if (fabs(a) < fabs(b)) *da = 1.0;
        else *db = 1.0;
    } else if constexpr (LIM == VAN_LEER) {
        // 2ab / (a + b + 1e-10)
        const double s = a + b + 1e-10;
        *da = 2.0 * b * (b + 1e-10) / (s * s);
        *db = 2.0 * a * (a + 1e-10) / (s * s);
    } else {
        // MC: the central difference, or twice the smaller one-sided slope
        const double cen = (a > 0) ? 0.5 * (a + b) : -0.5 * (a + b);
        const double m2 = 2.0 * fmin(fabs(a), fabs(b));
        if (cen <= m2) {
            *da = 0.5;
            *db = 0.5;
        } else if (fabs(a) <= fabs(b)) {
            *da = 2.0;
        } else {
            *db = 2.0;
        }
    }
}

// c2p (system.c:763-802) transposed: cotangent of (rho, u, v, p) -> of (m, mx, my, E)
HFV_HD __forceinline__ void c2p_vjp(double gamma, double vol, double m, const double* w, const double* mw,
                                        double* mu)
{
    const double u = w[1], v = w[2];
    const double ivol = 1.0 / vol, im = 1.0 / m;
    const double gv = (gamma - 1.0) * ivol;
    mu[0] = mw[0] * ivol - (mw[1] * u + mw[2] * v) * im + mw[3] * gv * 0.5 * (u * u + v * v);
    mu[1] = mw[1] * im - mw[3] * gv * u;
    mu[2] = mw[2] * im - mw[3] * gv * v;
    mu[3] = mw[3] * gv;
}

}  // namespace hfv
