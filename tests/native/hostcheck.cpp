// hostcheck.cpp -- the device math of hydrofv_amd (hfv_math.hpp,
// hfv_adjoint.hpp) compiled for the host, with batched C entry points, so
// that CPU tests can check the kernels' own source: forward pieces against
// the oracle, derivative pieces against automatic differentiation.
// Test infrastructure only; built by tests/native/Makefile.
#include <cstring>
#include "../../hydrofv_amd/csrc/device/hfv_adjoint.hpp"
#include "../../hydrofv_amd/csrc/device/hfv_pow.hpp"
#include "../../hydrofv_amd/csrc/device/hfv_strips.hpp"
#include "hydrofv/hfv_gpu.h"

#include <cstddef>

using namespace hfv;

namespace {

template <int LIM>
void limits(long long n, const double* a, const double* b, double* d, double* da, double* db)
{
    for (long long i = 0; i < n; i++) {
        d[i] = limit<LIM>(a[i], b[i]);
        limit_partials<LIM>(a[i], b[i], &da[i], &db[i]);
    }
}

}  // namespace

extern "C" {

// the exact solver's sampler with (shortcut = 1) or without its equal-state
// shortcut: w[n][4] = (rho, u, v, p) left and right states, speed[n]
void hc_exact_sample(int shortcut, double gamma, double tol, long long n, const double* wl, const double* wr,
                     const double* speed, double* out, double* us, int* rc)
{
    for (long long i = 0; i < n; i++) {
        const double* l = wl + 4 * i;
        const double* r = wr + 4 * i;
        const Sample S = shortcut ? exact_sample<true>(gamma, tol, l[0], l[1], l[2], l[3], r[0], r[1], r[2], r[3], speed[i])
                                  : exact_sample<false>(gamma, tol, l[0], l[1], l[2], l[3], r[0], r[1], r[2], r[3], speed[i]);
        out[4 * i + 0] = S.w.r;
        out[4 * i + 1] = S.w.u;
        out[4 * i + 2] = S.w.v;
        out[4 * i + 3] = S.w.p;
        us[i] = S.us;
        rc[i] = S.rc;
    }
}

// exact_sample_capped, then exact_sample_resume for the faces it set
// aside: the device's two-pass face solve, face by face
void hc_exact_sample_capped(int cap, double gamma, double tol, long long n, const double* wl, const double* wr,
                            const double* speed, double* out, double* us, int* rc, int* paused)
{
    for (long long i = 0; i < n; i++) {
        const double* l = wl + 4 * i;
        const double* r = wr + 4 * i;
        NewtonState st;
        Sample S = exact_sample_capped(gamma, tol, l[0], l[1], l[2], l[3], r[0], r[1], r[2], r[3], speed[i], cap, &st);
        paused[i] = S.rc == ERR_PAUSED;
        if (paused[i]) S = exact_sample_resume(gamma, tol, l[0], l[1], l[2], l[3], r[0], r[1], r[2], r[3], speed[i], st);
        out[4 * i + 0] = S.w.r;
        out[4 * i + 1] = S.w.u;
        out[4 * i + 2] = S.w.v;
        out[4 * i + 3] = S.w.p;
        us[i] = S.us;
        rc[i] = S.rc;
    }
}

// The set-aside faces' split solve, face by face: Newton to cap iterations;
// for a face still iterating, newton_resume from that state, and where it
// leaves the result to the finish, newton_finish of the set-aside state (the
// 2D step's k_solve_slow, k_solve_spec and k_solve_join).  same[n] = 1 when
// p* and the failure kind equal solve_p_star's bit for bit; took_spec[n] = 1
// when the finish of the set-aside state gave the result.
void hc_resume_split(int cap, double gamma, double tol, long long n, const double* wl, const double* wr, int* same,
                     int* took_spec)
{
    for (long long i = 0; i < n; i++) {
        const double* l = wl + 4 * i;
        const double* r = wr + 4 * i;
        same[i] = 1;
        took_spec[i] = 0;
        if (!(l[0] > 0.0) || !(r[0] > 0.0)) continue;
        const double al = sound_speed(gamma, l[0], l[3]), ar = sound_speed(gamma, r[0], r[3]);
        if (((al + ar) * 2.0 / (gamma + 1.0)) <= (r[1] - l[1])) continue;
        const Side L{l[0], l[1], l[3], al}, R{r[0], r[1], r[3], ar};
        const PStar full = solve_p_star(gamma, L, R, tol);
        const PressureFn F = pressure_fn_setup(gamma, L, R);
        NewtonState st = newton_begin(gamma, L, R, tol, F);
        double ps;
        const int rc = newton_run(F, tol, st, cap, &ps);
        PStar P;
        if (rc > 0) P = PStar{ps, ERR_NONE};
        else if (rc < 0) P = newton_finish(F, tol, st);
        else if (!newton_resume(F, tol, st, &P)) {
            P = newton_finish(F, tol, st);
            took_spec[i] = 1;
        }
        const bool eq = (P.rc == full.rc) && (P.rc != ERR_NONE || memcmp(&P.p, &full.p, sizeof(double)) == 0);
        same[i] = eq ? 1 : 0;
    }
}

// sizes and selected field offsets of the structs of hfv_gpu.h, for the
// ctypes mirrors: out[0..3] sizeof(mh2d, mh1d, run_status, adj2d config),
// out[4..] offsets of the fields added last to each config
void hc_struct_layout(long long* out)
{
    out[0] = sizeof(hfv_mh2d_config);
    out[1] = sizeof(hfv_mh1d_config);
    out[2] = sizeof(hfv_run_status);
    out[3] = sizeof(hfv_adj2d_config);
    out[4] = offsetof(hfv_mh2d_config, full_tiles);
    out[5] = offsetof(hfv_mh1d_config, t_pause);
    out[6] = offsetof(hfv_mh1d_config, random_choice);
    out[7] = offsetof(hfv_adj2d_config, slope_limiter);
}

// out[n] = pow_pos(x[n], y[n]) (the kernels' pow of the two-rarefaction guess)
void hc_pow_pos(long long n, const double* x, const double* y, double* out)
{
    for (long long i = 0; i < n; i++) out[i] = pow_pos(x[i], y[i]);
}

// f[n][nf] = HLLC flux (flux_hllc<nf>); err[n]
void hc_flux_hllc(int nf, double gamma, double tol, long long n, const double* wl, const double* wr, double* f,
                  int* err)
{
    for (long long i = 0; i < n; i++) {
        if (nf == 3) err[i] = flux_hllc<3>(gamma, tol, wl + 3 * i, wr + 3 * i, f + 3 * i);
        else err[i] = flux_hllc<4>(gamma, tol, wl + 4 * i, wr + 4 * i, f + 4 * i);
    }
}

// settled[n] = 1 where trrs_settle_direct settles the face's two-rarefaction
// guess without powers, 0 where it does not, -1 for faces of another kind
void hc_trrs_settle_direct(double gamma, double tol, long long n, const double* wl, const double* wr, int* settled)
{
    for (long long i = 0; i < n; i++) {
        HllcFace F;
        hllc_begin(gamma, tol, wl + 4 * i, wr + 4 * i, F);
        settled[i] = (F.kind == HLLC_TRRS) ? (trrs_settle_direct(gamma, tol, F) ? 1 : 0) : -1;
    }
}

// The 2D kernels' split form: hllc_begin, the two-rarefaction powers taken
// as the wavefront takes them (trrs_settle_direct when settle_direct, then
// the powers with trrs_settle), hllc_end.
void hc_flux_hllc_split(double gamma, double tol, long long n, const double* wl, const double* wr, double* f,
                        int* err, int settle_direct)
{
    for (long long i = 0; i < n; i++) {
        HllcFace F;
        hllc_begin(gamma, tol, wl + 4 * i, wr + 4 * i, F);
        if (F.kind == HLLC_TRRS && !(settle_direct && trrs_settle_direct(gamma, tol, F))) {
            const double h = 0.5 * (gamma - 1.0);
            const double pl = pow_pos(F.wl[3], h / gamma), pr = pow_pos(F.wr[3], h / gamma);
            double x;
            if (!trrs_settle(tol, F, pl, pr, &x)) {
                F.ps = pow_pos(x, gamma / h);
                if (F.ps < tol) F.ps = tol;
                F.kind = HLLC_READY;
            }
        }
        err[i] = hllc_end(gamma, F, f + 4 * i);
    }
}

void hc_flux_exact(int nf, double gamma, double tol, long long n, const double* wl, const double* wr, double* f,
                   int* err)
{
    for (long long i = 0; i < n; i++) {
        if (nf == 3) err[i] = flux_exact<3>(gamma, tol, wl + 3 * i, wr + 3 * i, f + 3 * i);
        else err[i] = flux_exact<4>(gamma, tol, wl + 4 * i, wr + 4 * i, f + 4 * i);
    }
}

// ml, mr [n][4]: cotangents of the HLLC input states for flux cotangent mf
void hc_hllc_vjp(double gamma, double tol, long long n, const double* wl, const double* wr, const double* mf,
                 double* ml, double* mr, int* flag)
{
    for (long long i = 0; i < n; i++)
        flag[i] = hllc_vjp(gamma, tol, wl + 4 * i, wr + 4 * i, mf + 4 * i, ml + 4 * i, mr + 4 * i);
}

// x (y = 0) or y (y = 1) predictor in the global order (rho, u, v, p)
void hc_predict(int y, long long n, const double* s, const double* d, double c, double gamma, double* lo,
                double* hi)
{
    for (long long i = 0; i < n; i++) {
        const double* w = s + 4 * i;
        const double a = sound_speed(gamma, w[0], w[3]);
        if (y) predict_y(w, d + 4 * i, c, a, lo + 4 * i, hi + 4 * i);
        else predict_x(w, d + 4 * i, c, a, lo + 4 * i, hi + 4 * i);
    }
}

// predictor VJP in sweep-local order; quirk = the y-sweep density row
void hc_predict_vjp(int quirk, long long n, const double* s, const double* d, double c, double gamma,
                    const double* mlo, const double* mhi, double* ms, double* md)
{
    for (long long i = 0; i < n; i++) {
        for (int k = 0; k < 4; k++) ms[4 * i + k] = 0.0;
        predict_vjp(s + 4 * i, d + 4 * i, c, gamma, quirk != 0, mlo + 4 * i, mhi + 4 * i, ms + 4 * i, md + 4 * i);
    }
}

void hc_limit(int lim, long long n, const double* a, const double* b, double* d, double* da, double* db)
{
    if (lim == MINMOD) limits<MINMOD>(n, a, b, d, da, db);
    else if (lim == VAN_LEER) limits<VAN_LEER>(n, a, b, d, da, db);
    else limits<MC>(n, a, b, d, da, db);
}

// mu [n][4] = J_c2p(U)^T mw, U = (m, mx, my, E)
void hc_c2p_vjp(double gamma, double vol, long long n, const double* U, const double* mw, double* mu)
{
    for (long long i = 0; i < n; i++) {
        const double* u = U + 4 * i;
        double w[4];
        c2p_2d(gamma, vol, u[0], u[1], u[2], u[3], w);
        c2p_vjp(gamma, vol, u[0], w, mw + 4 * i, mu + 4 * i);
    }
}

// the 2D step's strip height for an nx x ny grid on a device holding `slots`
// step wavefronts at once
int hc_strip_rows(int nx, int ny, long long slots) { return hfv::balanced_strip_rows(nx, ny, slots); }
int hc_exact_strip_rows(int nx, int ny, long long slots) { return hfv::exact_strip_rows(nx, ny, slots); }

}  // extern "C"
