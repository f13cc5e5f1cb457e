// adj2d.hip -- the fixed-dt 2D MUSCL-Hancock + HLLC step on a doubly periodic
// Cartesian grid, as a stateless operation on caller-owned device arrays, and
// its discrete adjoint (the transposed Jacobian of the step applied to a
// cotangent), for gradient-based use such as initial-condition inversion.
//
// The step is the reference's (integrator_muscl_hancock.c:552-1141 with a
// given dt): W = c2p(U); per sweep, limited slopes, half-step predictor and
// HLLC fluxes; U += (dt A) (F_in - F_out).  Its state is the interior of U;
// with periodic boundaries no ghost value is read.  The adjoint maps the
// cotangent lam of U^{n+1} to that of U^n:
//
//   lam^n = lam^{n+1} + J_c2p(U^n)^T (mu_W^x + mu_W^y)
//
// where mu_W^x, mu_W^y collect, per cell, the cotangents that reach its
// primitive state through the x and y sweeps (face flux cotangent
// mu_f = dt A (lam_R - lam_L), HLLC VJP, predictor VJP, limiter VJP).
// Derivatives are one-sided at the forward's branch points, with the branch
// the forward takes (hfv_adjoint.hpp).
//
// Two passes per adjoint step:
//   x pass: one wavefront per 58-column row segment (3 halo columns a side:
//           a cell's x cotangent depends on the faces of columns c-1..c+2,
//           whose states need columns c-3..c+3), neighbours by shuffles;
//           writes lam' = lam^{n+1} + J^T mu_W^x.
//   y pass: one wavefront per 64-column strip of rows, marching upward with
//           a register pipeline (face (i|i+1) at iteration i finishes row i's
//           predictor cotangent and row i-1's total); adds J^T mu_W^y to lam'
//           in place.
#include <hip/hip_runtime.h>

#include "hfv_adjoint.hpp"
#include "hfv_ctrl.hpp"
#include "hfv_gpu_internal.hpp"
#include "hfv_math.hpp"
#include "mh2d_strip.hpp"

namespace hfv {

constexpr int XHALO = 3;
constexpr int XCOLS = LANES - 2 * XHALO;  // 58 output columns per x-pass wavefront

struct AdjGrid {
    int nx, ny, g;
    long long pitch;
    double gamma, tol, volume;
    double cx, cy;   // dt*A/V: predictor coefficients (integrator_muscl_hancock.c:636, 921)
    double fx, fy;   // dt*A: weight of a face flux in the update (806-814, 1092-1100)
    int rows_x;      // rows per x-pass wavefront
    int nsx_x, nsy_x;
    int rows_y;      // rows per y-pass strip
    int nsx_y, nsy_y;
};

// periodic image of index c in [g, g+n)
__device__ __forceinline__ int wrap(int c, int n, int g)
{
    int k = (c - g) % n;
    if (k < 0) k += n;
    return g + k;
}

__device__ __forceinline__ int adj_bits(int e)
{
    return e == ADJ_VACUUM ? HFV_ADJ_VACUUM : (e == ADJ_INVALID ? HFV_ADJ_INVALID : 0);
}

// x pass ---------------------------------------------------------------------
template <int LIM>
__global__ void __launch_bounds__(LANES) k_adj_x(AdjGrid G, Fields4 U, Fields4 Lin, Fields4 Lout, int* flags)
{
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const int sx = b % G.nsx_x, sy = b / G.nsx_x;
    const int x0 = G.g + sx * XCOLS;
    const int c = x0 - XHALO + lane;
    const int cm = wrap(c, G.nx, G.g);
    const bool out = lane >= XHALO && lane < XHALO + XCOLS && c < G.g + G.nx;
    // the face at this lane's left feeds the outputs of lanes lane-2 .. lane+1
    const unsigned long long om = __ballot(out);
    const unsigned long long near = (lane >= 2 ? om >> (lane - 2) : om << (2 - lane)) & 0xFull;
    const bool face_used = lane >= 1 && near != 0ull;
    const int r0 = G.g + sy * G.rows_x, r1 = min(r0 + G.rows_x, G.g + G.ny);
    const double gamma = G.gamma, tol = G.tol;
    // c2p and sound speeds as the forward step takes them (c2p_2d_rn,
    // sound_speed_rn: the same bits as c2p_2d and sound_speed, cheaper)
    const double rvol = rcp_core(G.volume);
    const bool vol_ok = div_core_ok(G.volume);
    int flag = 0;

    for (int r = r0; r < r1; r++) {
        const long long k = (long long)r * G.pitch + cm;
        double u[4], lam[4], w[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            u[f] = U.f[f][k];
            lam[f] = Lin.f[f][k];
        }
        c2p_2d_rn(gamma, G.volume, rvol, vol_ok, u[0], u[1], u[2], u[3], w);
        double d[4], pa[4], pb[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            const double a = w[f] - shfl_from_left(w[f]);
            const double bb = shfl_from_right(w[f]) - w[f];
            d[f] = limit<LIM>(a, bb);
            limit_partials<LIM>(a, bb, &pa[f], &pb[f]);
        }
        double xlo[4], xhi[4];
        predict_x(w, d, G.cx, sound_speed_rn(gamma, w[0], w[3]), xlo, xhi);
        // the face at the lane's left: (left neighbour's xhi | own xlo)
        double hl[4], mf[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            hl[f] = shfl_from_left(xhi[f]);
            mf[f] = G.fx * (lam[f] - shfl_from_left(lam[f]));
        }
        double ml[4], mr[4];
        const int e = hllc_vjp(gamma, tol, hl, xlo, mf, ml, mr);
        if (face_used) flag |= adj_bits(e);
        double mhi[4];
#pragma unroll
        for (int f = 0; f < 4; f++) mhi[f] = shfl_from_right(ml[f]);
        double ms[4] = {0.0, 0.0, 0.0, 0.0}, md[4];
        predict_vjp(w, d, G.cx, gamma, false, mr, mhi, ms, md);
        // slopes d = limit(w - w_left, w_right - w)
        double mw[4], to_left[4], to_right[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            mw[f] = ms[f] + md[f] * (pa[f] - pb[f]);
            to_left[f] = -md[f] * pa[f];
            to_right[f] = md[f] * pb[f];
        }
#pragma unroll
        for (int f = 0; f < 4; f++) mw[f] += shfl_from_right(to_left[f]) + shfl_from_left(to_right[f]);
        if (out) {
            double mu[4];
            c2p_vjp(gamma, G.volume, u[0], w, mw, mu);
#pragma unroll
            for (int f = 0; f < 4; f++) Lout.f[f][k] = lam[f] + mu[f];
        }
    }
    if (__ballot(flag != 0)) {
        // fold the wave's flags into one atomic
        int all = flag;
#pragma unroll
        for (int s = 1; s < LANES; s <<= 1) all |= __shfl_xor(all, s, LANES);
        if (lane == 0) atomicOr(flags, all);
    }
}

// y pass ---------------------------------------------------------------------
// Sweep-local order for the y sweep: (rho, v, u, p); fluxes (mass, my, mx, E).
// rvol, vol_ok: rcp_core(G.volume), div_core_ok(G.volume) (c2p_2d_rn); every
// lane of the wavefront calls this
__device__ __forceinline__ void load_row_y(const AdjGrid& G, const Fields4& U, int q, int cm, double* m, double* wl,
                                           double rvol, bool vol_ok)
{
    const long long k = (long long)wrap(q, G.ny, G.g) * G.pitch + cm;
    double w[4];
    const double u0 = U.f[0][k];
    c2p_2d_rn(G.gamma, G.volume, rvol, vol_ok, u0, U.f[1][k], U.f[2][k], U.f[3][k], w);
    *m = u0;
    wl[0] = w[0];
    wl[1] = w[2];
    wl[2] = w[1];
    wl[3] = w[3];
}

// A copy of v that the compiler cannot identify with v, so that values
// computed from it are computed again rather than kept in registers from an
// earlier computation from v (an empty asm that claims to modify it).
__device__ __forceinline__ void unshare(double* v, int n)
{
    for (int f = 0; f < n; f++) asm volatile("" : "+v"(v[f]));
}

template <int LIM>
__device__ __forceinline__ void slopes_only_y(const double* below, const double* mid, const double* above, double* d)
{
#pragma unroll
    for (int f = 0; f < 4; f++) d[f] = limit<LIM>(mid[f] - below[f], above[f] - mid[f]);
}

template <int LIM>
__device__ __forceinline__ void slopes_y(const double* below, const double* mid, const double* above, double* d,
                                         double* pa, double* pb)
{
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double a = mid[f] - below[f], bb = above[f] - mid[f];
        d[f] = limit<LIM>(a, bb);
        limit_partials<LIM>(a, bb, &pa[f], &pb[f]);
    }
}

// y-sweep predictor in sweep-local order: integrator_muscl_hancock.c:977-1025
// (the density row takes the slope of the transverse velocity u)
__device__ __forceinline__ void predict_y_local(const double* s, const double* d, double c, double a, double* lo,
                                                double* hi)
{
    const double sg[4] = {s[0], s[2], s[1], s[3]};
    const double dg[4] = {d[0], d[2], d[1], d[3]};
    double log_[4], hig[4];
    predict_y(sg, dg, c, a, log_, hig);
    lo[0] = log_[0];
    lo[1] = log_[2];
    lo[2] = log_[1];
    lo[3] = log_[3];
    hi[0] = hig[0];
    hi[1] = hig[2];
    hi[2] = hig[1];
    hi[3] = hig[3];
}

template <int LIM>
__global__ void __launch_bounds__(LANES) __attribute__((amdgpu_waves_per_eu(2))) k_adj_y(AdjGrid G, Fields4 U, Fields4 Lin, Fields4 Lout, int* flags)
{
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const int sx = b % G.nsx_y, sy = b / G.nsx_y;
    const int c = G.g + sx * LANES + lane;
    const bool active = c < G.g + G.nx;
    const int cm = active ? c : G.g;
    const int y0 = G.g + sy * G.rows_y, y1 = min(y0 + G.rows_y, G.g + G.ny);
    const double gamma = G.gamma, tol = G.tol;
    int flag = 0;

    // rows i-1, i, i+1 of the pipeline (i = y0 - 2 at the start).  Only the
    // rows' primitive states and the cotangent sums travel from one
    // iteration to the next; slopes, limiter partials and face states are
    // recomputed from them where needed (the same expressions, so the same
    // values), and a row's mass and incoming cotangent are read again where
    // they are used, which keeps the pipeline within half the register file.
    const double rvol = rcp_core(G.volume);
    const bool vol_ok = div_core_ok(G.volume);
    double Wp[4], Wi[4], Wn[4], m_unused;
    load_row_y(G, U, y0 - 3, cm, &m_unused, Wp, rvol, vol_ok);
    load_row_y(G, U, y0 - 2, cm, &m_unused, Wi, rvol, vol_ok);
    load_row_y(G, U, y0 - 1, cm, &m_unused, Wn, rvol, vol_ok);
    double mloi[4] = {0.0, 0.0, 0.0, 0.0};     // cotangent of row i's low face state
    double accp[4] = {0.0, 0.0, 0.0, 0.0};     // mu_W^y of row i-1 (local order)
    double acci[4] = {0.0, 0.0, 0.0, 0.0};     // of row i

    for (int i = y0 - 2; i <= y1; i++) {
        // row i+2 enters
        double W2[4], m2;
        load_row_y(G, U, i + 2, cm, &m2, W2, rvol, vol_ok);
        double lami[4], lam1[4];
        {
            const long long k = (long long)wrap(i, G.ny, G.g) * G.pitch + cm;
            const long long k1 = (long long)wrap(i + 1, G.ny, G.g) * G.pitch + cm;
#pragma unroll
            for (int f = 0; f < 4; f++) {
                lami[f] = Lin.f[f][k];
                lam1[f] = Lin.f[f][k1];
            }
        }
        // face (i | i+1): row i's high state and row i+1's low state
        double hii[4], lo1[4];
        {
            double d1[4], hi1[4];
            slopes_only_y<LIM>(Wi, Wn, W2, d1);
            predict_y_local(Wn, d1, G.cy, sound_speed_rn(gamma, Wn[0], Wn[3]), lo1, hi1);
        }
        {
            double di[4], loi[4];
            slopes_only_y<LIM>(Wp, Wi, Wn, di);
            predict_y_local(Wi, di, G.cy, sound_speed_rn(gamma, Wi[0], Wi[3]), loi, hii);
        }
        double mf[4];
        mf[0] = G.fy * (lam1[0] - lami[0]);
        mf[1] = G.fy * (lam1[2] - lami[2]);
        mf[2] = G.fy * (lam1[1] - lami[1]);
        mf[3] = G.fy * (lam1[3] - lami[3]);
        double ml[4], mr[4];
        const int e = hllc_vjp(gamma, tol, hii, lo1, mf, ml, mr);
        if (active) flag |= adj_bits(e);
        // row i: predictor and slope cotangents, from its slopes computed
        // again
        unshare(Wp, 4);
        unshare(Wi, 4);
        unshare(Wn, 4);
        double di[4], pai[4], pbi[4];
        slopes_y<LIM>(Wp, Wi, Wn, di, pai, pbi);
        double ms[4] = {0.0, 0.0, 0.0, 0.0}, md[4];
        predict_vjp(Wi, di, G.cy, gamma, true, mloi, ml, ms, md);
        double accn[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            accp[f] += -md[f] * pai[f];
            acci[f] += ms[f] + md[f] * (pai[f] - pbi[f]);
            accn[f] = md[f] * pbi[f];
        }
        // row i-1 is complete
        const int ro = i - 1;
        if (active && ro >= y0 && ro < y1) {
            const long long k = (long long)ro * G.pitch + cm;
            const double w[4] = {Wp[0], Wp[2], Wp[1], Wp[3]};
            const double mw[4] = {accp[0], accp[2], accp[1], accp[3]};
            double mu[4];
            c2p_vjp(gamma, G.volume, U.f[0][k], w, mw, mu);
#pragma unroll
            for (int f = 0; f < 4; f++) Lout.f[f][k] += mu[f];
        }
        // advance the pipeline by one row
#pragma unroll
        for (int f = 0; f < 4; f++) {
            Wp[f] = Wi[f];
            Wi[f] = Wn[f];
            Wn[f] = W2[f];
            mloi[f] = mr[f];
            accp[f] = acci[f];
            acci[f] = accn[f];
        }
    }
    if (__ballot(flag != 0)) {
        int all = flag;
#pragma unroll
        for (int s = 1; s < LANES; s <<= 1) all |= __shfl_xor(all, s, LANES);
        if (lane == 0) atomicOr(flags, all);
    }
}

// forward step ---------------------------------------------------------------
template <int LIM>
__global__ void __launch_bounds__(LANES) k_fwd_step(Grid2D G, double dt, Fields4 U, Fields4 V, int* flags)
{
    const int ntiles = G.nsx * G.nsy;
    const int b = xcd_remap(blockIdx.x, ntiles);
    // the adjoint workload's states are smooth fields without tiles of one
    // state: the step is built without that form (the same results)
    const StripResult r =
        strip_step<LIM, HLLC, false, false, false, false>(G, dt, U, V, U, b % G.nsx, b / G.nsx, threadIdx.x);
    if (__ballot(r.err != ERR_NONE) && threadIdx.x == 0) atomicOr(flags, HFV_ADJ_FLUX_ERROR);
}

}  // namespace hfv

using namespace hfv;

namespace {

bool valid(const hfv_adj2d_config* c)
{
    if (!c || c->num_ghost_cells_side < 2) return false;
    if (c->num_cells_x < 3 || c->num_cells_y < 3) return false;
    if (c->pitch < (long long)c->num_cells_x + 2 * c->num_ghost_cells_side) return false;
    if (c->slope_limiter < MINMOD || c->slope_limiter > MC) return false;
    return true;
}

Fields4 fields(double* const a[4])
{
    Fields4 F;
    for (int f = 0; f < 4; f++) F.f[f] = a[f];
    return F;
}

}  // namespace

extern "C" {

int hfv_fwd2d_step(const hfv_adj2d_config* cfg, double* const U[4], double* const V[4], int* flags, void* stream)
{
    if (!valid(cfg)) return HFV_EINVAL;
    Grid2D G;
    G.nx = cfg->num_cells_x;
    G.ny = cfg->num_cells_y;
    G.g = cfg->num_ghost_cells_side;
    G.tx = G.nx + 2 * G.g;
    G.ty = G.ny + 2 * G.g;
    G.pitch = cfg->pitch;
    G.gamma = cfg->gamma;
    G.gravity = 0.0;
    G.tol = cfg->tol;
    G.area_x = cfg->area_x;
    G.area_y = cfg->area_y;
    G.volume = cfg->volume;
    for (int s = 0; s < 4; s++) G.bc[s] = BC_PERIODIC;
    G.nsx = (G.nx + STRIP_COLS - 1) / STRIP_COLS;
    G.strip_rows = default_strip_rows(G.nx, G.ny);
    G.nsy = (G.ny + G.strip_rows - 1) / G.strip_rows;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid(G.nsx * G.nsy), block(LANES);
    const Fields4 Uf = fields(U), Vf = fields(V);
    switch (cfg->slope_limiter) {
    case MINMOD: hipLaunchKernelGGL(k_fwd_step<MINMOD>, grid, block, 0, s, G, cfg->dt, Uf, Vf, flags); break;
    case VAN_LEER: hipLaunchKernelGGL(k_fwd_step<VAN_LEER>, grid, block, 0, s, G, cfg->dt, Uf, Vf, flags); break;
    default: hipLaunchKernelGGL(k_fwd_step<MC>, grid, block, 0, s, G, cfg->dt, Uf, Vf, flags); break;
    }
    return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_adj2d_step(const hfv_adj2d_config* cfg, double* const U[4], double* const lam_in[4],
                   double* const lam_out[4], int* flags, void* stream)
{
    if (!valid(cfg)) return HFV_EINVAL;
    AdjGrid G;
    G.nx = cfg->num_cells_x;
    G.ny = cfg->num_cells_y;
    G.g = cfg->num_ghost_cells_side;
    G.pitch = cfg->pitch;
    G.gamma = cfg->gamma;
    G.tol = cfg->tol;
    G.volume = cfg->volume;
    G.cx = cfg->dt * cfg->area_x / cfg->volume;
    G.cy = cfg->dt * cfg->area_y / cfg->volume;
    G.fx = cfg->dt * cfg->area_x;
    G.fy = cfg->dt * cfg->area_y;
    G.rows_x = 8;
    G.nsx_x = (G.nx + XCOLS - 1) / XCOLS;
    G.nsy_x = (G.ny + G.rows_x - 1) / G.rows_x;
    G.rows_y = 32;
    G.nsx_y = (G.nx + LANES - 1) / LANES;
    while (G.rows_y > 8 && (long long)G.nsx_y * ((G.ny + G.rows_y - 1) / G.rows_y) < 4096) G.rows_y /= 2;
    G.nsy_y = (G.ny + G.rows_y - 1) / G.rows_y;
    hipStream_t s = (hipStream_t)stream;
    const Fields4 Uf = fields(U), Li = fields(lam_in), Lo = fields(lam_out);
    const dim3 gx(G.nsx_x * G.nsy_x), gy(G.nsx_y * G.nsy_y), block(LANES);
    switch (cfg->slope_limiter) {
    case MINMOD:
        hipLaunchKernelGGL(k_adj_x<MINMOD>, gx, block, 0, s, G, Uf, Li, Lo, flags);
        hipLaunchKernelGGL(k_adj_y<MINMOD>, gy, block, 0, s, G, Uf, Li, Lo, flags);
        break;
    case VAN_LEER:
        hipLaunchKernelGGL(k_adj_x<VAN_LEER>, gx, block, 0, s, G, Uf, Li, Lo, flags);
        hipLaunchKernelGGL(k_adj_y<VAN_LEER>, gy, block, 0, s, G, Uf, Li, Lo, flags);
        break;
    default:
        hipLaunchKernelGGL(k_adj_x<MC>, gx, block, 0, s, G, Uf, Li, Lo, flags);
        hipLaunchKernelGGL(k_adj_y<MC>, gy, block, 0, s, G, Uf, Li, Lo, flags);
        break;
    }
    return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

}  // extern "C"
