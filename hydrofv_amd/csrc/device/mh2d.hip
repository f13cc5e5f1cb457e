// mh2d.hip -- fused 2D MUSCL-Hancock step for gfx950 (one launch per step).
//
// Work decomposition: one wavefront (64 lanes) owns a strip of 60 interior
// columns plus 2 halo columns on each side (lane l <-> column x0-2+l) and a
// run of STRIP_ROWS interior rows, which it sweeps from bottom to top.  Each
// lane keeps a 3-row window of primitive states in registers, so every cell's
// primitive state is derived once per strip from the conserved state in HBM
// (W is never stored between steps).  x-neighbours are exchanged with
// cross-lane shuffles; y-neighbours are the lane's own register window.
//
// Per output cell and step this reads the 4 conserved fields once and writes
// them once: 64 B of algorithmic HBM traffic; halo re-reads hit L2.
//
// The arithmetic per cell and per interface is that of
// integrator_muscl_hancock.c:552-1141 (see hfv_math.hpp), with the update
// order of the reference (x-sweep: left face then right face; y-sweep: bottom
// face then top face, 806-814 and 1092-1100) so that U is bitwise identical.
// Ghost primitives are formed by the boundary mapping of boundary.c:513-777
// (x-ghosts on interior rows, y-ghosts on interior columns) applied to the
// interior cell they copy, which yields the same bits as the reference's
// c2p-then-fill.  The first step of a run reads the caller's primitive
// arrays instead (the reference's first step uses the initial W as given).
#include <hip/hip_runtime.h>

#include "hfv_ctrl.hpp"
#include "hfv_math.hpp"
#include "hfv_gpu_internal.hpp"

namespace hfv {

constexpr int LANES = 64;
constexpr int HALO = 2;
constexpr int STRIP_COLS = LANES - 2 * HALO;  // 60 output columns per wave

struct Fields4 {
    double* f[4];
};

struct Grid2D {
    int nx, ny, g;
    int tx, ty;
    long long pitch;     // row pitch of every device field, in doubles
    double gamma, gravity, tol;
    double area_x, area_y, volume;  // uniform Cartesian geometry (system.c:400-415, 490-503)
    int bc[4];           // x_min, x_max, y_min, y_max
    int strip_rows;
    int nsx, nsy;        // strips in x and y
};

// Boundary mapping of a ghost column (row) index onto the interior cell whose
// primitive state it receives; *flip = 1 when the normal velocity is negated.
__device__ __forceinline__ int map_index(int c, int n, int g, int bc_lo, int bc_hi, int* flip)
{
    *flip = 0;
    if (c < g) {
        if (bc_lo == BC_PERIODIC) return c + n;
        if (bc_lo == BC_TRANSMISSIVE || bc_lo == BC_REFLECTIVE) {
            *flip = (bc_lo == BC_REFLECTIVE);
            return 2 * g - 1 - c;
        }
        return c;  // none: the ghost keeps c2p of its own conserved state
    }
    if (c >= g + n) {
        if (bc_hi == BC_PERIODIC) return c - n;
        if (bc_hi == BC_TRANSMISSIVE || bc_hi == BC_REFLECTIVE) {
            *flip = (bc_hi == BC_REFLECTIVE);
            return 2 * (g + n) - 1 - c;
        }
        return c;
    }
    return c;
}

// Primitive state of cell (row q, column c) as the reference's stencil sees it.
template <bool WMODE>
__device__ __forceinline__ void load_prim(const Grid2D& G, const Fields4& U, const Fields4& W, int q, int c,
                                          double* w)
{
    q = min(max(q, 0), G.ty - 1);
    c = min(max(c, 0), G.tx - 1);
    if constexpr (WMODE) {
        const long long k = (long long)q * G.pitch + c;
        w[0] = W.f[0][k];
        w[1] = W.f[1][k];
        w[2] = W.f[2][k];
        w[3] = W.f[3][k];
    } else {
        int fu, fv;
        const int cs = map_index(c, G.nx, G.g, G.bc[0], G.bc[1], &fu);
        const int qs = map_index(q, G.ny, G.g, G.bc[2], G.bc[3], &fv);
        const long long k = (long long)qs * G.pitch + cs;
        c2p_2d(G.gamma, G.volume, U.f[0][k], U.f[1][k], U.f[2][k], U.f[3][k], w);
        if (fu) w[1] = -w[1];
        if (fv) w[2] = -w[2];
    }
}

// y-predictor of the cell whose own state is wc, with wb below and wt above.
template <int LIM>
__device__ __forceinline__ void ypredict(const double* wb, const double* wc, const double* wt, double c, double gamma,
                                         double* lo, double* hi)
{
    double d[4];
#pragma unroll
    for (int f = 0; f < 4; f++) d[f] = limit<LIM>(wc[f] - wb[f], wt[f] - wc[f]);
    const double a = sound_speed(gamma, wc[0], wc[3]);
    predict_y(wc, d, c, a, lo, hi);
}

// y-sweep flux between the top face of the lower cell (hi) and the bottom
// face of the upper cell (lo): the solver sees (rho, v, u, p) and returns
// (mass, my, mx, E) (integrator_muscl_hancock.c:1051-1067); reordered here to
// (mass, mx, my, E).
template <int SOLVER>
__device__ __forceinline__ int yflux(double gamma, double tol, const double* hi, const double* lo, double* f)
{
    const double l[4] = {hi[0], hi[2], hi[1], hi[3]};
    const double r[4] = {lo[0], lo[2], lo[1], lo[3]};
    double o[4];
    const int e = flux<SOLVER, 4>(gamma, tol, l, r, o);
    f[0] = o[0];
    f[1] = o[2];
    f[2] = o[1];
    f[3] = o[3];
    return e;
}

__device__ __forceinline__ double shfl_from_left(double v) { return __shfl_up(v, 1, LANES); }
__device__ __forceinline__ double shfl_from_right(double v) { return __shfl_down(v, 1, LANES); }

template <int LIM, int SOLVER, bool WMODE>
__global__ void __launch_bounds__(LANES) k_mh2d_step(Grid2D G, RunParams rp, Fields4 Ua, Fields4 Ub, Fields4 W,
                                                     Ctrl* ctrl, int slot)
{
    const Ctrl* in = ctrl + slot;
    Ctrl* out = ctrl + (slot + 1) % 3;
    Ctrl* clear = ctrl + (slot + 2) % 3;
    int dt_zero;
    const StepPlan plan = plan_step(in, rp, &dt_zero);
    const int lane = threadIdx.x;
    if (blockIdx.x == 0 && lane == 0) publish(in, out, clear, plan, dt_zero);
    if (!plan.run) return;

    const Fields4 U = plan.cur ? Ub : Ua;   // state n
    const Fields4 V = plan.cur ? Ua : Ub;   // state n+1
    const double dt = plan.dt;
    const double cx = dt * G.area_x / G.volume;  // dt*A/V, integrator_muscl_hancock.c:636
    const double cy = dt * G.area_y / G.volume;  // :921

    // XCD-aware tile order: consecutive workgroups go to different XCDs, so
    // give each XCD a contiguous range of strips (neighbours share halos in L2).
    const int ntiles = G.nsx * G.nsy;
    const int b = blockIdx.x;
    const int xcd = b % 8, idx = b / 8;
    const int base = ntiles / 8, extra = ntiles % 8;
    const int tile = xcd * base + min(xcd, extra) + idx;
    const int sx = tile % G.nsx;
    const int sy = tile / G.nsx;
    const int g = G.g;
    const int x0 = g + sx * STRIP_COLS;        // first output column of the strip
    const int y0 = g + sy * G.strip_rows;      // first output row
    const int y1 = min(y0 + G.strip_rows, g + G.ny);
    const int col = x0 - HALO + lane;
    const bool out_col = (lane >= HALO) && (lane < HALO + STRIP_COLS) && (col < g + G.nx);
    const bool ghost_lo_col = (sx == 0) && (col == g - 1);               // x_min face flux only
    const bool ghost_hi_col = (sx == G.nsx - 1) && (col == g + G.nx);    // x_max face flux only
    const double gamma = G.gamma, tol = G.tol;

    int err = ERR_NONE;
    long long err_cell = -1;
    double smx = -INFINITY, smy = -INFINITY;

    // ---- prologue: rows y0-2 .. y0+1 ------------------------------------
    double wm2[4], wm1[4], wa[4], wb[4], wc[4];
    load_prim<WMODE>(G, U, W, y0 - 2, col, wm2);
    load_prim<WMODE>(G, U, W, y0 - 1, col, wm1);
    load_prim<WMODE>(G, U, W, y0, col, wa);
    load_prim<WMODE>(G, U, W, y0 + 1, col, wb);
    double yt[4], fyb[4];
    {
        double lo[4], hi_m1[4], lo0[4];
        ypredict<LIM>(wm2, wm1, wa, cy, gamma, lo, hi_m1);   // row y0-1: top face state
        ypredict<LIM>(wm1, wa, wb, cy, gamma, lo0, yt);      // row y0: bottom and top face states
        const int e = yflux<SOLVER>(gamma, tol, hi_m1, lo0, fyb);
        if (e && out_col) { err = e; err_cell = (long long)(y0 - 1) * G.tx + col; }
    }
    if (y0 == g && out_col) {
        // bottom ghost row receives the y_min face flux (1092-1095)
        const long long k = (long long)(g - 1) * G.pitch + col;
        V.f[0][k] = U.f[0][k] - fyb[0] * dt * G.area_y;
        V.f[1][k] = U.f[1][k] - fyb[1] * dt * G.area_y;
        V.f[2][k] = U.f[2][k] - fyb[2] * dt * G.area_y;
        V.f[3][k] = U.f[3][k] - fyb[3] * dt * G.area_y;
    }

    for (int r = y0; r < y1; r++) {
        load_prim<WMODE>(G, U, W, r + 2, col, wc);
        // y: row r+1's face states, then the flux through the top face of row r
        double ylo[4], yhi[4], fyt[4];
        ypredict<LIM>(wa, wb, wc, cy, gamma, ylo, yhi);
        {
            const int e = yflux<SOLVER>(gamma, tol, yt, ylo, fyt);
            if (e && out_col) { err = e; err_cell = (long long)r * G.tx + col; }
        }
        // x: face states of row r from the lane's neighbours
        double wl[4], wr[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            wl[f] = shfl_from_left(wa[f]);
            wr[f] = shfl_from_right(wa[f]);
        }
        double xlo[4], xhi[4];
        {
            double d[4];
#pragma unroll
            for (int f = 0; f < 4; f++) d[f] = limit<LIM>(wa[f] - wl[f], wr[f] - wa[f]);
            const double a = sound_speed(gamma, wa[0], wa[3]);
            predict_x(wa, d, cx, a, xlo, xhi);
        }
        double hl[4];
#pragma unroll
        for (int f = 0; f < 4; f++) hl[f] = shfl_from_left(xhi[f]);
        double fxl[4];
        {
            const int e = flux<SOLVER, 4>(gamma, tol, hl, xlo, fxl);
            if (e && (out_col || ghost_hi_col) && lane >= 1) { err = e; err_cell = (long long)r * G.tx + col - 1; }
        }
        double fxr[4];
#pragma unroll
        for (int f = 0; f < 4; f++) fxr[f] = shfl_from_right(fxl[f]);

        const long long k = (long long)r * G.pitch + col;
        if (out_col) {
            double u[4];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                double x = U.f[f][k];
                x += fxl[f] * dt * G.area_x;
                x -= fxr[f] * dt * G.area_x;
                x += fyb[f] * dt * G.area_y;
                x -= fyt[f] * dt * G.area_y;
                u[f] = x;
            }
            if (G.gravity != 0.0) {
                // source_term.c:213-240 on the updated state
                const double pre = -dt * G.gravity;
                const double kmy = pre * u[0];
                const double k1 = pre * u[2];
                const double k2 = pre * (u[2] + 0.5 * kmy);
                const double k3 = k2;
                const double k4 = pre * (u[2] + kmy);
                u[2] += kmy;
                u[3] += (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0;
            }
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = u[f];
            // CFL maxima of the new state (utils.c:75-92)
            double w[4];
            c2p_2d(gamma, G.volume, u[0], u[1], u[2], u[3], w);
            const double a = sound_speed(gamma, w[0], w[3]);
            const double sx_ = a + fabs(w[1]);
            const double sy_ = a + fabs(w[2]);
            if (sx_ > smx) smx = sx_;
            if (sy_ > smy) smy = sy_;
        } else if (ghost_lo_col) {
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] - fxr[f] * dt * G.area_x;
        } else if (ghost_hi_col) {
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] + fxl[f] * dt * G.area_x;
        }

#pragma unroll
        for (int f = 0; f < 4; f++) {
            wa[f] = wb[f];
            wb[f] = wc[f];
            yt[f] = yhi[f];
            fyb[f] = fyt[f];
        }
    }
    if (y1 == g + G.ny && out_col) {
        // top ghost row receives the y_max face flux (1097-1100)
        const long long k = (long long)(g + G.ny) * G.pitch + col;
#pragma unroll
        for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] + fyb[f] * dt * G.area_y;
    }

    // ---- reductions ----------------------------------------------------
    smx = wave_max(smx);
    smy = wave_max(smy);
    const unsigned long long any_err = __ballot(err != ERR_NONE);
    if (lane == 0) {
        const unsigned long long kx = smax_key(smx), ky = smax_key(smy);
        if (kx) atomicMax(&out->key[0], kx);
        if (ky) atomicMax(&out->key[1], ky);
    }
    if (any_err) {
        const int src = __ffsll((long long)any_err) - 1;
        const int ek = __shfl(err, src, LANES);
        const long long ec = __shfl(err_cell, src, LANES);
        if (lane == 0) report_error(out, ek, ec);
    }
}

// CFL maxima of a primitive state held in memory (the first step's dt).
__global__ void k_smax_from_w(Grid2D G, Fields4 W, Ctrl* ctrl, int slot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    double sx = -INFINITY, sy = -INFINITY;
    if (i < G.nx) {
        const long long k = (long long)(j + G.g) * G.pitch + (i + G.g);
        const double a = sound_speed(G.gamma, W.f[0][k], W.f[3][k]);
        const double vx = a + fabs(W.f[1][k]);
        const double vy = a + fabs(W.f[2][k]);
        if (vx > sx) sx = vx;
        if (vy > sy) sy = vy;
    }
    sx = wave_max(sx);
    sy = wave_max(sy);
    if ((threadIdx.x & 63) == 0) {
        const unsigned long long kx = smax_key(sx), ky = smax_key(sy);
        if (kx) atomicMax(&ctrl[slot].key[0], kx);
        if (ky) atomicMax(&ctrl[slot].key[1], ky);
    }
}

// Primitive state of every cell after a step: c2p everywhere, then the ghost
// fill (x-ghosts on interior rows, y-ghosts on interior columns; the corner
// blocks keep c2p of their own conserved state), i.e. system.c:763-802
// followed by boundary.c:513-777.
__global__ void k_prim_from_u(Grid2D G, Fields4 U, Fields4 W)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (c >= G.tx) return;
    const bool gx = (c < G.g) || (c >= G.g + G.nx);
    const bool gy = (q < G.g) || (q >= G.g + G.ny);
    int cs = c, qs = q, fu = 0, fv = 0;
    if (gx && !gy) cs = map_index(c, G.nx, G.g, G.bc[0], G.bc[1], &fu);
    if (gy && !gx) qs = map_index(q, G.ny, G.g, G.bc[2], G.bc[3], &fv);
    const long long ks = (long long)qs * G.pitch + cs;
    double w[4];
    c2p_2d(G.gamma, G.volume, U.f[0][ks], U.f[1][ks], U.f[2][ks], U.f[3][ks], w);
    if (fu) w[1] = -w[1];
    if (fv) w[2] = -w[2];
    const long long k = (long long)q * G.pitch + c;
    W.f[0][k] = w[0];
    W.f[1][k] = w[1];
    W.f[2][k] = w[2];
    W.f[3][k] = w[3];
}

// controller initialisation for a run: slot 0 = t 0 / step 0 / buffer 0,
// slots 1 and 2 cleared.  (A kernel, not a memset, so it is graph-safe.)
__global__ void k_ctrl_init(Ctrl* ctrl, double t0, long long n0)
{
    const int s = threadIdx.x;
    if (s < 3) {
        ctrl[s].t = t0;
        ctrl[s].dt = 0.0;
        ctrl[s].t_prev = t0;
        ctrl[s].num_steps = n0;
        ctrl[s].key[0] = 0ull;
        ctrl[s].key[1] = 0ull;
        ctrl[s].cur = 0;
        ctrl[s].err_kind = ERR_NONE;
        ctrl[s].err_cell = -1;
    }
}

}  // namespace hfv

// ===========================================================================
// Host side: state object and launchers (C ABI in hfv_gpu.h)
// ===========================================================================
using namespace hfv;

struct hfv_mh2d {
    hfv_mh2d_config cfg;
    Grid2D G;
    RunParams rp;
    Fields4 U[2];
    Fields4 W;
    Ctrl* ctrl;
    double* slab;          // one allocation for all fields
    long long launches;    // step launches issued (selects the controller slot)
    hipGraphExec_t graph;
    int graph_len;
    int graph_phase;
    hipStream_t graph_stream;
};

namespace {

template <int LIM, int SOLVER>
void launch_step_t(hfv_mh2d* h, bool wmode, hipStream_t s)
{
    const int slot = (int)(h->launches % 3);
    const dim3 grid(h->G.nsx * h->G.nsy), block(LANES);
    if (wmode) {
        hipLaunchKernelGGL((k_mh2d_step<LIM, SOLVER, true>), grid, block, 0, s, h->G, h->rp, h->U[0], h->U[1],
                           h->W, h->ctrl, slot);
    } else {
        hipLaunchKernelGGL((k_mh2d_step<LIM, SOLVER, false>), grid, block, 0, s, h->G, h->rp, h->U[0], h->U[1],
                           h->W, h->ctrl, slot);
    }
    h->launches++;
}

void launch_step(hfv_mh2d* h, bool wmode, hipStream_t s)
{
    const int lim = h->cfg.slope_limiter, sol = h->cfg.riemann_solver;
    if (sol == HLLC) {
        if (lim == MINMOD) launch_step_t<MINMOD, HLLC>(h, wmode, s);
        else if (lim == VAN_LEER) launch_step_t<VAN_LEER, HLLC>(h, wmode, s);
        else launch_step_t<MC, HLLC>(h, wmode, s);
    } else {
        if (lim == MINMOD) launch_step_t<MINMOD, EXACT>(h, wmode, s);
        else if (lim == VAN_LEER) launch_step_t<VAN_LEER, EXACT>(h, wmode, s);
        else launch_step_t<MC, EXACT>(h, wmode, s);
    }
}

}  // namespace

extern "C" {

int hfv_mh2d_create(const hfv_mh2d_config* cfg, hfv_mh2d** out)
{
    *out = nullptr;
    if (cfg->num_cells_x < 2 || cfg->num_cells_y < 2 || cfg->num_ghost_cells_side < 2) return HFV_EINVAL;
    if (cfg->slope_limiter < 1 || cfg->slope_limiter > 3) return HFV_EINVAL;
    if (cfg->riemann_solver != HLLC && cfg->riemann_solver != EXACT) return HFV_EINVAL;
    for (int s = 0; s < 4; s++)
        if (cfg->bc[s] < BC_NONE || cfg->bc[s] > BC_PERIODIC) return HFV_EINVAL;
    if (cfg->strip_rows < 0) return HFV_EINVAL;
    hfv_mh2d* h = new (std::nothrow) hfv_mh2d();
    if (!h) return HFV_ENOMEM;
    h->cfg = *cfg;
    Grid2D& G = h->G;
    G.nx = cfg->num_cells_x;
    G.ny = cfg->num_cells_y;
    G.g = cfg->num_ghost_cells_side;
    G.tx = G.nx + 2 * G.g;
    G.ty = G.ny + 2 * G.g;
    G.pitch = ((long long)G.tx + 31) / 32 * 32;
    G.gamma = cfg->gamma;
    G.gravity = cfg->gravity;
    G.tol = cfg->tol;
    G.area_x = cfg->area_x;
    G.area_y = cfg->area_y;
    G.volume = cfg->volume;
    for (int s = 0; s < 4; s++) G.bc[s] = cfg->bc[s];
    G.strip_rows = cfg->strip_rows > 0 ? cfg->strip_rows : 64;
    G.nsx = (G.nx + STRIP_COLS - 1) / STRIP_COLS;
    G.nsy = (G.ny + G.strip_rows - 1) / G.strip_rows;
    RunParams& rp = h->rp;
    rp.cfl = cfg->cfl;
    rp.shrink_factor = cfg->cfl_initial_shrink_factor;
    rp.shrink_steps = cfg->cfl_initial_shrink_num_steps;
    rp.dim = 2;
    rp.tf = cfg->tf;
    rp.max_steps = cfg->max_steps;
    rp.fixed_dt = cfg->fixed_dt;
    rp.dx = cfg->dx;
    rp.dy = cfg->dy;

    const size_t field = (size_t)G.pitch * G.ty;
    const size_t bytes = field * 12 * sizeof(double) + 3 * sizeof(Ctrl) + 256;
    if (hipMalloc(&h->slab, bytes) != hipSuccess) {
        delete h;
        return HFV_ENOMEM;
    }
    for (int f = 0; f < 4; f++) {
        h->U[0].f[f] = h->slab + field * f;
        h->U[1].f[f] = h->slab + field * (4 + f);
        h->W.f[f] = h->slab + field * (8 + f);
    }
    h->ctrl = reinterpret_cast<Ctrl*>(h->slab + field * 12);
    h->launches = 0;
    h->graph = nullptr;
    h->graph_len = 0;
    *out = h;
    return HFV_OK;
}

void hfv_mh2d_destroy(hfv_mh2d* h)
{
    if (!h) return;
    if (h->graph) hipGraphExecDestroy(h->graph);
    hipFree(h->slab);
    delete h;
}

int hfv_mh2d_upload(hfv_mh2d* h, const double* const W[4], const double* const U[4], void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    const Grid2D& G = h->G;
    const size_t hp = (size_t)G.tx * sizeof(double), dp = (size_t)G.pitch * sizeof(double);
    for (int f = 0; f < 4; f++) {
        if (hipMemcpy2DAsync(h->W.f[f], dp, W[f], hp, hp, G.ty, hipMemcpyHostToDevice, s) != hipSuccess)
            return HFV_EDEVICE;
        // both buffers start from the same conserved state: cells a step does
        // not write (deeper ghost layers, corners) stay valid in either one
        if (hipMemcpy2DAsync(h->U[0].f[f], dp, U[f], hp, hp, G.ty, hipMemcpyHostToDevice, s) != hipSuccess)
            return HFV_EDEVICE;
        if (hipMemcpy2DAsync(h->U[1].f[f], dp, U[f], hp, hp, G.ty, hipMemcpyHostToDevice, s) != hipSuccess)
            return HFV_EDEVICE;
    }
    return HFV_OK;
}

int hfv_mh2d_start(hfv_mh2d* h, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    h->launches = 0;
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(64), 0, s, h->ctrl, 0.0, 0LL);
    const dim3 grid((h->G.nx + 255) / 256, h->G.ny), block(256);
    hipLaunchKernelGGL(k_smax_from_w, grid, block, 0, s, h->G, h->W, h->ctrl, 0);
    launch_step(h, true, s);  // first step from the caller's primitive state
    return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_mh2d_advance(hfv_mh2d* h, long long launches, int use_graph, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) {
        for (long long i = 0; i < launches; i++) launch_step(h, false, s);
        return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
    }
    // Replays of a captured chain of GRAPH_LEN launches (a multiple of 3, so a
    // replay ends on the controller slot it started from).
    const int GRAPH_LEN = 48;
    while (launches > 0) {
        const int phase = (int)(h->launches % 3);
        if (launches < GRAPH_LEN) {
            launch_step(h, false, s);
            launches--;
            continue;
        }
        if (!h->graph || h->graph_phase != phase || h->graph_stream != s) {
            if (h->graph) hipGraphExecDestroy(h->graph);
            h->graph = nullptr;
            hipGraph_t graph;
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return HFV_EDEVICE;
            const long long saved = h->launches;
            for (int i = 0; i < GRAPH_LEN; i++) launch_step(h, false, s);
            h->launches = saved;
            if (hipStreamEndCapture(s, &graph) != hipSuccess) return HFV_EDEVICE;
            if (hipGraphInstantiate(&h->graph, graph, nullptr, nullptr, 0) != hipSuccess) {
                hipGraphDestroy(graph);
                return HFV_EDEVICE;
            }
            hipGraphDestroy(graph);
            h->graph_phase = phase;
            h->graph_stream = s;
        }
        if (hipGraphLaunch(h->graph, s) != hipSuccess) return HFV_EDEVICE;
        h->launches += GRAPH_LEN;
        launches -= GRAPH_LEN;
    }
    return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_mh2d_status(hfv_mh2d* h, hfv_run_status* st, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    Ctrl c;
    const int slot = (int)(h->launches % 3);
    if (hipMemcpyAsync(&c, h->ctrl + slot, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess) return HFV_EDEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return HFV_EDEVICE;
    st->t = c.t;
    st->dt = c.dt;
    st->t_prev = c.t_prev;
    st->num_steps = c.num_steps;
    st->err_kind = c.err_kind;
    st->err_cell = c.err_cell;
    st->cur = c.cur;
    st->smax[0] = smax_from_key(c.key[0]);
    st->smax[1] = smax_from_key(c.key[1]);
    const bool finished = c.err_kind != ERR_NONE || !(c.t < h->rp.tf) ||
                          (h->rp.max_steps >= 0 && c.num_steps >= h->rp.max_steps);
    st->finished = finished ? 1 : 0;
    return HFV_OK;
}

int hfv_mh2d_download(hfv_mh2d* h, double* const W[4], double* const U[4], void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    hfv_run_status st;
    int rc = hfv_mh2d_status(h, &st, stream);
    if (rc) return rc;
    const Grid2D& G = h->G;
    const Fields4 cur = h->U[st.cur];
    const dim3 grid((G.tx + 63) / 64, G.ty), block(64);
    hipLaunchKernelGGL(k_prim_from_u, grid, block, 0, s, G, cur, h->W);
    const size_t hp = (size_t)G.tx * sizeof(double), dp = (size_t)G.pitch * sizeof(double);
    for (int f = 0; f < 4; f++) {
        if (W && hipMemcpy2DAsync(W[f], hp, h->W.f[f], dp, hp, G.ty, hipMemcpyDeviceToHost, s) != hipSuccess)
            return HFV_EDEVICE;
        if (U && hipMemcpy2DAsync(U[f], hp, cur.f[f], dp, hp, G.ty, hipMemcpyDeviceToHost, s) != hipSuccess)
            return HFV_EDEVICE;
    }
    return hipStreamSynchronize(s) == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

const char* hfv_device_error_string(void)
{
    return hipGetErrorString(hipGetLastError());
}

int hfv_mh2d_fields(hfv_mh2d* h, double** U0, double** U1, double** W, long long* pitch)
{
    for (int f = 0; f < 4; f++) {
        if (U0) U0[f] = h->U[0].f[f];
        if (U1) U1[f] = h->U[1].f[f];
        if (W) W[f] = h->W.f[f];
    }
    if (pitch) *pitch = h->G.pitch;
    return HFV_OK;
}

}  // extern "C"
