// mh1d.hip -- 1D MUSCL-Hancock run (muscl_hancock_1d,
// integrator_muscl_hancock.c:26-383) as one resident workgroup.
//
// A shock tube of ~10^3 cells cannot fill the chip, and its time loop is a
// sequence of ~500 dependent steps each made of four grid-wide phases
// (reconstruct, flux, update, c2p + boundary fill [+ geometric source]).
// One workgroup of 1024 lanes runs the whole loop with __syncthreads()
// between phases: no per-step launch or host round trip, and the CFL maximum
// is a workgroup reduction.  The arrays live in global memory (L2-resident
// at these sizes), so any grid length is accepted.
#include <hip/hip_runtime.h>

#include <climits>

#include "hfv_ctrl.hpp"
#include "hfv_math.hpp"
#include "hfv_gpu_internal.hpp"
#include "mh2d_strip.hpp"

namespace hfv {

constexpr int T1D = 1024;

struct Grid1D {
    int n, g, tx;
    int coord;  // 1 cartesian, 4 cylindrical, 5 spherical
    double gamma, tol;
    int bc[2];
};

struct Arrays1D {
    double* w[3];
    double* u[3];
    double* lo[3];
    double* hi[3];
    double* fl[3];
    const double* area;
    const double* mid;
    const double* vol;
    // exact-solver solves that failed after 5000 bisection iterations where
    // the reference's loop goes on past a failure (MUSCL-Hancock faces;
    // random-choice cells), for the host's verbose warnings
    unsigned long long* bisect_fails;
};

// boundary.c:411-511 (x_min first, then x_max, per ghost layer)
__device__ void bc_1d(const Grid1D& G, double* const* w, int t)
{
    const int g = G.g, n = G.n;
    for (int i = t; i < g; i += T1D) {
        if (G.bc[0] == BC_REFLECTIVE || G.bc[0] == BC_TRANSMISSIVE) {
            w[0][g - 1 - i] = w[0][g + i];
            w[1][g - 1 - i] = (G.bc[0] == BC_REFLECTIVE) ? -w[1][g + i] : w[1][g + i];
            w[2][g - 1 - i] = w[2][g + i];
        }
        if (G.bc[1] == BC_REFLECTIVE || G.bc[1] == BC_TRANSMISSIVE) {
            w[0][g + n + i] = w[0][g + n - 1 - i];
            w[1][g + n + i] = (G.bc[1] == BC_REFLECTIVE) ? -w[1][g + n - 1 - i] : w[1][g + n - 1 - i];
            w[2][g + n + i] = w[2][g + n - 1 - i];
        } else if (G.bc[1] == BC_PERIODIC) {
            for (int f = 0; f < 3; f++) {
                w[f][g + n + i] = w[f][g + i];
                w[f][g - 1 - i] = w[f][g + n - 1 - i];
            }
        }
    }
}

__device__ void c2p_all(const Grid1D& G, const Arrays1D& A, int t)
{
    for (int k = t; k < G.tx; k += T1D) {
        double w[3];
        c2p_1d(G.gamma, A.vol[k], A.u[0][k], A.u[1][k], A.u[2][k], w);
        A.w[0][k] = w[0];
        A.w[1][k] = w[1];
        A.w[2][k] = w[2];
    }
}

// system.c:849-917 (1D): conserved state from the primitive one, every cell
__device__ void p2c_all(const Grid1D& G, const Arrays1D& A, int t)
{
    for (int k = t; k < G.tx; k += T1D) {
        const double r = A.w[0][k], vx = A.w[1][k], pr = A.w[2][k], vol = A.vol[k];
        const double m = r * vol;
        A.u[0][k] = m;
        A.u[1][k] = m * vx;
        A.u[2][k] = vol * (0.5 * r * vx * vx + pr / (G.gamma - 1.0));
    }
}

// In place of a limiter: Glimm's random choice method (random_choice_1d,
// integrator_random_choice_1d.c), which samples exact Riemann solutions.
constexpr int RANDOM_CHOICE = -1;

// integrator_random_choice_1d.c:28-43: van der Corput sequence, k1 = 5, k2 = 3
__device__ __forceinline__ double van_der_corput(long long n)
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

// source_term.c:28-43
// Runs steps until the run finishes or 'budget' steps have been taken (the
// host relaunches until finished, so no single launch is unbounded).
// LIM == FIRST_ORDER: godunov_first_order_1d (integrator_godunov_first_order.c:
// 128-231): no reconstruction, and a failed flux ends the run with its error
// (the serial loop's goto at :172-176), reported for the lowest failing
// interface; the update, c2p, boundary fill and geometric source are those
// of the MUSCL-Hancock step.  LIM == RANDOM_CHOICE: random_choice_1d, which
// samples W directly and then fills the boundary, converts to conserved
// variables and adds the geometric source.
// lds_stride > 0: the primitive and flux arrays live in LDS (6 arrays of
// lds_stride doubles of dynamic shared memory) for the launch; W is copied in
// at its start and back at its end.
template <int LIM, int SOLVER>
__global__ void __launch_bounds__(T1D) k_mh1d_run(Grid1D G, RunParams rp, Arrays1D Ain, Ctrl* ctrl, long long budget,
                                                  int lds_stride)
{
    extern __shared__ double lds[];
    __shared__ double red[T1D / 64];
    __shared__ int fail_i, fail_last, fail_kind;
    const int t = threadIdx.x;
    const int g = G.g, n = G.n;
    const double gamma = G.gamma;
    Arrays1D A = Ain;
    if (lds_stride > 0) {
        for (int f = 0; f < 3; f++) {
            A.w[f] = lds + f * lds_stride;
            A.fl[f] = lds + (3 + f) * lds_stride;
        }
        for (int k = t; k < G.tx; k += T1D)
            for (int f = 0; f < 3; f++) A.w[f][k] = Ain.w[f][k];
    }
    if (t == 0) {
        fail_i = INT_MAX;
        fail_last = -1;
        fail_kind = ERR_NONE;
    }
    __syncthreads();

    // CFL maximum of the interior (utils.c:28-53)
    auto cfl_max = [&]() -> double {
        double s = -INFINITY;
        for (int k = g + t; k < g + n; k += T1D) {
            const double a = sound_speed(gamma, A.w[0][k], A.w[2][k]);
            const double v = a + fabs(A.w[1][k]);
            if (v > s) s = v;
        }
        s = wave_max(s);
        if ((t & 63) == 0) red[t >> 6] = s;
        __syncthreads();
        double m = red[0];
        for (int k = 1; k < T1D / 64; k++) m = (red[k] > m) ? red[k] : m;
        __syncthreads();
        return m;
    };

    Ctrl c = *ctrl;
    c.key[0] = smax_key(cfl_max());
    for (long long taken = 0; taken < budget; taken++) {
        int dt_zero;
        const StepPlan plan = plan_step(&c, rp, &dt_zero);
        if (!plan.run) {
            if (dt_zero) c.err_kind = ERR_DT_ZERO;
            break;
        }
        const double dt = plan.dt;
        if constexpr (LIM == RANDOM_CHOICE) {
            // integrator_random_choice_1d.c:160-231: a copy of W (in the lo
            // arrays), then every cell g..g+n samples the exact solution of
            // the problem with its left (theta <= 1/2) or right neighbour
            const double theta = van_der_corput(plan.n + 1);
            for (int k = t; k < G.tx; k += T1D)
                for (int f = 0; f < 3; f++) A.lo[f][k] = A.w[f][k];
            __syncthreads();
            int my_i = -1, my_kind = ERR_NONE;
            for (int i = g + t; i <= g + n; i += T1D) {
                int a, b;
                double speed;
                if (theta <= 0.5) {
                    a = i - 1;
                    b = i;
                    speed = theta * rp.dx / dt;
                } else {
                    // with one ghost layer the reference reads one cell past
                    // its arrays here; the read is clamped to the last cell
                    a = i;
                    b = min(i + 1, G.tx - 1);
                    speed = (theta - 1.0) * rp.dx / dt;
                }
                const Sample S = exact_sample(gamma, G.tol, A.lo[0][a], A.lo[1][a], 0.0, A.lo[2][a], A.lo[0][b],
                                              A.lo[1][b], 0.0, A.lo[2][b], speed);
                A.w[0][i] = S.w.r;
                A.w[1][i] = S.w.u;
                A.w[2][i] = S.w.p;
                if (S.rc != ERR_NONE) {
                    my_i = i;  // the last failure of the serial loop is reported
                    my_kind = S.rc;
                    if (S.rc == ERR_EXACT_BISECT) atomicAdd(A.bisect_fails, 1ull);
                }
            }
            if (my_i >= 0) atomicMax(&fail_last, my_i);
            __syncthreads();
            if (my_i >= 0 && my_i == fail_last) fail_kind = my_kind;
            __syncthreads();
            if (fail_last >= 0) {
                c.err_kind = fail_kind;
                c.err_cell = fail_last;
                c.t_prev = c.t;
                c.t = plan.t + dt;
                c.dt = dt;
                c.num_steps = plan.n + 1;
                break;
            }
            bc_1d(G, A.w, t);
            __syncthreads();
            p2c_all(G, A, t);
            __syncthreads();
        } else if constexpr (LIM == FIRST_ORDER) {
            // interface fluxes 0..n from the cells' own states (:153-171)
            int my_i = INT_MAX, my_kind = ERR_NONE;
            for (int i = t; i <= n; i += T1D) {
                const int k = g + i;
                const double l[3] = {A.w[0][k - 1], A.w[1][k - 1], A.w[2][k - 1]};
                const double r[3] = {A.w[0][k], A.w[1][k], A.w[2][k]};
                double f[3];
                const int e = flux<SOLVER, 3>(gamma, G.tol, l, r, f);
                if (e != ERR_NONE && my_i == INT_MAX) {
                    my_i = i;
                    my_kind = e;
                }
                for (int q = 0; q < 3; q++) A.fl[q][i] = f[q] * dt;
            }
            if (my_i != INT_MAX) atomicMin(&fail_i, my_i);
            __syncthreads();
            if (my_i != INT_MAX && my_i == fail_i) fail_kind = my_kind;
            __syncthreads();
            if (fail_i != INT_MAX) {
                c.err_kind = fail_kind;
                c.err_cell = g + fail_i - 1;
                c.t_prev = c.t;
                c.t = plan.t + dt;
                c.dt = dt;
                c.num_steps = plan.n + 1;
                break;
            }
        } else {
        // interface fluxes 0..n (252-295), each from the reconstructions
        // (146-246) of its two cells g+i-1 and g+i, which the lane computes
        // itself: no reconstruction arrays and no barrier between the two.
        // A failed flux is not reported: the reference overwrites the flux
        // loop's error status with that of the boundary fill that follows
        // (296-321), so the run continues.
        auto reconstruct = [&](int k, double* lo, double* hi) {
            const double s[3] = {A.w[0][k], A.w[1][k], A.w[2][k]};
            const double cc = dt * A.area[k] / A.vol[k];
            const double a = sound_speed(gamma, s[0], s[2]);
            double d[3];
            for (int f = 0; f < 3; f++) d[f] = limit<LIM>(s[f] - A.w[f][k - 1], A.w[f][k + 1] - s[f]);
            predict_1d(s, d, cc, a, lo, hi);
        };
        for (int i = t; i <= n; i += T1D) {
            double lo_l[3], hi_l[3], lo_r[3], hi_r[3];
            reconstruct(g + i - 1, lo_l, hi_l);
            reconstruct(g + i, lo_r, hi_r);
            double f[3];
            const int e = flux<SOLVER, 3>(gamma, G.tol, hi_l, lo_r, f);
            if (e != ERR_NONE) {
                // the reference's loop goes on (its error is overwritten)
                f[0] = f[1] = f[2] = __builtin_nan("");
                if (e == ERR_EXACT_BISECT) atomicAdd(A.bisect_fails, 1ull);
            }
            for (int q = 0; q < 3; q++) A.fl[q][i] = f[q] * dt;
        }
        }
        bool key_done = false;
        if constexpr (LIM != RANDOM_CHOICE) {
            __syncthreads();
            // per cell, by one lane: the update in the reference's order (a
            // cell first gains its left face's flux, then loses its right
            // face's, both scaled by its own face-array entry, 287-294), then
            // c2p (297-321), then, without a source term, its share of the
            // CFL maximum of the interior (utils.c:28-53; the boundary fill
            // leaves the interior as it is)
            const bool fused_cfl = !(G.coord == 4 || G.coord == 5);
            double smax = -INFINITY;
            for (int k = t; k < G.tx; k += T1D) {
                if (k >= g - 1 && k <= g + n) {
                    const double ak = A.area[k];
                    for (int q = 0; q < 3; q++) {
                        double x = A.u[q][k];
                        if (k >= g) x += A.fl[q][k - g] * ak;
                        if (k < g + n) x -= A.fl[q][k - g + 1] * ak;
                        A.u[q][k] = x;
                    }
                }
                double w[3];
                c2p_1d(G.gamma, A.vol[k], A.u[0][k], A.u[1][k], A.u[2][k], w);
                A.w[0][k] = w[0];
                A.w[1][k] = w[1];
                A.w[2][k] = w[2];
                if (fused_cfl && k >= g && k < g + n) {
                    const double v = sound_speed(gamma, w[0], w[2]) + fabs(w[1]);
                    if (v > smax) smax = v;
                }
            }
            __syncthreads();
            bc_1d(G, A.w, t);
            if (fused_cfl) {
                smax = wave_max(smax);
                if ((t & 63) == 0) red[t >> 6] = smax;
            }
            __syncthreads();
            if (fused_cfl) {
                double m = red[0];
                for (int k = 1; k < T1D / 64; k++) m = (red[k] > m) ? red[k] : m;
                c.key[0] = smax_key(m);
                key_done = true;
            }
        }
        if (G.coord == 4 || G.coord == 5) {
            // source_term.c:45-175: RK4 per interior cell, then c2p and fill
            const double alpha = (G.coord == 4) ? 1.0 : 2.0;
            for (int k = g + t; k < g + n; k += T1D) {
                double u[3] = {A.u[0][k], A.u[1][k], A.u[2][k]};
                geometry_source_rk4(alpha, gamma, A.vol[k], A.mid[k], A.w[0][k], dt, u);
                A.u[0][k] = u[0];
                A.u[1][k] = u[1];
                A.u[2][k] = u[2];
            }
            __syncthreads();
            c2p_all(G, A, t);
            __syncthreads();
            bc_1d(G, A.w, t);
            __syncthreads();
        }
        c.t_prev = c.t;
        c.t = plan.t + dt;
        c.dt = dt;
        c.num_steps = plan.n + 1;
        if (!key_done) c.key[0] = smax_key(cfl_max());
    }
    if (lds_stride > 0) {
        __syncthreads();
        for (int k = t; k < G.tx; k += T1D)
            for (int f = 0; f < 3; f++) Ain.w[f][k] = A.w[f][k];
    }
    if (t == 0) *ctrl = c;
}

__global__ void k_probe_flux3(int solver, double gamma, double tol, long long n, const double* wl, const double* wr,
                              double* f, int* err)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double l[3], r[3], o[3];
    for (int q = 0; q < 3; q++) {
        l[q] = wl[3 * i + q];
        r[q] = wr[3 * i + q];
    }
    const int e = (solver == HLLC) ? flux<HLLC, 3>(gamma, tol, l, r, o) : flux<EXACT, 3>(gamma, tol, l, r, o);
    for (int q = 0; q < 3; q++) f[3 * i + q] = o[q];
    err[i] = e;
}

__global__ void k_probe_flux4(int solver, double gamma, double tol, long long n, const double* wl, const double* wr,
                              double* f, int* err)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double l[4], r[4], o[4];
    for (int q = 0; q < 4; q++) {
        l[q] = wl[4 * i + q];
        r[q] = wr[4 * i + q];
    }
    const int e = (solver == HLLC) ? flux<HLLC, 4>(gamma, tol, l, r, o) : flux<EXACT, 4>(gamma, tol, l, r, o);
    for (int q = 0; q < 4; q++) f[4 * i + q] = o[q];
    err[i] = e;
}

constexpr int PROBE_SPLIT_MISMATCH = 99;

__device__ __forceinline__ bool same_bits(double a, double b)
{
    return __double_as_longlong(a) == __double_as_longlong(b) || (isnan(a) && isnan(b));
}

// The 2D kernels' HLLC form: hllc_begin, the wavefront's two-rarefaction
// powers, hllc_end.  Lane l of a one-wavefront block takes state pairs i and
// n + i (when they exist), so every lane holds two faces like a strip row.
// Lanes past the end work on pair 0 and store nothing.
__global__ void __launch_bounds__(LANES) k_probe_flux4_wave(double gamma, double tol, long long n, long long half,
                                                            const double* wl, const double* wr, double* f, int* err)
{
    __shared__ double q[6 * LANES];
    __shared__ double pow_lds[POW_TABS_LEN];
    const int lane = threadIdx.x;
    stage_pow_tables(pow_lds, lane);
    const long long i = (long long)blockIdx.x * LANES + lane;
    const long long ia = (i < half) ? i : 0, ib = (half + i < n) ? half + i : 0;
    double la[4], ra[4], lb[4], rb[4];
    for (int k = 0; k < 4; k++) {
        la[k] = wl[4 * ia + k];
        ra[k] = wr[4 * ia + k];
        lb[k] = wl[4 * ib + k];
        rb[k] = wr[4 * ib + k];
    }
    HllcFace A, B;
    hllc_begin(gamma, tol, la, ra, A);
    hllc_begin(gamma, tol, lb, rb, B);
    trrs_powers(gamma, tol, A, B, q, lane, pow_lds);
    double oa[4], ob[4];
    int ea = hllc_end(gamma, A, oa);
    int eb = hllc_end(gamma, B, ob);
    // the split form must give the plain flux_hllc<4> bits (the kernels use
    // both for the same faces, depending on the strip layout)
    double pa[4], pb[4];
    const int fa = flux_hllc<4>(gamma, tol, la, ra, pa);
    const int fb = flux_hllc<4>(gamma, tol, lb, rb, pb);
    bool sa = fa == ea, sb = fb == eb;
    for (int k = 0; k < 4 && fa == 0 && ea == 0; k++) sa = sa && same_bits(pa[k], oa[k]);
    for (int k = 0; k < 4 && fb == 0 && eb == 0; k++) sb = sb && same_bits(pb[k], ob[k]);
    if (!sa) ea = PROBE_SPLIT_MISMATCH;
    if (!sb) eb = PROBE_SPLIT_MISMATCH;
    if (i < half) {
        for (int k = 0; k < 4; k++) f[4 * ia + k] = oa[k];
        err[ia] = ea;
    }
    if (half + i < n) {
        for (int k = 0; k < 4; k++) f[4 * ib + k] = ob[k];
        err[ib] = eb;
    }
}

__global__ void k_probe_pow(int impl, long long n, const double* x, const double* y, double* out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = impl ? pow_pos(x[i], y[i]) : pow(x[i], y[i]);
}

// impl 0: sqrt(); 1: sqrt_rn over the wavefront (one argument per lane);
// 2: sqrt_core alone, for arguments in [2^-767, +inf) only.
__global__ void __launch_bounds__(LANES) k_probe_sqrt(int impl, long long n, const double* x, double* out)
{
    const long long i = (long long)blockIdx.x * LANES + threadIdx.x;
    const double v = x[(i < n) ? i : 0];
    double r;
    if (impl == 1) {
        double g[1] = {v};
        sqrt_rn<1>(g);
        r = g[0];
    } else {
        r = (impl == 2) ? sqrt_core(v) : sqrt(v);
    }
    if (i < n) out[i] = r;
}

__global__ void __launch_bounds__(LANES) k_probe_div(int impl, long long n, const double* a, const double* b,
                                                     double* out)
{
    const long long i = (long long)blockIdx.x * LANES + threadIdx.x;
    const double x = a[(i < n) ? i : 0], y = b[(i < n) ? i : 0];
    double r;
    if (impl == 1) {
        if (wave_all(div_core_ok(x) & div_core_ok(y))) r = div_core(x, y, rcp_core(y));
        else r = x / y;
    } else {
        r = (impl == 2) ? div_core(x, y, rcp_core(y)) : x / y;
    }
    if (i < n) out[i] = r;
}

__global__ void k_probe_limit(int lim, long long n, const double* a, const double* b, double* out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = limit_rt(lim, a[i], b[i]);
}

}  // namespace hfv

using namespace hfv;

namespace {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t n) { return hipMalloc(&p, n * sizeof(T) + 64) == hipSuccess; }
};

}  // namespace

extern "C" {

int hfv_mh1d_run(const hfv_mh1d_config* cfg, double* const W[3], double* const U[3], const double* area,
                 const double* mid, const double* volume, hfv_run_status* st, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (cfg->first_order != 0 && cfg->first_order != 1) return HFV_EINVAL;
    if (cfg->random_choice != 0 && cfg->random_choice != 1) return HFV_EINVAL;
    if (cfg->steps_start < 0 || !(cfg->t_start >= 0.0) || cfg->t_pause < 0.0) return HFV_EINVAL;
    const bool order1 = cfg->first_order || cfg->random_choice;  // one ghost layer suffices
    if (cfg->num_cells_x < 2 || cfg->num_ghost_cells_side < (order1 ? 1 : 2)) return HFV_EINVAL;
    // the boundary fill mirrors and wraps interior cells onto the ghosts
    if (cfg->num_cells_x < cfg->num_ghost_cells_side) return HFV_EINVAL;
    if (!order1 && (cfg->slope_limiter < 1 || cfg->slope_limiter > 3)) return HFV_EINVAL;
    if (!cfg->random_choice && cfg->riemann_solver != HLLC && cfg->riemann_solver != EXACT) return HFV_EINVAL;
    for (int q = 0; q < 2; q++)
        // BC_CUSTOM: the device leaves those ghosts as c2p of their own
        // state; the host applies the hook between steps (hfv_integrator.c)
        if (cfg->bc[q] < BC_NONE || cfg->bc[q] > BC_CUSTOM) return HFV_EINVAL;
    Grid1D G;
    G.n = cfg->num_cells_x;
    G.g = cfg->num_ghost_cells_side;
    G.tx = G.n + 2 * G.g;
    G.coord = cfg->coord_sys_flag;
    G.gamma = cfg->gamma;
    G.tol = cfg->tol;
    G.bc[0] = cfg->bc[0];
    G.bc[1] = cfg->bc[1];
    RunParams rp;
    rp.cfl = cfg->cfl;
    rp.shrink_factor = cfg->cfl_initial_shrink_factor;
    rp.shrink_steps = cfg->cfl_initial_shrink_num_steps;
    rp.dim = 1;
    rp.tf = cfg->tf;
    rp.max_steps = cfg->max_steps;
    rp.fixed_dt = cfg->fixed_dt;
    rp.dx = cfg->dx;
    rp.dy = 0.0;
    if (cfg->t_pause > 0.0) {
        rp.t_pause = cfg->t_pause;
        rp.resume_n = cfg->steps_start;
    }

    const size_t tx = (size_t)G.tx;
    const size_t stride = (tx + 2 + 7) / 8 * 8;
    DevBuf<double> buf;
    DevBuf<Ctrl> dctrl;
    DevBuf<unsigned long long> dfails;
    if (!buf.alloc(stride * 18) || !dctrl.alloc(1) || !dfails.alloc(1)) return HFV_ENOMEM;
    Arrays1D A;
    for (int f = 0; f < 3; f++) {
        A.w[f] = buf.p + stride * f;
        A.u[f] = buf.p + stride * (3 + f);
        A.lo[f] = buf.p + stride * (6 + f);
        A.hi[f] = buf.p + stride * (9 + f);
        A.fl[f] = buf.p + stride * (12 + f);
    }
    double* geo = buf.p + stride * 15;
    A.bisect_fails = dfails.p;
    if (hipMemsetAsync(dfails.p, 0, sizeof(unsigned long long), s) != hipSuccess) return HFV_EDEVICE;
    A.area = geo;
    A.mid = geo + stride;
    A.vol = geo + 2 * stride;
    for (int f = 0; f < 3; f++) {
        if (hipMemcpyAsync(A.w[f], W[f], tx * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) return HFV_EDEVICE;
        if (hipMemcpyAsync(A.u[f], U[f], tx * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) return HFV_EDEVICE;
    }
    if (hipMemcpyAsync(geo, area, (tx + 1) * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(geo + stride, mid, tx * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(geo + 2 * stride, volume, tx * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
        return HFV_EDEVICE;
    Ctrl c0 = {};
    c0.err_cell = -1;
    c0.t = cfg->t_start;
    c0.t_prev = cfg->t_start;
    c0.num_steps = cfg->steps_start;
    if (hipMemcpyAsync(dctrl.p, &c0, sizeof(Ctrl), hipMemcpyHostToDevice, s) != hipSuccess) return HFV_EDEVICE;

    const int lim = cfg->slope_limiter;
    const bool hllc = cfg->riemann_solver == HLLC;
    const long long budget = 1LL << 16;  // steps per launch
    // W and the fluxes in LDS when they fit in 60 KiB (about 1250 cells)
    const int lds_stride = tx + 2;
    const bool lds_fits = (size_t)6 * lds_stride * sizeof(double) <= 60 * 1024;
    const unsigned lds_bytes = lds_fits ? (unsigned)(6 * lds_stride * sizeof(double)) : 0u;
    const int lds_arg = lds_fits ? lds_stride : 0;
    Ctrl c;
    for (;;) {
#define HFV_L1D(L, S) \
    hipLaunchKernelGGL((k_mh1d_run<L, S>), dim3(1), dim3(T1D), lds_bytes, s, G, rp, A, dctrl.p, budget, lds_arg)
        if (cfg->random_choice) {
            HFV_L1D(RANDOM_CHOICE, EXACT);
        } else if (cfg->first_order) {
            if (hllc) HFV_L1D(FIRST_ORDER, HLLC);
            else HFV_L1D(FIRST_ORDER, EXACT);
        } else if (hllc) {
            if (lim == MINMOD) HFV_L1D(MINMOD, HLLC);
            else if (lim == VAN_LEER) HFV_L1D(VAN_LEER, HLLC);
            else HFV_L1D(MC, HLLC);
        } else {
            if (lim == MINMOD) HFV_L1D(MINMOD, EXACT);
            else if (lim == VAN_LEER) HFV_L1D(VAN_LEER, EXACT);
            else HFV_L1D(MC, EXACT);
        }
#undef HFV_L1D
        if (hipGetLastError() != hipSuccess) return HFV_EDEVICE;
        if (hipMemcpyAsync(&c, dctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess) return HFV_EDEVICE;
        if (hipStreamSynchronize(s) != hipSuccess) return HFV_EDEVICE;
        const bool done = c.err_kind != ERR_NONE || !(c.t < rp.tf) || (rp.max_steps >= 0 && c.num_steps >= rp.max_steps) ||
                          (c.num_steps > rp.resume_n && c.t >= rp.t_pause);
        if (done) break;
    }
    for (int f = 0; f < 3; f++) {
        if (hipMemcpyAsync(W[f], A.w[f], tx * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return HFV_EDEVICE;
        if (hipMemcpyAsync(U[f], A.u[f], tx * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return HFV_EDEVICE;
    }
    unsigned long long fails = 0;
    if (hipMemcpyAsync(&fails, dfails.p, sizeof fails, hipMemcpyDeviceToHost, s) != hipSuccess) return HFV_EDEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return HFV_EDEVICE;
    st->bisect_fails = (long long)fails;
    st->t = c.t;
    st->dt = c.dt;
    st->t_prev = c.t_prev;
    st->num_steps = c.num_steps;
    st->err_kind = c.err_kind;
    st->err_cell = c.err_cell;
    st->cur = 0;
    st->smax[0] = smax_from_key(c.key[0]);
    st->smax[1] = 0.0;
    // a run stopped at a pause has not finished
    st->finished = (c.err_kind != ERR_NONE || !(c.t < rp.tf) || (rp.max_steps >= 0 && c.num_steps >= rp.max_steps)) ? 1 : 0;
    return HFV_OK;
}

int hfv_probe_pow(int impl, long long n, const double* x, const double* y, double* out)
{
    if (n < 0 || (impl != 0 && impl != 1)) return HFV_EINVAL;
    if (n == 0) return HFV_OK;
    DevBuf<double> dx, dy, dout;
    if (!dx.alloc((size_t)n) || !dy.alloc((size_t)n) || !dout.alloc((size_t)n)) return HFV_ENOMEM;
    if (hipMemcpy(dx.p, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dy.p, y, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return HFV_EDEVICE;
    hipLaunchKernelGGL(k_probe_pow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, impl, n, dx.p, dy.p, dout.p);
    if (hipGetLastError() != hipSuccess) return HFV_EDEVICE;
    if (hipMemcpy(out, dout.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return HFV_EDEVICE;
    return HFV_OK;
}

int hfv_probe_flux(int nf, int solver, double gamma, double tol, long long n, const double* wl, const double* wr,
                   double* f, int* err)
{
    if ((nf != 3 && nf != 4) || n < 0) return HFV_EINVAL;
    if (n == 0) return HFV_OK;
    DevBuf<double> dl, dr, df;
    DevBuf<int> de;
    const size_t m = (size_t)n * nf;
    if (!dl.alloc(m) || !dr.alloc(m) || !df.alloc(m) || !de.alloc((size_t)n)) return HFV_ENOMEM;
    if (hipMemcpy(dl.p, wl, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dr.p, wr, m * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return HFV_EDEVICE;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (nf == 3) {
        hipLaunchKernelGGL(k_probe_flux3, grid, block, 0, 0, solver, gamma, tol, n, dl.p, dr.p, df.p, de.p);
    } else if (solver == HLLC) {
        const long long half = (n + 1) / 2;
        const dim3 wgrid((unsigned)((half + LANES - 1) / LANES)), wblock(LANES);
        hipLaunchKernelGGL(k_probe_flux4_wave, wgrid, wblock, 0, 0, gamma, tol, n, half, dl.p, dr.p, df.p, de.p);
    } else {
        hipLaunchKernelGGL(k_probe_flux4, grid, block, 0, 0, solver, gamma, tol, n, dl.p, dr.p, df.p, de.p);
    }
    if (hipGetLastError() != hipSuccess) return HFV_EDEVICE;
    if (hipMemcpy(f, df.p, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(err, de.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return HFV_EDEVICE;
    return HFV_OK;
}

int hfv_probe_sqrt(int impl, long long n, const double* x, double* out)
{
    if (n < 0 || impl < 0 || impl > 2) return HFV_EINVAL;
    if (n == 0) return HFV_OK;
    DevBuf<double> dx, dout;
    if (!dx.alloc((size_t)n) || !dout.alloc((size_t)n)) return HFV_ENOMEM;
    if (hipMemcpy(dx.p, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return HFV_EDEVICE;
    hipLaunchKernelGGL(k_probe_sqrt, dim3((unsigned)((n + LANES - 1) / LANES)), dim3(LANES), 0, 0, impl, n, dx.p,
                       dout.p);
    if (hipGetLastError() != hipSuccess) return HFV_EDEVICE;
    if (hipMemcpy(out, dout.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return HFV_EDEVICE;
    return HFV_OK;
}

int hfv_probe_div(int impl, long long n, const double* a, const double* b, double* out)
{
    if (n < 0 || impl < 0 || impl > 2) return HFV_EINVAL;
    if (n == 0) return HFV_OK;
    DevBuf<double> da, db, dout;
    if (!da.alloc((size_t)n) || !db.alloc((size_t)n) || !dout.alloc((size_t)n)) return HFV_ENOMEM;
    if (hipMemcpy(da.p, a, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db.p, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return HFV_EDEVICE;
    hipLaunchKernelGGL(k_probe_div, dim3((unsigned)((n + LANES - 1) / LANES)), dim3(LANES), 0, 0, impl, n, da.p, db.p,
                       dout.p);
    if (hipGetLastError() != hipSuccess) return HFV_EDEVICE;
    if (hipMemcpy(out, dout.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return HFV_EDEVICE;
    return HFV_OK;
}

int hfv_probe_limit(int limiter, long long n, const double* a, const double* b, double* out)
{
    if (limiter < 1 || limiter > 3 || n < 0) return HFV_EINVAL;
    if (n == 0) return HFV_OK;
    DevBuf<double> da, db, dout;
    if (!da.alloc((size_t)n) || !db.alloc((size_t)n) || !dout.alloc((size_t)n)) return HFV_ENOMEM;
    if (hipMemcpy(da.p, a, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db.p, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return HFV_EDEVICE;
    hipLaunchKernelGGL(k_probe_limit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, limiter, n, da.p, db.p,
                       dout.p);
    if (hipGetLastError() != hipSuccess) return HFV_EDEVICE;
    if (hipMemcpy(out, dout.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return HFV_EDEVICE;
    return HFV_OK;
}

}  // extern "C"
