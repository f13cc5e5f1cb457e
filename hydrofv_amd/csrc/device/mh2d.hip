// mh2d.hip -- fused 2D MUSCL-Hancock step for gfx950 (one launch per step;
// the per-wavefront work is in mh2d_strip.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "hfv_ctrl.hpp"
#include "hfv_gpu_internal.hpp"
#include "hfv_math.hpp"
#include "mh2d_strip.hpp"

namespace hfv {

template <int LIM, int SOLVER, bool WMODE, bool BRANCHY>
__global__ void __launch_bounds__(LANES) __attribute__((amdgpu_waves_per_eu(2))) k_mh2d_step(Grid2D G, RunParams rp, Fields4 Ua, Fields4 Ub, Fields4 W,
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
    const int tile = xcd_remap(blockIdx.x, G.nsx * G.nsy);
    const StripResult r =
        strip_step<LIM, SOLVER, WMODE, BRANCHY>(G, plan.dt, U, V, W, tile % G.nsx, tile / G.nsx, lane);
    strip_finish(r, out, lane);
}

// The exact-solver step with queued faces (Deferred, mh2d_strip.hpp).
//
// Plain form (a step that reads the caller's primitive state): the strip
// pass, then k_solve_faces, k_solve_slow and k_update_deferred.
//
// Split form (every other step): the faces that need the iteration lie in a
// small share of the strip tiles (the blast region: about 4% of them in a
// Sedov run), and the slowest of them take the reference's 100 Newton
// iterations, a serial chain of ~100 us.  So the tiles are split in two
// passes with queues of their own (views A and B of the queue arrays): the
// previous step's deferred-cell update (plan_tiles) lists as A the tiles that
// queued faces in that step and their neighbours, and the rest as B.  Pass A
// runs first, its tiles cut in DEFER_NSUB runs of rows so that it ends
// early, then the first pass over its faces (k_solve_faces); its set-aside
// faces (k_solve_slow) then run on a second stream while pass B runs and
// solves its own faces (usually none); k_update_deferred takes both views.
// Grids of more than DEFER_PLAN_MAX_TILES tiles take the plain form.
// Every face and cell gets the same operations on the same values in either
// form: which pass handles a tile changes no result.
//
// The queue counters come in three sets per view, one per controller slot: a
// step uses its slot's sets, and its strip pass clears the next slot's, which
// no launch of this step reads (k_defer_reset clears the sets of a slot that
// no step has cleared: the first step of a run description).
constexpr int DEFER_NSUB = 6;     // row runs per tile in pass A
constexpr int DEFER_MAX_A = 1024; // tiles in pass A at most (the rest go to B)

__device__ __forceinline__ Deferred defer_slot(Deferred D, int slot)
{
    D.count += 3 * slot;
    return D;
}

// Block 0 of a step's strip pass: publish the controller state and clear the
// next slot's queue counters of both views (count_all: [view][slot][3]).
__device__ __forceinline__ void step_publish(const Ctrl* in, Ctrl* out, Ctrl* clear, const StepPlan& plan,
                                             int dt_zero, unsigned long long* count_all, int slot)
{
    publish(in, out, clear, plan, dt_zero);
    const int next = 3 * ((slot + 1) % 3);
    for (int v = 0; v < 2; v++)
        for (int k = 0; k < 3; k++) count_all[9 * v + next + k] = 0ull;
}

template <int LIM, bool WMODE>
__global__ void __launch_bounds__(LANES) k_mh2d_step_defer(Grid2D G, RunParams rp, Fields4 Ua, Fields4 Ub,
                                                           Fields4 W, Ctrl* ctrl, int slot, Deferred D,
                                                           unsigned long long* count_all)
{
    const Ctrl* in = ctrl + slot;
    Ctrl* out = ctrl + (slot + 1) % 3;
    Ctrl* clear = ctrl + (slot + 2) % 3;
    int dt_zero;
    const StepPlan plan = plan_step(in, rp, &dt_zero);
    const int lane = threadIdx.x;
    if (blockIdx.x == 0 && lane == 0) step_publish(in, out, clear, plan, dt_zero, count_all, slot);
    if (!plan.run) return;
    const Fields4 U = plan.cur ? Ub : Ua;
    const Fields4 V = plan.cur ? Ua : Ub;
    const int tile = xcd_remap(blockIdx.x, G.nsx * G.nsy);
    const StripResult r = strip_step<LIM, EXACT, WMODE, false, true>(G, plan.dt, U, V, W, tile % G.nsx,
                                                                     tile / G.nsx, lane, defer_slot(D, slot));
    strip_finish(r, out, lane);
}

// The split form's passes (plan: plan_tiles).  PASS 0 (A): block b takes run
// b % DEFER_NSUB of tile A[b / DEFER_NSUB]; PASS 1 (B): block b takes tile
// xcd_remap(b) unless it is in A, and its block 0 publishes the step.
template <int LIM, int PASS>
__global__ void __launch_bounds__(LANES) k_mh2d_step_split(Grid2D G, RunParams rp, Fields4 Ua, Fields4 Ub,
                                                           Ctrl* ctrl, int slot, Deferred D,
                                                           unsigned long long* count_all, const int* plan)
{
    const Ctrl* in = ctrl + slot;
    Ctrl* out = ctrl + (slot + 1) % 3;
    Ctrl* clear = ctrl + (slot + 2) % 3;
    int dt_zero;
    const StepPlan sp = plan_step(in, rp, &dt_zero);
    const int lane = threadIdx.x;
    if (PASS == 1 && blockIdx.x == 0 && lane == 0) step_publish(in, out, clear, sp, dt_zero, count_all, slot);
    if (!sp.run) return;
    const int ntiles = G.nsx * G.nsy;
    int tile, sub = 0, nsub = 1;
    if constexpr (PASS == 0) {
        const int pos = blockIdx.x / DEFER_NSUB;
        if (pos >= plan[0]) return;
        tile = plan[2 + pos];
        if (tile < 0 || tile >= ntiles) return;
        sub = blockIdx.x % DEFER_NSUB;
        nsub = DEFER_NSUB;
        int y0, y1;
        strip_rows_of(G, tile / G.nsx, sub, nsub, &y0, &y1);
        if (y0 >= y1) return;
    } else {
        tile = xcd_remap(blockIdx.x, ntiles);
        if (plan[2 + ntiles + tile] != 0) return;
    }
    const Fields4 U = sp.cur ? Ub : Ua;
    const Fields4 V = sp.cur ? Ua : Ub;
    const Fields4 W{};  // unused: these steps read the conserved state
    const StripResult r = strip_step<LIM, EXACT, false, false, true>(G, sp.dt, U, V, W, tile % G.nsx, tile / G.nsx,
                                                                     lane, defer_slot(D, slot), sub, nsub);
    strip_finish(r, out, lane);
}

// The split of the tiles for the next step, by one workgroup of T threads:
// A holds the tiles that queued faces in this step and their eight
// neighbours, at most cap_a of them (in no particular order); B the rest.
// plan: [0] tiles in A, [2, 2 + ntiles) the A list, [2 + ntiles, 2 + 2 ntiles)
// a flag per tile, 1 when the tile is in A.  The counts are read into LDS (one
// byte per tile, up to DEFER_PLAN_MAX_TILES) and cleared for the next step.
constexpr int DEFER_PLAN_MAX_TILES = 32768;

template <int T>
__device__ void plan_tiles(const Grid2D& G, unsigned long long* tcount, int* plan, int cap_a)
{
    __shared__ unsigned char had[DEFER_PLAN_MAX_TILES];
    __shared__ int n_a;
    const int t = threadIdx.x;
    const int nsx = G.nsx, nsy = G.nsy, ntiles = nsx * nsy;
    for (int k = t; k < ntiles; k += T) had[k] = tcount[k] != 0ull;
    if (t == 0) n_a = 0;
    __syncthreads();
    for (int k = t; k < ntiles; k += T) tcount[k] = 0ull;
    for (int k = t; k < ntiles; k += T) {
        const int sx = k % nsx, sy = k / nsx;
        bool act = false;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int x = sx + dx, y = sy + dy;
                if (x >= 0 && x < nsx && y >= 0 && y < nsy && had[y * nsx + x]) act = true;
            }
        int in_a = 0;
        if (act) {
            const int pos = atomicAdd(&n_a, 1);
            if (pos < cap_a) {
                plan[2 + pos] = k;
                in_a = 1;
            }
        }
        plan[2 + ntiles + k] = in_a;
    }
    __syncthreads();
    if (t == 0) plan[0] = min(n_a, cap_a);
}

// The first step of a run description: the queue counters of both views
// for its slot, the per-tile counts, and a plan with every tile in B.
__global__ void k_defer_reset(unsigned long long* count_all, int slot, unsigned long long* tcount, int* plan,
                              int tiles)
{
    const int t = threadIdx.x;
    if (t < 6) count_all[9 * (t / 3) + 3 * slot + t % 3] = 0ull;
    if (t == 0) plan[0] = 0;
    for (int k = t; k < tiles; k += blockDim.x) {
        tcount[k] = 0ull;
        plan[2 + tiles + k] = 0;
    }
}

// The flux of queued face i from its sample, stored at the face; a failure
// is reported with the cell the strip pass would name.
__device__ __forceinline__ void finish_face(const Grid2D& G, const Deferred& D, long long i, const Sample& S,
                                            Ctrl* out)
{
    double o[4];
    const int e = flux_from_sample<4>(G.gamma, S, o);
    const long long id = D.qid[i];
    const long long k = id >> 1;
    if (id & 1) {
        // y face: (mass, my, mx, E) back to (mass, mx, my, E)
        D.fy[0][k] = o[0];
        D.fy[1][k] = o[2];
        D.fy[2][k] = o[1];
        D.fy[3][k] = o[3];
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) D.fx[q][k] = o[q];
    }
    if (e != ERR_NONE) {
        const long long row = k / G.pitch, c = k % G.pitch;
        report_error(out, e, row * G.tx + c - ((id & 1) ? 0 : 1));
    }
}

// Newton iterations a queued face may take in k_solve_faces.  Most faces
// converge within a few; the reference's iteration can also run its 100
// iterations and bisect (about 1 face in 100 of a blast), and one such face
// would hold its whole wavefront.  Those faces are set aside with their
// iteration state and finished together by k_solve_slow.
constexpr int NEWTON_CAP = 8;

// One queued face per lane: the exact solve (riemann_solver_exact.c) of
// exact_sample_capped, with at most cap Newton iterations before a face is
// set aside (cap 100: none is).  The queue length was written by the strip
// pass, which ran before this launch on the same stream.
__global__ void __launch_bounds__(LANES) k_solve_faces(Grid2D G, Deferred D0, Ctrl* ctrl, int slot, int cap)
{
    const Deferred D = defer_slot(D0, slot);
    const long long n = (long long)min(D.count[0], (unsigned long long)D.face_cap);
    Ctrl* out = ctrl + (slot + 1) % 3;
    const int lane = threadIdx.x;
    for (long long i = (long long)blockIdx.x * LANES + lane; i < n; i += (long long)gridDim.x * LANES) {
        double l[4], r[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            l[q] = D.qw[q][i];
            r[q] = D.qw[4 + q][i];
        }
        NewtonState st;
        Sample S = exact_sample_capped(G.gamma, G.tol, l[0], l[1], l[2], l[3], r[0], r[1], r[2], r[3], 0.0,
                                       cap, &st);
        const bool paused = S.rc == ERR_PAUSED;
        const long long j = wave_append(&D.count[2], paused, lane);
        if (paused && j < D.slow_cap) {
            D.sidx[j] = i;
            D.sst[0][j] = st.p0;
            D.sst[1][j] = st.p_hi;
            D.sst[2][j] = st.f_hi;
            D.sst[3][j] = st.p_lo;
            D.sst[4][j] = st.f_lo;
            D.sit[j] = st.it;
            D.sbr[j] = st.bracketed;
            continue;
        }
        if (paused)  // no room to set it aside: finish here
            S = exact_sample_resume(G.gamma, G.tol, l[0], l[1], l[2], l[3], r[0], r[1], r[2], r[3], 0.0, st);
        finish_face(G, D, i, S, out);
    }
}

// newton_finish (hfv_math.hpp; exact.c:664-760) by one wavefront, with the
// same operations on the same values and so the same result:
// - the bracket scan evaluates f at p_lo + j dp, j = 0..100, 64 points at a
//   time (lane l: j = 63 r + l in round r, so that consecutive rounds share
//   an end point), and takes the first j whose interval [j, j + 1] changes
//   sign, the interval where the serial scan stops;
// - the bisection evaluates, per round, the 63 midpoints of the six levels
//   below the current interval (lane l: node l + 1 of the tree in heap
//   order, its interval formed by the serial loop's own halvings along the
//   node's path), then takes six serial steps that read them.
// Every lane must call this; all return the same value.
__device__ PStar newton_finish_wave(const PressureFn& F, double tol, const NewtonState& s, int lane)
{
    double p_lo = s.p_lo, f_lo = s.f_lo, p_hi = s.p_hi, f_hi = s.f_hi;
    if (!s.bracketed) {
        p_lo = 0.0;
        f_lo = pressure_fn(F, p_lo);
        if (f_lo * f_hi >= 0.0) {
            p_hi *= 10.0;
            const double dp = (p_hi - p_lo) / 100;
            for (int base = 0; base < 100; base += LANES - 1) {
                const int j = base + lane;
                const double fj = (j <= 100) ? pressure_fn(F, p_lo + j * dp) : 0.0;
                const double fn = __shfl_down(fj, 1, LANES);
                const bool cross = (lane < LANES - 1) && (j < 100) && (fj * fn < 0.0);
                const unsigned long long m = __ballot(cross);
                if (m) {
                    const int l = __ffsll((long long)m) - 1;
                    const int k = base + l;
                    f_lo = __shfl(fj, l, LANES);
                    f_hi = __shfl(fj, l + 1, LANES);
                    p_hi = p_lo + (k + 1) * dp;
                    p_lo = p_lo + k * dp;
                    break;
                }
            }
            if (f_lo * f_hi >= 0.0) return PStar{0.0, ERR_EXACT_BRACKET};
        }
    }
    int count = 0;
    const int node = lane + 1;
    const int depth = 31 - __clz(node);
    for (;;) {
        double nlo = p_lo, nhi = p_hi;
        for (int d = depth - 1; d >= 0; d--) {
            const double m = 0.5 * (nlo + nhi);
            if ((node >> d) & 1) nlo = m;
            else nhi = m;
        }
        const double fm_l = (lane < LANES - 1) ? pressure_fn(F, 0.5 * (nlo + nhi)) : 0.0;
        int cur = 1;
        for (int lev = 0; lev < 6; lev++) {
            const double pm = 0.5 * (p_lo + p_hi);
            if (p_hi - p_lo < tol) return PStar{pm, ERR_NONE};
            const double fm = __shfl(fm_l, cur - 1, LANES);
            if (fm * f_lo < 0.0) {
                p_hi = pm;
                f_hi = fm;
                cur = 2 * cur;
            } else {
                p_lo = pm;
                f_lo = fm;
                cur = 2 * cur + 1;
            }
            count++;
            if (count > 5000) return PStar{0.0, ERR_EXACT_BISECT};
        }
    }
}

// One side's wave function f_K (df false) or its derivative f'_K (df true)
// at ps, by the expression pressure_fn_df evaluates for it (shock_f_df,
// rarefaction_f_df; exact.c:512-565), computed alone: y is the value's
// rarefaction exponent (F.zf for f_K, F.zd for f'_K), and pow_pos takes the
// logarithm that rarefaction_f_df shares between its two powers.  Both
// branches' values are computed and the side's branch selects one, so that
// the shock value's chain runs beside the power's.
__device__ __forceinline__ double wave_fn_part(const PressureFn& F, const WaveSide& S, double y, bool df, double ps)
{
    const double x = ps / S.p;
    const DD l = pow_log(x, F.lt);
    const double s = sqrt(S.A / (ps + S.B));
    const double d = ps - S.p;
    double vs = df ? (1.0 - 0.5 * d / (S.B - ps)) * s : d * s;
    // an empty statement that takes vs: the compiler keeps the shock
    // value's computation here, ahead of the power's special-case branch,
    // where it can interleave with the power, instead of sinking it below
    asm volatile("" : "+v"(vs));
    const double r = pow_with_log(x, y, l, F.et);
    // r / p a, or 2 a (r - 1) / (gamma - 1): one division either way
    const double num = df ? r : S.a2 * (r - 1.0);
    const double vr = num / (df ? S.pa : F.gm1);
    return (ps > S.p) ? vs : vr;
}

// v of lane K of the calling lane's quad (lanes 4q .. 4q + 3), by a DPP
// quad_perm broadcast; the four lanes of the quad must be active.
template <int K>
__device__ __forceinline__ double quad_lane(double v)
{
    constexpr int CTRL = K * 0x55;  // quad_perm [K, K, K, K]
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// newton_resume (hfv_math.hpp) by a quad of lanes: each iteration's f_0,
// f'_0, f_1 and f'_1 (the sides in PressureFn's order) are computed by lanes
// 0, 1, 2 and 3 of the quad (lane & 3) and summed in pressure_fn_df's order,
// so every lane of the quad takes newton_run's iterates, and the latency of
// an iteration is one side's value instead of all four in turn.  Returns 2
// with *P when the loop converged, 1 when it bracketed a root it had not
// bracketed when it started (*P: newton_finish's result), 0 when
// newton_finish of the state it started from applies.  The four lanes of a
// quad call this with the same face.
__device__ int newton_resume_quad(const PressureFn& F, double tol, NewtonState st, PStar* P, int lane)
{
    const int role = lane & 3;
    const bool df = (role & 1) != 0;
    const WaveSide S = (role & 2) ? F.s1 : F.s0;
    const double y = df ? F.zd : F.zf;
    const int br0 = st.bracketed;
    // v: this lane's value at the iterate st.p0, evaluated an iteration
    // ahead (at the end of the previous one), so that each iterate's
    // convergence test runs beside the evaluation at the next; iteration 0
    // takes the stored f and f' at the guess
    double v = (st.it > 0 && st.it < 100) ? wave_fn_part(F, S, y, df, st.p0) : 0.0;
    for (; st.it < 100; st.it++) {
        double f, d;
        if (st.it == 0) {
            f = st.f_hi;
            d = st.df0;
        } else {
            const double v0 = quad_lane<0>(v), v1 = quad_lane<1>(v);
            const double v2 = quad_lane<2>(v), v3 = quad_lane<3>(v);
            f = v0 + v2 + F.du;
            d = v1 + v3;
        }
        if (!st.bracketed && f * st.f_hi < 0.0) {
            st.bracketed = 1;
            st.f_lo = f;
            st.p_lo = st.p0;
        }
        const double p1 = st.p0 - f / d;
        if (p1 < 0.0) break;
        const bool conv = 2.0 * fabs(p1 - st.p0) / (p1 + st.p0) < tol;
        v = wave_fn_part(F, S, y, df, p1);  // the next iteration's (unused if p1 is the last)
        if (conv) {
            *P = PStar{p1, ERR_NONE};
            return 2;
        }
        st.p0 = p1;
    }
    if (st.bracketed == br0) return 0;
    *P = newton_finish(F, tol, st);
    return 1;
}

// k_solve_slow's outcome of a face whose result is newton_finish of its
// set-aside state (k_solve_spec's)
constexpr int SLOW_TAKE_SPEC = -100;

// The bracket scan and bisection of the set-aside faces whose Newton loop
// k_solve_slow ended without converging and with the bracket unchanged (sit
// SLOW_TAKE_SPEC), one face per wavefront: newton_finish of the state the
// face was set aside with, to Deferred::spp and sprc.  The guess, f there and
// the bracket (sst[1..4], sbr) are k_solve_faces's; k_solve_slow, which ran
// before this launch, wrote only sst[0] and sit.
__global__ void __launch_bounds__(LANES) k_solve_spec(Grid2D G, Deferred D0, int slot)
{
    const Deferred D = defer_slot(D0, slot);
    const long long n = (long long)min(D.count[2], (unsigned long long)D.slow_cap);
    if ((long long)blockIdx.x >= n) return;
    // these waves are a latency-bound chain that runs beside the strip pass:
    // their instructions win the SIMD's issue arbitration when ready
    __builtin_amdgcn_s_setprio(3);
    const int lane = threadIdx.x;
    __shared__ double pow_lds[POW_TABS_LEN];
    stage_pow_tables(pow_lds, lane);
    for (long long j = blockIdx.x; j < n; j += gridDim.x) {
        if (D.sit[j] != SLOW_TAKE_SPEC) continue;
        const long long i = D.sidx[j];
        const double rl = D.qw[0][i], ul = D.qw[1][i], pl = D.qw[3][i];
        const double rr = D.qw[4][i], ur = D.qw[5][i], pr = D.qw[7][i];
        const double al = (rl > 0.0) ? sound_speed(G.gamma, rl, pl) : 0.0;
        const double ar = (rr > 0.0) ? sound_speed(G.gamma, rr, pr) : 0.0;
        const Side L{rl, ul, pl, al}, R{rr, ur, pr, ar};
        const PressureFn F = pressure_fn_setup(G.gamma, L, R, pow_lds, pow_lds + POW_LOG_TAB_LEN);
        // newton_finish reads the guess, f there and the bracket, as the
        // face was set aside with them
        NewtonState st;
        st.p0 = 0.0;
        st.p_hi = D.sst[1][j];
        st.f_hi = D.sst[2][j];
        st.df0 = 0.0;
        st.p_lo = D.sst[3][j];
        st.f_lo = D.sst[4][j];
        st.it = 0;
        st.bracketed = D.sbr[j];
        const PStar P = newton_finish_wave(F, G.tol, st, lane);
        if (lane == 0) {
            D.spp[j] = P.p;
            D.sprc[j] = P.rc;
        }
    }
}

// Set-aside faces per wavefront of k_solve_slow: one per quad of lanes
constexpr int SLOW_FACES_PER_WAVE = LANES / 4;

// The set-aside faces' Newton loops, resumed from their saved state, one
// face per quad of lanes (newton_resume_quad): the outcome goes to sst[0]
// (p*) and sit (the failure kind, or SLOW_TAKE_SPEC).  A face that brackets
// a root in this loop is finished here, serially.  A set-aside face has
// taken at least one iteration (k_solve_faces' cap is at least 1).
__global__ void __launch_bounds__(LANES) k_solve_slow(Grid2D G, Deferred D0, int slot)
{
    const Deferred D = defer_slot(D0, slot);
    const long long n = (long long)min(D.count[2], (unsigned long long)D.slow_cap);
    if ((long long)blockIdx.x * SLOW_FACES_PER_WAVE >= n) return;
    // these waves are a latency-bound chain that runs beside the strip pass:
    // their instructions win the SIMD's issue arbitration when ready
    __builtin_amdgcn_s_setprio(3);
    const int lane = threadIdx.x;
    __shared__ double pow_lds[POW_TABS_LEN];
    stage_pow_tables(pow_lds, lane);
    const long long stride = (long long)gridDim.x * SLOW_FACES_PER_WAVE;
    for (long long j = (long long)blockIdx.x * SLOW_FACES_PER_WAVE + lane / 4; j < n; j += stride) {
        const long long i = D.sidx[j];
        const double rl = D.qw[0][i], ul = D.qw[1][i], pl = D.qw[3][i];
        const double rr = D.qw[4][i], ur = D.qw[5][i], pr = D.qw[7][i];
        const double al = (rl > 0.0) ? sound_speed(G.gamma, rl, pl) : 0.0;
        const double ar = (rr > 0.0) ? sound_speed(G.gamma, rr, pr) : 0.0;
        const Side L{rl, ul, pl, al}, R{rr, ur, pr, ar};
        const PressureFn F = pressure_fn_setup(G.gamma, L, R, pow_lds, pow_lds + POW_LOG_TAB_LEN);
        NewtonState st;
        st.p0 = D.sst[0][j];
        st.p_hi = D.sst[1][j];
        st.f_hi = D.sst[2][j];
        st.df0 = 0.0;  // used by iteration 0 only
        st.p_lo = D.sst[3][j];
        st.f_lo = D.sst[4][j];
        st.it = D.sit[j];
        st.bracketed = D.sbr[j];
        PStar P;
        const int took = newton_resume_quad(F, G.tol, st, &P, lane);
        if ((lane & 3) == 0) {
            if (took) {
                D.sst[0][j] = P.p;
                D.sit[j] = P.rc;
            } else {
                D.sit[j] = SLOW_TAKE_SPEC;
            }
        }
    }
}

// The set-aside faces' samples and fluxes, from k_solve_slow's outcome or
// k_solve_spec's finish (both complete before this launch), one face per
// lane.
__global__ void __launch_bounds__(LANES) k_solve_join(Grid2D G, Deferred D0, Ctrl* ctrl, int slot)
{
    const Deferred D = defer_slot(D0, slot);
    const long long n = (long long)min(D.count[2], (unsigned long long)D.slow_cap);
    Ctrl* out = ctrl + (slot + 1) % 3;
    for (long long j = (long long)blockIdx.x * LANES + threadIdx.x; j < n; j += (long long)gridDim.x * LANES) {
        const long long i = D.sidx[j];
        double l[4], r[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            l[q] = D.qw[q][i];
            r[q] = D.qw[4 + q][i];
        }
        const int code = D.sit[j];
        const PStar P = (code == SLOW_TAKE_SPEC) ? PStar{D.spp[j], D.sprc[j]} : PStar{D.sst[0][j], code};
        const double al = (l[0] > 0.0) ? sound_speed(G.gamma, l[0], l[3]) : 0.0;
        const double ar = (r[0] > 0.0) ? sound_speed(G.gamma, r[0], r[3]) : 0.0;
        const Side L{l[0], l[1], l[3], al}, R{r[0], r[1], r[3], ar};
        finish_face(G, D, i, exact_sample_star(G.gamma, L, l[2], R, r[2], P, 0.0), out);
    }
}

// The deferred cells' updates, in the strip pass's operation order
// (integrator_muscl_hancock.c:806-814, 1092-1100; gravity source_term.c:
// 213-240), and their share of the CFL maxima.
constexpr int UPD_THREADS = 256;

// views: the queue views whose deferred cells this step updates (1: the
// plain form's, 2: both of the split form).  planner = 1: workgroup 0 plans
// the next step's split (plan_tiles) and the others update the cells.
__global__ void __launch_bounds__(UPD_THREADS) k_update_deferred(Grid2D G, RunParams rp, Fields4 Ua, Fields4 Ub,
                                                                 Deferred DA, Deferred DB, int views, Ctrl* ctrl,
                                                                 int slot, int* tile_plan, int cap_a, int planner)
{
    if (planner && blockIdx.x == 0) {
        plan_tiles<UPD_THREADS>(G, DA.tcount, tile_plan, cap_a);
        return;
    }
    const long long blk = (long long)blockIdx.x - planner, nblk = (long long)gridDim.x - planner;
    int dt_zero;
    const StepPlan plan = plan_step(ctrl + slot, rp, &dt_zero);
    if (!plan.run) return;
    const Fields4 U = plan.cur ? Ub : Ua;
    const Fields4 V = plan.cur ? Ua : Ub;
    const double dt = plan.dt, gamma = G.gamma;
    double smx = -INFINITY, smy = -INFINITY;
    for (int v = 0; v < views; v++) {
    const Deferred D = defer_slot(v == 0 ? DA : DB, slot);
    const long long n = (long long)min(D.count[1], (unsigned long long)D.cell_cap);
    for (long long i = blk * UPD_THREADS + threadIdx.x; i < n; i += nblk * UPD_THREADS) {
        const long long e = D.cells[i];
        const int kind = (int)(e & 7);
        const long long k = e >> 3;
        if (kind == DEF_INTERIOR) {
            double u[4];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                double x = U.f[f][k];
                x += D.fx[f][k] * dt * G.area_x;
                x -= D.fx[f][k + 1] * dt * G.area_x;
                x += D.fy[f][k - G.pitch] * dt * G.area_y;
                x -= D.fy[f][k] * dt * G.area_y;
                u[f] = x;
            }
            if (G.gravity != 0.0) gravity_source(dt, G.gravity, u[0], &u[2], &u[3]);
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = u[f];
            double w[4];
            c2p_2d(gamma, G.volume, u[0], u[1], u[2], u[3], w);
            const double a = sound_speed(gamma, w[0], w[3]);
            const double sx = a + fabs(w[1]);
            const double sy = a + fabs(w[2]);
            if (sx > smx) smx = sx;
            if (sy > smy) smy = sy;
        } else if (kind == DEF_GHOST_XLO) {
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] - D.fx[f][k + 1] * dt * G.area_x;
        } else if (kind == DEF_GHOST_XHI) {
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] + D.fx[f][k] * dt * G.area_x;
        } else if (kind == DEF_GHOST_YLO) {
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] - D.fy[f][k] * dt * G.area_y;
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] + D.fy[f][k - G.pitch] * dt * G.area_y;
        }
    }
    }
    block_smax_to<UPD_THREADS>(ctrl + (slot + 1) % 3, smx, smy);
}

// CFL maxima of a primitive state held in memory (the first step's dt):
// workgroups stride over the interior rows, reduce in registers and LDS, and
// add one atomic per key (the comparisons of utils.c:75-92 skip NaN).
constexpr int SMAX_THREADS = 256;

__global__ void __launch_bounds__(SMAX_THREADS) k_smax_from_w(Grid2D G, Fields4 W, Ctrl* ctrl, int slot)
{
    double sx = -INFINITY, sy = -INFINITY;
    for (int j = blockIdx.x; j < G.ny; j += gridDim.x) {
        const long long row = (long long)(j + G.g) * G.pitch + G.g;
        for (int i = threadIdx.x; i < G.nx; i += SMAX_THREADS) {
            const long long k = row + i;
            const double a = sound_speed(G.gamma, W.f[0][k], W.f[3][k]);
            const double vx = a + fabs(W.f[1][k]);
            const double vy = a + fabs(W.f[2][k]);
            if (vx > sx) sx = vx;
            if (vy > sy) sy = vy;
        }
    }
    block_smax_to<SMAX_THREADS>(ctrl + slot, sx, sy);
}

// Primitive state of every cell after a step: c2p everywhere, then the ghost
// fill (x-ghosts on interior rows, y-ghosts on interior columns; the corner
// blocks keep c2p of their own conserved state), i.e. system.c:763-802
// followed by boundary.c:513-777.
// With a controller slot: nothing when the run's first step failed in its
// flux, whose primitive state is the caller's own.
__global__ void k_prim_from_u(Grid2D G, Fields4 U, Fields4 W, int fill, const Ctrl* ctrl, int slot)
{
    if (ctrl && ctrl[slot].err_kind != ERR_NONE && ctrl[slot].err_kind != ERR_DT_ZERO && ctrl[slot].num_steps == 1)
        return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (c >= G.tx) return;
    const bool gx = fill && ((c < G.g) || (c >= G.g + G.nx));
    const bool gy = fill && ((q < G.g) || (q >= G.g + G.ny));
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

// fill = 0: c2p of every cell's own conserved state (system.c:763-802) with
// no boundary fill
void prim_from_u(const Grid2D& G, const Fields4& U, const Fields4& W, hipStream_t s, int fill,
                 const Ctrl* ctrl, int slot)
{
    const dim3 grid((G.tx + 63) / 64, G.ty), block(64);
    hipLaunchKernelGGL(k_prim_from_u, grid, block, 0, s, G, U, W, fill, ctrl, slot);
}

// c2p of every cell (system.c:763-802) from the conserved buffer that holds
// the current state, as named by the controller slot the next launch reads
__global__ void k_prim_from_slot(Grid2D G, Fields4 Ua, Fields4 Ub, Fields4 W, const Ctrl* ctrl, int slot)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (c >= G.tx) return;
    const Fields4 U = ctrl[slot].cur ? Ub : Ua;
    const long long k = (long long)q * G.pitch + c;
    double w[4];
    c2p_2d(G.gamma, G.volume, U.f[0][k], U.f[1][k], U.f[2][k], U.f[3][k], w);
    W.f[0][k] = w[0];
    W.f[1][k] = w[1];
    W.f[2][k] = w[2];
    W.f[3][k] = w[3];
}

// Built-in boundary fills of the primitive buffer, from its own interior
// cells (boundary.c:513-777): the x sides on the interior rows, the y sides
// on the interior columns.  Each side reads interior cells only and writes
// its own ghost cells, so the sides in `mask` (bit s = side s, order x_min,
// x_max, y_min, y_max) run in one launch.  A periodic pair is filled from
// its max side, both ghost strips, as in the reference; the min side's
// periodic case does nothing.  Needs num_cells >= ghost cells per direction.
__global__ void k_fill_sides(Grid2D G, Fields4 W, int mask)
{
    const int side = blockIdx.y;
    if (!((mask >> side) & 1)) return;
    const int bc = G.bc[side];
    const bool x = side < 2;
    const int span = x ? G.ny : G.nx;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= G.g * span) return;
    const int layer = e / span, along = G.g + e % span;
    const int n = x ? G.nx : G.ny;
    const int flip = x ? 1 : 2;  // the normal velocity
    // (ghost index, source index) pairs along the normal; at most two for a
    // periodic pair
    int dst[2], src[2], cnt = 0;
    if (bc == BC_REFLECTIVE || bc == BC_TRANSMISSIVE) {
        if (side == 0 || side == 2) {
            dst[0] = G.g - 1 - layer;
            src[0] = G.g + layer;
        } else {
            dst[0] = G.g + n + layer;
            src[0] = G.g + n - 1 - layer;
        }
        cnt = 1;
    } else if (bc == BC_PERIODIC && (side == 1 || side == 3)) {
        dst[0] = G.g + n + layer;
        src[0] = G.g + layer;
        dst[1] = G.g - 1 - layer;
        src[1] = G.g + n - 1 - layer;
        cnt = 2;
    }
    for (int m = 0; m < cnt; m++) {
        const long long kd = x ? (long long)along * G.pitch + dst[m] : (long long)dst[m] * G.pitch + along;
        const long long ks = x ? (long long)along * G.pitch + src[m] : (long long)src[m] * G.pitch + along;
        for (int f = 0; f < 4; f++) {
            const double v = W.f[f][ks];
            W.f[f][kd] = (bc == BC_REFLECTIVE && f == flip) ? -v : v;
        }
    }
}

__global__ void k_clear_keys(Ctrl* ctrl, int slot)
{
    if (threadIdx.x == 0) {
        ctrl[slot].key[0] = 0ull;
        ctrl[slot].key[1] = 0ull;
    }
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

// Copy of the 4 fields of every cell (ghosts included) from src to dst; with
// a controller slot, only when the slot names buffer `when_cur` as holding
// the run's result.
__global__ void k_copy_fields(Grid2D G, Fields4 dst, Fields4 src, const Ctrl* ctrl, int slot, int when_cur)
{
    if (ctrl) {
        // after a failed flux the result is the state at the start of the
        // failing step, in the other buffer (as hfv_mh2d_download returns)
        const Ctrl& c = ctrl[slot];
        const int result = (c.err_kind != ERR_NONE && c.err_kind != ERR_DT_ZERO) ? c.cur ^ 1 : c.cur;
        if (result != when_cur) return;
    }
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G.tx) return;
    const long long k = (long long)blockIdx.y * G.pitch + c;
#pragma unroll
    for (int f = 0; f < 4; f++) dst.f[f][k] = src.f[f][k];
}

// Probe of the pressure iteration's finish (hfv_probe_exact_finish): the
// Newton loop from the guess for cap iterations, then, for a face it has not
// converged, newton_finish (WAVE 0: one face per lane) or newton_finish_wave
// (WAVE 1: one face per wavefront) from the state it left.  rc: the failure
// kind, or -1 for a face whose loop converged (p: the converged iterate), -3
// for a vacuum face (no iteration).
template <bool WAVE>
__global__ void __launch_bounds__(LANES) k_probe_finish(double gamma, double tol, int cap, long long n,
                                                        const double* wl, const double* wr, double* p, int* rc)
{
    const int lane = threadIdx.x;
    const long long i = WAVE ? (long long)blockIdx.x : (long long)blockIdx.x * LANES + lane;
    if (i >= n) return;
    const double rl = wl[4 * i], ul = wl[4 * i + 1], pl = wl[4 * i + 3];
    const double rr = wr[4 * i], ur = wr[4 * i + 1], pr = wr[4 * i + 3];
    const double al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0;
    double out_p = 0.0;
    int out_rc = -3;
    if (!(rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul))) {
        const Side L{rl, ul, pl, al}, R{rr, ur, pr, ar};
        const PressureFn F = pressure_fn_setup(gamma, L, R);
        NewtonState st = newton_begin(gamma, L, R, tol, F);
        double ps;
        if (newton_run(F, tol, st, cap, &ps) > 0) {
            out_p = ps;
            out_rc = -1;
        } else {
            PStar P;
            if constexpr (WAVE) P = newton_finish_wave(F, tol, st, lane);
            else P = newton_finish(F, tol, st);
            out_p = P.p;
            out_rc = P.rc;
        }
    }
    if (!WAVE || lane == 0) {
        p[i] = out_p;
        rc[i] = out_rc;
    }
}

// hfv_probe_exact_finish impl 2: Newton for cap iterations by one lane,
// the rest of the loop as k_solve_slow takes it (newton_resume_quad, a quad
// of lanes per face), then newton_finish where its outcome says so.
__global__ void __launch_bounds__(LANES) k_probe_resume_quad(double gamma, double tol, int cap, long long n,
                                                             const double* wl, const double* wr, double* p, int* rc)
{
    const int lane = threadIdx.x;
    const long long i = (long long)blockIdx.x * SLOW_FACES_PER_WAVE + lane / 4;
    if (i >= n) return;
    const double rl = wl[4 * i], ul = wl[4 * i + 1], pl = wl[4 * i + 3];
    const double rr = wr[4 * i], ur = wr[4 * i + 1], pr = wr[4 * i + 3];
    const double al = (rl > 0.0) ? sound_speed(gamma, rl, pl) : 0.0;
    const double ar = (rr > 0.0) ? sound_speed(gamma, rr, pr) : 0.0;
    double out_p = 0.0;
    int out_rc = -3;
    if (!(rl <= 0.0 || rr <= 0.0 || ((al + ar) * 2.0 / (gamma + 1.0)) <= (ur - ul))) {
        const Side L{rl, ul, pl, al}, R{rr, ur, pr, ar};
        const PressureFn F = pressure_fn_setup(gamma, L, R);
        NewtonState st = newton_begin(gamma, L, R, tol, F);
        double ps;
        const int r = newton_run(F, tol, st, cap, &ps);
        PStar P{0.0, ERR_NONE};
        if (r > 0) {
            out_p = ps;
            out_rc = -1;
        } else if (r < 0) {
            P = newton_finish(F, tol, st);
            out_p = P.p;
            out_rc = P.rc;
        } else {
            const int took = newton_resume_quad(F, tol, st, &P, lane);
            if (took == 0) P = newton_finish(F, tol, st);
            out_p = P.p;
            out_rc = (took == 2) ? -1 : P.rc;
        }
    }
    if ((lane & 3) == 0) {
        p[i] = out_p;
        rc[i] = out_rc;
    }
}

// Working memory of hfv_mh2d_run_noalloc on each GPU: the three controller
// slots in its first 256 bytes, then the second conserved buffer.  Host code
// takes its address per device with hipGetSymbolAddress (scratch_base) and
// never uses the name itself.
constexpr long long NOALLOC_DOUBLES = HFV_NOALLOC_SCRATCH_BYTES / 8;
constexpr long long NOALLOC_HEAD = 256 / 8;  // doubles before the fields
static_assert(3 * sizeof(Ctrl) <= 256, "controller slots fit the scratch head");
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

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
    hipStream_t own;       // used when the caller passes no stream (graphs cannot be captured on the null stream)
    int w_pending;         // the next step reads the primitive state set by hfv_mh2d_set_primitive
    hfv_device_bc_2d_func dev_fn[4];  // device hooks of the custom sides
    void* dev_user[4];
    long long w_filled;    // launch count at which the primitive buffer was last filled (device hooks)
    double* stage;         // one contiguous field for host copies (allocated at first use)
    Deferred D;            // exact solver: queued faces and deferred cells (view A; the plain form's)
    Deferred DB;           // exact solver: view B of the split form
    unsigned long long* count_all;  // exact solver: both views' queue counters, [view][slot][3]
    int* plan;             // exact solver: the split of the tiles (plan_tiles)
    hipStream_t aux;       // exact solver: the second stream of the split form
    hipEvent_t ev_fork, ev_join;
    long long defer_clear; // exact solver: the launch whose queue counters are known to be clear
    double* defer_slab;
};

namespace {

hipStream_t pick(hfv_mh2d* h, void* stream) { return stream ? (hipStream_t)stream : h->own; }

// Host copies of one field (reference layout, tx x ty, x fastest) from and
// to a pitched device field.  A pitched copy to pageable host memory goes
// row by row, so unless the pitch equals the width the field passes through
// a contiguous device buffer: a device-side 2D copy plus one linear copy.
bool staged(hfv_mh2d* h)
{
    if (h->G.pitch == h->G.tx) return false;
    if (!h->stage && hipMalloc(&h->stage, (size_t)h->G.tx * h->G.ty * sizeof(double)) != hipSuccess) {
        h->stage = nullptr;
        (void)hipGetLastError();
    }
    return h->stage != nullptr;
}

hipError_t field_to_host(hfv_mh2d* h, double* host, const double* dev, hipStream_t s)
{
    const Grid2D& G = h->G;
    const size_t hp = (size_t)G.tx * sizeof(double), dp = (size_t)G.pitch * sizeof(double);
    if (G.pitch == G.tx) return hipMemcpyAsync(host, dev, hp * G.ty, hipMemcpyDeviceToHost, s);
    if (!staged(h)) return hipMemcpy2DAsync(host, hp, dev, dp, hp, G.ty, hipMemcpyDeviceToHost, s);
    hipError_t e = hipMemcpy2DAsync(h->stage, hp, dev, dp, hp, G.ty, hipMemcpyDeviceToDevice, s);
    return e != hipSuccess ? e : hipMemcpyAsync(host, h->stage, hp * G.ty, hipMemcpyDeviceToHost, s);
}

hipError_t field_to_device(hfv_mh2d* h, double* dev, const double* host, hipStream_t s)
{
    const Grid2D& G = h->G;
    const size_t hp = (size_t)G.tx * sizeof(double), dp = (size_t)G.pitch * sizeof(double);
    if (G.pitch == G.tx) return hipMemcpyAsync(dev, host, hp * G.ty, hipMemcpyHostToDevice, s);
    if (!staged(h)) return hipMemcpy2DAsync(dev, dp, host, hp, hp, G.ty, hipMemcpyHostToDevice, s);
    hipError_t e = hipMemcpyAsync(h->stage, host, hp * G.ty, hipMemcpyHostToDevice, s);
    return e != hipSuccess ? e : hipMemcpy2DAsync(dev, dp, h->stage, hp, hp, G.ty, hipMemcpyDeviceToDevice, s);
}

template <int LIM, int SOLVER, bool BRANCHY>
void launch_step_t(hfv_mh2d* h, bool wmode, hipStream_t s)
{
    const int slot = (int)(h->launches % 3);
    const dim3 grid(h->G.nsx * h->G.nsy), block(LANES);
    if (wmode) {
        hipLaunchKernelGGL((k_mh2d_step<LIM, SOLVER, true, BRANCHY>), grid, block, 0, s, h->G, h->rp, h->U[0],
                           h->U[1], h->W, h->ctrl, slot);
    } else {
        hipLaunchKernelGGL((k_mh2d_step<LIM, SOLVER, false, BRANCHY>), grid, block, 0, s, h->G, h->rp, h->U[0],
                           h->U[1], h->W, h->ctrl, slot);
    }
    h->launches++;
}

// exact solver: strip pass, queued faces, deferred cells (and the first
// step's queue reset)
template <int LIM>
void launch_exact(hfv_mh2d* h, bool wmode, hipStream_t s)
{
    const int slot = (int)(h->launches % 3);
    const Grid2D& G = h->G;
    const int ntiles = G.nsx * G.nsy;
    const long long cells = (long long)G.tx * G.ty;
    const unsigned solve_blocks = (unsigned)std::min<long long>(4096, (2 * cells + LANES - 1) / LANES);
    const unsigned update_blocks = (unsigned)std::min<long long>(1024, (cells + UPD_THREADS - 1) / UPD_THREADS);
    const unsigned slow_blocks =
        (unsigned)std::min<long long>(4096, (h->D.slow_cap + SLOW_FACES_PER_WAVE - 1) / SLOW_FACES_PER_WAVE);
    const unsigned join_blocks = (unsigned)std::min<long long>(1024, (h->D.slow_cap + LANES - 1) / LANES);
    const unsigned spec_blocks = (unsigned)std::min<long long>(2048, h->D.slow_cap);
    const int cap_a = std::min(ntiles, DEFER_MAX_A);
    if (h->defer_clear != h->launches)
        hipLaunchKernelGGL(k_defer_reset, dim3(1), dim3(256), 0, s, h->count_all, slot, h->D.tcount, h->plan, ntiles);
    const dim3 block(LANES), ublock(UPD_THREADS);
    if (wmode || ntiles > DEFER_PLAN_MAX_TILES) {
        // plain form
        if (wmode)
            hipLaunchKernelGGL((k_mh2d_step_defer<LIM, true>), dim3(ntiles), block, 0, s, G, h->rp, h->U[0],
                               h->U[1], h->W, h->ctrl, slot, h->D, h->count_all);
        else
            hipLaunchKernelGGL((k_mh2d_step_defer<LIM, false>), dim3(ntiles), block, 0, s, G, h->rp, h->U[0],
                               h->U[1], h->W, h->ctrl, slot, h->D, h->count_all);
        hipLaunchKernelGGL(k_solve_faces, dim3(solve_blocks), block, 0, s, G, h->D, h->ctrl, slot, NEWTON_CAP);
        hipLaunchKernelGGL(k_solve_slow, dim3(slow_blocks), block, 0, s, G, h->D, slot);
        hipLaunchKernelGGL(k_solve_spec, dim3(spec_blocks), block, 0, s, G, h->D, slot);
        hipLaunchKernelGGL(k_solve_join, dim3(join_blocks), block, 0, s, G, h->D, h->ctrl, slot);
        const int planner = ntiles <= DEFER_PLAN_MAX_TILES;
        hipLaunchKernelGGL(k_update_deferred, dim3(update_blocks + planner), ublock, 0, s, G, h->rp, h->U[0],
                           h->U[1], h->D, h->DB, 1, h->ctrl, slot, h->plan, cap_a, planner);
    } else {
        // split form: pass A and its first face pass on s; its set-aside
        // faces' Newton loops, the finish of those that need it and the join
        // on aux, while pass B runs on s and solves its own faces (rarely
        // any: a tile with none in the previous step and none next to it) in
        // their first pass
        hipLaunchKernelGGL((k_mh2d_step_split<LIM, 0>), dim3(cap_a * DEFER_NSUB), block, 0, s, G, h->rp, h->U[0],
                           h->U[1], h->ctrl, slot, h->D, h->count_all, (const int*)h->plan);
        hipLaunchKernelGGL(k_solve_faces, dim3(solve_blocks), block, 0, s, G, h->D, h->ctrl, slot, NEWTON_CAP);
        (void)hipEventRecord(h->ev_fork, s);
        (void)hipStreamWaitEvent(h->aux, h->ev_fork, 0);
        hipLaunchKernelGGL(k_solve_slow, dim3(slow_blocks), block, 0, h->aux, G, h->D, slot);
        hipLaunchKernelGGL(k_solve_spec, dim3(spec_blocks), block, 0, h->aux, G, h->D, slot);
        hipLaunchKernelGGL(k_solve_join, dim3(join_blocks), block, 0, h->aux, G, h->D, h->ctrl, slot);
        (void)hipEventRecord(h->ev_join, h->aux);
        hipLaunchKernelGGL((k_mh2d_step_split<LIM, 1>), dim3(ntiles), block, 0, s, G, h->rp, h->U[0], h->U[1],
                           h->ctrl, slot, h->DB, h->count_all, (const int*)h->plan);
        hipLaunchKernelGGL(k_solve_faces, dim3(solve_blocks), block, 0, s, G, h->DB, h->ctrl, slot, 100);
        (void)hipStreamWaitEvent(s, h->ev_join, 0);
        hipLaunchKernelGGL(k_update_deferred, dim3(update_blocks + 1), ublock, 0, s, G, h->rp, h->U[0], h->U[1],
                           h->D, h->DB, 2, h->ctrl, slot, h->plan, cap_a, 1);
    }
    h->launches++;
    h->defer_clear = h->launches;  // the strip pass cleared the next slot's counters
}

template <int LIM>
void launch_hllc(hfv_mh2d* h, bool wmode, hipStream_t s)
{
    if (h->cfg.hllc_form == HFV_HLLC_BRANCHES) launch_step_t<LIM, HLLC, true>(h, wmode, s);
    else launch_step_t<LIM, HLLC, false>(h, wmode, s);
}

// The step kernel a configuration launches (primitive-state mode aside,
// which has the same register budget): for its occupancy.
template <int LIM>
const void* step_kernel_lim(const hfv_mh2d_config* c)
{
    if (c->riemann_solver == EXACT) return reinterpret_cast<const void*>(&k_mh2d_step_split<LIM, 1>);
    if (c->hllc_form == HFV_HLLC_BRANCHES) return reinterpret_cast<const void*>(&k_mh2d_step<LIM, HLLC, false, true>);
    return reinterpret_cast<const void*>(&k_mh2d_step<LIM, HLLC, false, false>);
}

const void* step_kernel(const hfv_mh2d_config* c)
{
    if (c->first_order) return step_kernel_lim<FIRST_ORDER>(c);
    if (c->slope_limiter == MINMOD) return step_kernel_lim<MINMOD>(c);
    if (c->slope_limiter == VAN_LEER) return step_kernel_lim<VAN_LEER>(c);
    return step_kernel_lim<MC>(c);
}

// Step wavefronts the current device holds at once: CUs x resident
// workgroups (one wavefront each) per CU; 0 when the runtime cannot say.
long long resident_step_waves(const hfv_mh2d_config* c)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, step_kernel(c), LANES, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return (long long)cus * per_cu;
}

void launch_step(hfv_mh2d* h, bool wmode, hipStream_t s)
{
    const int lim = h->cfg.slope_limiter, sol = h->cfg.riemann_solver;
    if (h->cfg.first_order) {
        if (sol == HLLC) launch_hllc<FIRST_ORDER>(h, wmode, s);
        else launch_exact<FIRST_ORDER>(h, wmode, s);
    } else if (sol == HLLC) {
        if (lim == MINMOD) launch_hllc<MINMOD>(h, wmode, s);
        else if (lim == VAN_LEER) launch_hllc<VAN_LEER>(h, wmode, s);
        else launch_hllc<MC>(h, wmode, s);
    } else {
        if (lim == MINMOD) launch_exact<MINMOD>(h, wmode, s);
        else if (lim == VAN_LEER) launch_exact<VAN_LEER>(h, wmode, s);
        else launch_exact<MC>(h, wmode, s);
    }
}

int device_bc_mode(const hfv_mh2d* h)
{
    for (int s = 0; s < 4; s++)
        if (h->dev_fn[s]) return 1;
    return 0;
}

int device_bc_complete(const hfv_mh2d* h)
{
    for (int s = 0; s < 4; s++)
        if (h->G.bc[s] == BC_CUSTOM && !h->dev_fn[s]) return 0;
    return 1;
}

void fill_sides(hfv_mh2d* h, int mask, hipStream_t s)
{
    if (!mask) return;
    const Grid2D& G = h->G;
    const int span = std::max(G.nx, G.ny);
    const dim3 grid((unsigned)((G.g * span + 255) / 256), 4), block(256);
    hipLaunchKernelGGL(k_fill_sides, grid, block, 0, s, G, h->W, mask);
}

// Device-hook mode: the primitive buffer for the next launch, i.e. the
// reference's c2p then set_boundary_condition (boundary.c:513-777, sides in
// its order) at the end of a step, then the CFL maxima of that state
// (utils.c:55-95) into the slot the next launch reads.  Once per launch
// count, so a snapshot's download and the next step share one fill.
void fill_primitive(hfv_mh2d* h, hipStream_t s)
{
    if (h->w_filled == h->launches) return;
    const Grid2D& G = h->G;
    const int slot = (int)(h->launches % 3);
    hipLaunchKernelGGL(k_prim_from_slot, dim3((G.tx + 63) / 64, G.ty), dim3(64), 0, s, G, h->U[0], h->U[1], h->W,
                       h->ctrl, slot);
    int mask = 0;
    for (int side = 0; side < 4; side++) {
        if (G.bc[side] != BC_CUSTOM) {
            mask |= 1 << side;
            continue;
        }
        fill_sides(h, mask, s);
        mask = 0;
        h->dev_fn[side](h->W.f[0], h->W.f[1], h->W.f[2], h->W.f[3], G.g, G.nx, G.ny, G.pitch, (void*)s,
                        h->dev_user[side]);
    }
    fill_sides(h, mask, s);
    hipLaunchKernelGGL(k_clear_keys, dim3(1), dim3(64), 0, s, h->ctrl, slot);
    const dim3 grid((unsigned)std::min(G.ny, 2048)), block(SMAX_THREADS);
    hipLaunchKernelGGL(k_smax_from_w, grid, block, 0, s, G, h->W, h->ctrl, slot);
    h->w_filled = h->launches;
}

// Validation shared by the 2D entry points.  Fewer interior cells than ghost
// cells in a direction is refused: the boundary mapping (map_index) and the
// side fills need num_cells >= num_ghost_cells_side.
int check_config(const hfv_mh2d_config* cfg)
{
    if (cfg->first_order != 0 && cfg->first_order != 1) return HFV_EINVAL;
    if (cfg->num_cells_x < 2 || cfg->num_cells_y < 2 || cfg->num_ghost_cells_side < (cfg->first_order ? 1 : 2))
        return HFV_EINVAL;
    if (cfg->num_cells_x < cfg->num_ghost_cells_side || cfg->num_cells_y < cfg->num_ghost_cells_side)
        return HFV_EINVAL;
    if (!cfg->first_order && (cfg->slope_limiter < 1 || cfg->slope_limiter > 3)) return HFV_EINVAL;
    if (cfg->riemann_solver != HLLC && cfg->riemann_solver != EXACT) return HFV_EINVAL;
    for (int s = 0; s < 4; s++)
        // BC_CUSTOM: the device leaves those ghosts as c2p of their own
        // state; the host applies the hook between steps (hfv_integrator.c)
        if (cfg->bc[s] < BC_NONE || cfg->bc[s] > BC_CUSTOM) return HFV_EINVAL;
    if (cfg->strip_rows < 0) return HFV_EINVAL;
    if (cfg->hllc_form != HFV_HLLC_SELECTS && cfg->hllc_form != HFV_HLLC_BRANCHES) return HFV_EINVAL;
    return HFV_OK;
}

// Grid description and time-control parameters of a configuration, for
// fields whose rows are `pitch` doubles apart.
void grid_from_config(const hfv_mh2d_config* cfg, long long pitch, Grid2D& G, RunParams& rp)
{
    G.nx = cfg->num_cells_x;
    G.ny = cfg->num_cells_y;
    G.g = cfg->num_ghost_cells_side;
    G.tx = G.nx + 2 * G.g;
    G.ty = G.ny + 2 * G.g;
    G.pitch = pitch;
    G.gamma = cfg->gamma;
    G.gravity = cfg->gravity;
    G.tol = cfg->tol;
    G.area_x = cfg->area_x;
    G.area_y = cfg->area_y;
    G.volume = cfg->volume;
    for (int s = 0; s < 4; s++) G.bc[s] = cfg->bc[s];
    G.nsx = (G.nx + STRIP_COLS - 1) / STRIP_COLS;
    G.strip_rows = cfg->strip_rows > 0                ? cfg->strip_rows
                   : cfg->riemann_solver == EXACT ? exact_strip_rows(G.nx, G.ny, resident_step_waves(cfg))
                                                  : balanced_strip_rows(G.nx, G.ny, resident_step_waves(cfg));
    G.nsy = (G.ny + G.strip_rows - 1) / G.strip_rows;
    G.full_tiles = cfg->full_tiles != 0;
    rp = RunParams();
    rp.cfl = cfg->cfl;
    rp.shrink_factor = cfg->cfl_initial_shrink_factor;
    rp.shrink_steps = cfg->cfl_initial_shrink_num_steps;
    rp.dim = 2;
    rp.tf = cfg->tf;
    rp.max_steps = cfg->max_steps;
    rp.fixed_dt = cfg->fixed_dt;
    rp.dx = cfg->dx;
    rp.dy = cfg->dy;
}

void fill_status(const Ctrl& c, const RunParams& rp, hfv_run_status* st)
{
    st->t = c.t;
    st->dt = c.dt;
    st->t_prev = c.t_prev;
    st->num_steps = c.num_steps;
    st->err_kind = c.err_kind;
    st->err_cell = c.err_cell;
    st->cur = c.cur;
    st->smax[0] = smax_from_key(c.key[0]);
    st->smax[1] = smax_from_key(c.key[1]);
    const bool finished =
        c.err_kind != ERR_NONE || !(c.t < rp.tf) || (rp.max_steps >= 0 && c.num_steps >= rp.max_steps);
    st->finished = finished ? 1 : 0;
    st->bisect_fails = 0;
}

// This device's address of `scratch`, looked up on the first call on the
// device (nullptr if the lookup fails).
double* scratch_base()
{
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static double* base[MAX_DEVICES] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!base[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        base[dev] = static_cast<double*>(p);
    }
    return base[dev];
}

}  // namespace

extern "C" {

int hfv_mh2d_create(const hfv_mh2d_config* cfg, hfv_mh2d** out)
{
    *out = nullptr;
    const int rc = check_config(cfg);
    if (rc != HFV_OK) return rc;
    hfv_mh2d* h = new (std::nothrow) hfv_mh2d();
    if (!h) return HFV_ENOMEM;
    h->cfg = *cfg;
    Grid2D& G = h->G;
    grid_from_config(cfg, ((long long)cfg->num_cells_x + 2 * cfg->num_ghost_cells_side + 31) / 32 * 32, G, h->rp);

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
    if (cfg->riemann_solver == EXACT) {
        // every face can be queued (each has one owning lane) and every cell
        // deferred (each has one updating lane)
        Deferred& D = h->D;
        D.cell_cap = (long long)G.tx * G.ty;
        D.face_cap = 2 * D.cell_cap;
        // a face that cannot be set aside is finished in the first pass
        D.slow_cap = std::max<long long>(D.face_cap / 8, 64);
        const size_t ntiles = (size_t)G.nsx * G.nsy;
        const size_t view = (size_t)D.face_cap * 9 + (size_t)D.cell_cap + (size_t)D.slow_cap * 9;
        const size_t dbytes = (field * 8 + 2 * view + 18 + ntiles + (2 + 2 * ntiles + 1) / 2 + 1) * sizeof(double);
        if (hipMalloc(&h->defer_slab, dbytes) != hipSuccess) {
            (void)hipFree(h->slab);
            delete h;
            return HFV_ENOMEM;
        }
        double* p = h->defer_slab;
        for (int f = 0; f < 4; f++) {
            D.fx[f] = p + field * f;
            D.fy[f] = p + field * (4 + f);
        }
        p += field * 8;
        // the queue arrays of one view
        const auto carve = [&](Deferred& V) {
            for (int q = 0; q < 8; q++) V.qw[q] = p + (size_t)D.face_cap * q;
            p += (size_t)D.face_cap * 8;
            V.qid = reinterpret_cast<long long*>(p);
            p += D.face_cap;
            V.cells = reinterpret_cast<long long*>(p);
            p += D.cell_cap;
            V.sidx = reinterpret_cast<long long*>(p);
            p += D.slow_cap;
            for (int q = 0; q < 5; q++) V.sst[q] = p + (size_t)D.slow_cap * q;
            p += (size_t)D.slow_cap * 5;
            V.sit = reinterpret_cast<int*>(p);
            V.sbr = V.sit + D.slow_cap;
            p += D.slow_cap;  // two int arrays in one double-sized stride
            V.spp = p;
            p += D.slow_cap;
            V.sprc = reinterpret_cast<int*>(p);
            p += D.slow_cap;  // one int array in a double-sized stride
        };
        Deferred& DB = h->DB;
        DB = D;
        carve(D);
        carve(DB);
        h->count_all = reinterpret_cast<unsigned long long*>(p);
        p += 18;
        D.count = h->count_all;
        DB.count = h->count_all + 9;
        D.tcount = DB.tcount = reinterpret_cast<unsigned long long*>(p);
        p += ntiles;
        h->plan = reinterpret_cast<int*>(p);
        if (hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess) {
            // the handle was value-initialised: what was not created is null
            if (h->aux) (void)hipStreamDestroy(h->aux);
            if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
            if (h->ev_join) (void)hipEventDestroy(h->ev_join);
            (void)hipFree(h->defer_slab);
            (void)hipFree(h->slab);
            delete h;
            return HFV_EDEVICE;
        }
    }
    h->launches = 0;
    h->defer_clear = -1;
    h->graph = nullptr;
    h->graph_len = 0;
    h->w_filled = -1;
    if (hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking) != hipSuccess) {
        if (h->aux) {
            (void)hipStreamDestroy(h->aux);
            (void)hipEventDestroy(h->ev_fork);
            (void)hipEventDestroy(h->ev_join);
        }
        if (h->defer_slab) (void)hipFree(h->defer_slab);
        (void)hipFree(h->slab);
        delete h;
        return HFV_EDEVICE;
    }
    *out = h;
    return HFV_OK;
}

void hfv_mh2d_destroy(hfv_mh2d* h)
{
    if (!h) return;
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    (void)hipStreamSynchronize(h->own);
    (void)hipStreamDestroy(h->own);
    if (h->aux) {
        (void)hipStreamSynchronize(h->aux);
        (void)hipStreamDestroy(h->aux);
        (void)hipEventDestroy(h->ev_fork);
        (void)hipEventDestroy(h->ev_join);
    }
    if (h->stage) (void)hipFree(h->stage);
    if (h->defer_slab) (void)hipFree(h->defer_slab);
    (void)hipFree(h->slab);
    delete h;
}

int hfv_mh2d_upload(hfv_mh2d* h, const double* const W[4], const double* const U[4], void* stream)
{
    hipStream_t s = pick(h, stream);
    const Grid2D& G = h->G;
    const size_t dp = (size_t)G.pitch * sizeof(double), wb = (size_t)G.tx * sizeof(double);
    for (int f = 0; f < 4; f++) {
        if (field_to_device(h, h->W.f[f], W[f], s) != hipSuccess) return HFV_EDEVICE;
        // both buffers start from the same conserved state: cells a step does
        // not write (deeper ghost layers, corners) stay valid in either one
        if (field_to_device(h, h->U[0].f[f], U[f], s) != hipSuccess) return HFV_EDEVICE;
        if (hipMemcpy2DAsync(h->U[1].f[f], dp, h->U[0].f[f], dp, wb, G.ty, hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
            return HFV_EDEVICE;
    }
    return HFV_OK;
}

int hfv_mh2d_start(hfv_mh2d* h, void* stream)
{
    hipStream_t s = pick(h, stream);
    h->launches = 0;
    h->w_filled = -1;
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(64), 0, s, h->ctrl, 0.0, 0LL);
    const dim3 grid((unsigned)std::min(h->G.ny, 2048)), block(SMAX_THREADS);
    hipLaunchKernelGGL(k_smax_from_w, grid, block, 0, s, h->G, h->W, h->ctrl, 0);
    launch_step(h, true, s);  // first step from the caller's primitive state
    return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_mh2d_advance(hfv_mh2d* h, long long launches, int use_graph, void* stream)
{
    hipStream_t s = pick(h, stream);
    if (device_bc_mode(h)) {
        // every step reads the primitive buffer that the device fill, hooks
        // included, has just refreshed
        if (!device_bc_complete(h)) return HFV_EINVAL;
        for (long long i = 0; i < launches; i++) {
            if (h->w_pending) h->w_pending = 0;
            else fill_primitive(h, s);
            launch_step(h, true, s);
        }
        return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
    }
    if (h->w_pending && launches > 0) {
        launch_step(h, true, s);
        h->w_pending = 0;
        launches--;
    }
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
            if (h->graph) (void)hipGraphExecDestroy(h->graph);
            h->graph = nullptr;
            hipGraph_t graph;
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return HFV_EDEVICE;
            const long long saved = h->launches, saved_clear = h->defer_clear;
            for (int i = 0; i < GRAPH_LEN; i++) launch_step(h, false, s);
            h->launches = saved;
            h->defer_clear = saved_clear;
            if (hipStreamEndCapture(s, &graph) != hipSuccess) return HFV_EDEVICE;
            if (hipGraphInstantiate(&h->graph, graph, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return HFV_EDEVICE;
            }
            (void)hipGraphDestroy(graph);
            h->graph_phase = phase;
            h->graph_stream = s;
        }
        if (hipGraphLaunch(h->graph, s) != hipSuccess) return HFV_EDEVICE;
        h->launches += GRAPH_LEN;
        h->defer_clear = h->launches;
        launches -= GRAPH_LEN;
    }
    return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_mh2d_set_pause(hfv_mh2d* h, double t_pause, long long resume_n)
{
    if (!(t_pause > 0.0) || resume_n < 0) return HFV_EINVAL;
    h->rp.t_pause = t_pause;
    h->rp.resume_n = resume_n;
    if (h->graph) (void)hipGraphExecDestroy(h->graph);  // its launches carry the old parameters
    h->graph = nullptr;
    return HFV_OK;
}

int hfv_mh2d_status(hfv_mh2d* h, hfv_run_status* st, void* stream)
{
    hipStream_t s = pick(h, stream);
    Ctrl c;
    const int slot = (int)(h->launches % 3);
    if (hipMemcpyAsync(&c, h->ctrl + slot, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess) return HFV_EDEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return HFV_EDEVICE;
    fill_status(c, h->rp, st);
    return HFV_OK;
}

int hfv_mh2d_download(hfv_mh2d* h, double* const W[4], double* const U[4], void* stream)
{
    hipStream_t s = pick(h, stream);
    hfv_run_status st;
    int rc = hfv_mh2d_status(h, &st, stream);
    if (rc) return rc;
    const Grid2D& G = h->G;
    // A step whose flux failed ran to its end on the device, but the
    // reference stops inside it; the state returned is the one at the start
    // of that step, which the other buffer still holds.
    const bool flux_failed = st.err_kind != ERR_NONE && st.err_kind != ERR_DT_ZERO;
    const Fields4 cur = h->U[flux_failed ? st.cur ^ 1 : st.cur];
    // only when asked for: W may hold a state set by hfv_mh2d_set_primitive
    if (W && device_bc_mode(h) && !flux_failed) {
        if (!device_bc_complete(h)) return HFV_EINVAL;
        if (!h->w_pending) fill_primitive(h, s);
    } else if (W && !(flux_failed && st.num_steps == 1)) {
        // (a first step that failed leaves the caller's primitive state in W)
        prim_from_u(G, cur, h->W, s);
    }
    for (int f = 0; f < 4; f++) {
        if (W && field_to_host(h, W[f], h->W.f[f], s) != hipSuccess) return HFV_EDEVICE;
        if (U && field_to_host(h, U[f], cur.f[f], s) != hipSuccess) return HFV_EDEVICE;
    }
    return hipStreamSynchronize(s) == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_mh2d_primitive(hfv_mh2d* h, double* const W[4], void* stream)
{
    hipStream_t s = pick(h, stream);
    hfv_run_status st;
    int rc = hfv_mh2d_status(h, &st, stream);
    if (rc) return rc;
    const Grid2D& G = h->G;
    prim_from_u(G, h->U[st.cur], h->W, s, 0);
    for (int f = 0; f < 4; f++)
        if (field_to_host(h, W[f], h->W.f[f], s) != hipSuccess) return HFV_EDEVICE;
    return hipStreamSynchronize(s) == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_mh2d_set_primitive(hfv_mh2d* h, const double* const W[4], void* stream)
{
    hipStream_t s = pick(h, stream);
    const Grid2D& G = h->G;
    for (int f = 0; f < 4; f++)
        if (field_to_device(h, h->W.f[f], W[f], s) != hipSuccess) return HFV_EDEVICE;
    // the next step's CFL maxima come from this state (utils.c:55-95)
    const int slot = (int)(h->launches % 3);
    hipLaunchKernelGGL(k_clear_keys, dim3(1), dim3(64), 0, s, h->ctrl, slot);
    const dim3 grid((unsigned)std::min(G.ny, 2048)), block(SMAX_THREADS);
    hipLaunchKernelGGL(k_smax_from_w, grid, block, 0, s, G, h->W, h->ctrl, slot);
    h->w_pending = 1;
    if (hipGetLastError() != hipSuccess) return HFV_EDEVICE;
    return hipStreamSynchronize(s) == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_mh2d_set_device_bc(hfv_mh2d* h, int side, hfv_device_bc_2d_func fn, void* user)
{
    if (side < 0 || side > 3 || h->G.bc[side] != BC_CUSTOM) return HFV_EINVAL;
    h->dev_fn[side] = fn;
    h->dev_user[side] = fn ? user : nullptr;
    h->w_filled = -1;
    return HFV_OK;
}

int hfv_probe_exact_finish(int impl, double gamma, double tol, int cap, long long n, const double* wl,
                           const double* wr, double* p, int* rc)
{
    if (impl < 0 || impl > 2 || cap < 0 || cap > 100 || n < 0) return HFV_EINVAL;
    if (n == 0) return HFV_OK;
    double *dl = nullptr, *dr = nullptr, *dp = nullptr;
    int* dc = nullptr;
    int ret = HFV_OK;
    if (hipMalloc(&dl, (size_t)n * 4 * sizeof(double)) != hipSuccess ||
        hipMalloc(&dr, (size_t)n * 4 * sizeof(double)) != hipSuccess ||
        hipMalloc(&dp, (size_t)n * sizeof(double)) != hipSuccess || hipMalloc(&dc, (size_t)n * sizeof(int)) != hipSuccess) {
        ret = HFV_ENOMEM;
    } else if (hipMemcpy(dl, wl, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dr, wr, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        ret = HFV_EDEVICE;
    } else {
        if (impl == 2)
            hipLaunchKernelGGL(k_probe_resume_quad, dim3((unsigned)((n + SLOW_FACES_PER_WAVE - 1) / SLOW_FACES_PER_WAVE)),
                               dim3(LANES), 0, 0, gamma, tol, cap, n, dl, dr, dp, dc);
        else if (impl == 1)
            hipLaunchKernelGGL(k_probe_finish<true>, dim3((unsigned)n), dim3(LANES), 0, 0, gamma, tol, cap, n, dl, dr,
                               dp, dc);
        else
            hipLaunchKernelGGL(k_probe_finish<false>, dim3((unsigned)((n + LANES - 1) / LANES)), dim3(LANES), 0, 0,
                               gamma, tol, cap, n, dl, dr, dp, dc);
        if (hipGetLastError() != hipSuccess || hipMemcpy(p, dp, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(rc, dc, (size_t)n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            ret = HFV_EDEVICE;
    }
    (void)hipFree(dl);
    (void)hipFree(dr);
    (void)hipFree(dp);
    (void)hipFree(dc);
    return ret;
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

int hfv_mh2d_run_noalloc(const hfv_mh2d_config* cfg, double* const W[4], double* const U[4], long long pitch,
                         long long launches, hfv_run_status* st, void* stream)
{
    if (!cfg || !W || !U || launches < 1) return HFV_EINVAL;
    int rc = check_config(cfg);
    if (rc != HFV_OK) return rc;
    // the exact solver's face queues scale with the grid; hooks need the host
    if (cfg->riemann_solver != HLLC) return HFV_EINVAL;
    for (int s = 0; s < 4; s++)
        if (cfg->bc[s] == BC_CUSTOM) return HFV_EINVAL;
    for (int f = 0; f < 4; f++)
        if (!W[f] || !U[f]) return HFV_EINVAL;
    const long long tx = (long long)cfg->num_cells_x + 2 * cfg->num_ghost_cells_side;
    const long long ty = (long long)cfg->num_cells_y + 2 * cfg->num_ghost_cells_side;
    if (pitch < tx) return HFV_EINVAL;
    if (4 * pitch * ty > NOALLOC_DOUBLES - NOALLOC_HEAD) return HFV_ETOOBIG;
    double* base = scratch_base();
    if (!base) return HFV_EDEVICE;

    // a run description on the host stack; its buffers are the caller's
    // arrays and the scratch
    hfv_mh2d h{};
    h.cfg = *cfg;
    grid_from_config(cfg, pitch, h.G, h.rp);
    const Grid2D& G = h.G;
    h.ctrl = reinterpret_cast<Ctrl*>(base);
    for (int f = 0; f < 4; f++) {
        h.U[0].f[f] = U[f];
        h.U[1].f[f] = base + NOALLOC_HEAD + pitch * ty * f;
        h.W.f[f] = W[f];
    }
    hipStream_t s = (hipStream_t)stream;
    const dim3 cgrid((unsigned)((G.tx + 255) / 256), (unsigned)G.ty), cblock(256);
    // both buffers start from the same conserved state (cells a step does
    // not write stay valid in either)
    hipLaunchKernelGGL(k_copy_fields, cgrid, cblock, 0, s, G, h.U[1], h.U[0], (const Ctrl*)nullptr, 0, 0);
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(64), 0, s, h.ctrl, 0.0, 0LL);
    hipLaunchKernelGGL(k_smax_from_w, dim3((unsigned)std::min(G.ny, 2048)), dim3(SMAX_THREADS), 0, s, G, h.W,
                       h.ctrl, 0);
    launch_step(&h, true, s);  // first step from the caller's primitive state
    for (long long i = 1; i < launches; i++) launch_step(&h, false, s);
    const int slot = (int)(launches % 3);
    hipLaunchKernelGGL(k_copy_fields, cgrid, cblock, 0, s, G, h.U[0], h.U[1], (const Ctrl*)h.ctrl, slot, 1);
    prim_from_u(G, h.U[0], h.W, s, 1, h.ctrl, slot);
    if (hipGetLastError() != hipSuccess) return HFV_EDEVICE;
    if (st) {
        Ctrl c;
        if (hipMemcpyAsync(&c, h.ctrl + slot, sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return HFV_EDEVICE;
        fill_status(c, h.rp, st);
        st->cur = 0;  // the state is back in the caller's U
    }
    return HFV_OK;
}

}  // extern "C"
