// mh2d_strip.hpp -- one wavefront's share of a 2D MUSCL-Hancock step.
//
// A wavefront (64 lanes) owns a strip of 60 interior columns plus 2 halo
// columns on each side (lane l <-> column x0-2+l) and a run of strip_rows
// interior rows, which it sweeps from bottom to top.  Each lane keeps a
// 3-row window of primitive states in registers, so every cell's primitive
// state is derived once per strip from the conserved state in HBM (W is never
// stored between steps).  x-neighbours are exchanged with cross-lane
// shuffles; y-neighbours are the lane's own register window.
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
#pragma once

#include <hip/hip_runtime.h>

#include "hfv_ctrl.hpp"
#include "hfv_math.hpp"
#include "hfv_pow.hpp"
#include "hfv_strips.hpp"

namespace hfv {


constexpr int LANES = 64;
constexpr int HALO = 2;
constexpr int STRIP_COLS = LANES - 2 * HALO;  // 60 output columns per wave
static_assert(STRIP_COLS == STRIP_COLS_2D, "strip width of hfv_strips.hpp");

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
    int full_tiles = 0;  // 1: no uniform-row form (uniform_rows)
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

// A cell's stored values as loaded, before conversion: the conserved state
// of the cell the boundary mapping names (the cell itself inside the grid)
// and the velocity flips, or (WMODE) the primitive state itself.
struct RawCell {
    double v[4];
    int fu, fv;
};

template <bool WMODE>
__device__ __forceinline__ void load_raw(const Grid2D& G, const Fields4& U, const Fields4& W, int q, int c,
                                         RawCell& rc)
{
    q = min(max(q, 0), G.ty - 1);
    c = min(max(c, 0), G.tx - 1);
    if constexpr (WMODE) {
        const long long k = (long long)q * G.pitch + c;
#pragma unroll
        for (int f = 0; f < 4; f++) rc.v[f] = W.f[f][k];
        rc.fu = rc.fv = 0;
    } else {
        const int cs = map_index(c, G.nx, G.g, G.bc[0], G.bc[1], &rc.fu);
        const int qs = map_index(q, G.ny, G.g, G.bc[2], G.bc[3], &rc.fv);
        const long long k = (long long)qs * G.pitch + cs;
#pragma unroll
        for (int f = 0; f < 4; f++) rc.v[f] = U.f[f][k];
    }
}

// Primitive state from loaded values: c2p (system.c:763-802) and the
// boundary flips.
// rvol, vol_ok: rcp_core(G.volume), div_core_ok(G.volume) (c2p_2d_rn).
template <bool WMODE>
__device__ __forceinline__ void raw_to_prim(const Grid2D& G, const RawCell& rc, double* w, double rvol, bool vol_ok)
{
    if constexpr (WMODE) {
#pragma unroll
        for (int f = 0; f < 4; f++) w[f] = rc.v[f];
    } else {
        c2p_2d_rn(G.gamma, G.volume, rvol, vol_ok, rc.v[0], rc.v[1], rc.v[2], rc.v[3], w);
        if (rc.fu) w[1] = -w[1];
        if (rc.fv) w[2] = -w[2];
    }
}

// Primitive state of cell (row q, column c) as the reference's stencil sees it.
template <bool WMODE>
__device__ __forceinline__ void load_prim(const Grid2D& G, const Fields4& U, const Fields4& W, int q, int c,
                                          double* w, double rvol, bool vol_ok)
{
    RawCell rc;
    load_raw<WMODE>(G, U, W, q, c, rc);
    raw_to_prim<WMODE>(G, rc, w, rvol, vol_ok);
}

// y-predictor of the cell whose own state is wc, with wb below and wt above
// (c = dt A_y / V).  Returns the cell's sound speed and *cxr = cx / rho
// (cx = dt A_x / V), which its x-predictor uses a row later.  The three
// divisions by the cell's density (the sound speed's gamma p / rho, c / rho
// and cx / rho) share one reciprocal when every lane's operands allow
// (div_core); the bits of sound_speed and predict_y / predict_x.
template <int LIM>
__device__ __forceinline__ double ypredict(const double* wb, const double* wc, const double* wt, double c,
                                           double cx, double gamma, double* lo, double* hi, double* cxr)
{
    double d[4];
#pragma unroll
    for (int f = 0; f < 4; f++) d[f] = limit<LIM>(wc[f] - wb[f], wt[f] - wc[f]);
    const double rho = wc[0], gp = gamma * wc[3];
    double a2[1], cyr;
    if (wave_all(div_core_ok(rho) & div_core_ok(gp) & div_core_ok(c) & div_core_ok(cx))) {
        const double rr = rcp_core(rho);
        a2[0] = div_core(gp, rho, rr);
        cyr = div_core(c, rho, rr);
        *cxr = div_core(cx, rho, rr);
    } else {
        a2[0] = gp / rho;
        cyr = c / rho;
        *cxr = cx / rho;
    }
    sqrt_rn<1>(a2);
    predict_y_cr(wc, d, c, cyr, a2[0], lo, hi);
    return a2[0];
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

// pow_pos's log and exp tables copied into LDS (pow_lds: POW_TABS_LEN
// doubles, log table first).  Read from LDS, a table lookup waits on the LDS
// counter; read from global memory it would wait on the vector-memory
// counter, which completes in order and so would also wait for the row loads
// issued ahead of it.  Every lane of the one-wavefront workgroup calls this.
constexpr int POW_LOG_TAB_LEN = 48 * 3, POW_TABS_LEN = 48 * 3 + 64 * 2;
__device__ __forceinline__ void stage_pow_tables(double* pow_lds, int lane)
{
    for (int i = lane; i < POW_LOG_TAB_LEN; i += LANES) pow_lds[i] = POW_LOG_TAB_D[i];
    for (int i = lane; i < POW_TABS_LEN - POW_LOG_TAB_LEN; i += LANES) pow_lds[POW_LOG_TAB_LEN + i] = POW_EXP_TAB_D[i];
    __syncthreads();
}

// The powers of the two-rarefaction pressure guesses pending on two faces per
// lane (riemann_solver.c:191-199) that trrs_settle_direct cannot settle
// (about 1 face in 1000 on a Kelvin-Helmholtz run), evaluated 64 at a time
// over the wavefront: the lanes' pending bases are packed into LDS, each lane raises
// one packed value per pass (pow_pos, hfv_pow.hpp), and the results are
// read back.  A row of a
// strip has about 30 such faces in 128, so the first two powers take one or
// two passes; the last power is needed only by the rare faces trrs_settle
// cannot settle (about 1 in 1000 on a Kelvin-Helmholtz run), and its pass
// is skipped when no lane has one.  q: 384 doubles of LDS; the workgroup is
// this one wavefront, and every lane must call this.
__device__ __forceinline__ void trrs_powers(double gamma, double tol, HllcFace& A, HllcFace& B, double* q, int lane,
                                            const double* pow_lds)
{
    // most faces are settled without powers (trrs_settle_direct); the
    // passes below take the others
    bool ta = A.kind == HLLC_TRRS, tb = B.kind == HLLC_TRRS;
    if (__ballot(ta | tb) == 0ull) return;
    if (ta) ta = !trrs_settle_direct(gamma, tol, A);
    if (tb) tb = !trrs_settle_direct(gamma, tol, B);
    const unsigned long long ma = __ballot(ta), mb = __ballot(tb);
    if ((ma | mb) == 0ull) return;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int na = __popcll(ma), n = na + __popcll(mb);
    const int ia = __popcll(ma & below), ib = na + __popcll(mb & below);
    const double h = 0.5 * (gamma - 1.0);

    // pow(pl, h/gamma) and pow(pr, h/gamma) of face i at q[2i], q[2i+1]
    if (ta) { q[2 * ia] = A.wl[3]; q[2 * ia + 1] = A.wr[3]; }
    if (tb) { q[2 * ib] = B.wl[3]; q[2 * ib + 1] = B.wr[3]; }
    __syncthreads();
    const double e1 = h / gamma;
    for (int b = 0; b < 2 * n; b += LANES) {
        const int i = b + lane;
        if (i < 2 * n) q[i] = pow_pos(q[i], e1, pow_lds, pow_lds + POW_LOG_TAB_LEN);
    }
    __syncthreads();
    double xa = 0.0, xb = 0.0;
    const bool la = ta && !trrs_settle(tol, A, q[2 * ia], q[2 * ia + 1], &xa);
    const bool lb = tb && !trrs_settle(tol, B, q[2 * ib], q[2 * ib + 1], &xb);
    const unsigned long long ra = __ballot(la), rb = __ballot(lb);
    if ((ra | rb) == 0ull) return;
    // pow(X, gamma/h) of the unsettled faces at q[256 + j]
    const int nra = __popcll(ra), m = nra + __popcll(rb);
    const int ja = __popcll(ra & below), jb = nra + __popcll(rb & below);
    double* q2 = q + 4 * LANES;
    if (la) q2[ja] = xa;
    if (lb) q2[jb] = xb;
    __syncthreads();
    const double e2 = gamma / h;
    for (int b = 0; b < m; b += LANES) {
        const int i = b + lane;
        if (i < m) q2[i] = pow_pos(q2[i], e2, pow_lds, pow_lds + POW_LOG_TAB_LEN);
    }
    __syncthreads();
    if (la) {
        A.ps = q2[ja];
        if (A.ps < tol) A.ps = tol;
        A.kind = HLLC_READY;
    }
    if (lb) {
        B.ps = q2[jb];
        if (B.ps < tol) B.ps = tol;
        B.kind = HLLC_READY;
    }
    __syncthreads();
}

// The two-shock guesses of the two faces per lane (riemann_solver.c:
// 201-215), taken only for the faces that use them: their inputs are packed
// into LDS, each lane computes one packed face per pass, and the results are
// read back.  About 30 faces in 128 take it on a Kelvin-Helmholtz run, so one
// pass replaces evaluating it on every face.  q: 6 * 64 doubles of LDS; the
// workgroup is this one wavefront, and every lane must call this.
__device__ __forceinline__ void tsrs_compact(double gamma, const HllcFace (&F)[2], HllcGuess<2>& G, double* q,
                                             int lane)
{
    const bool ta = needs_tsrs<2>(G, 0), tb = needs_tsrs<2>(G, 1);
    const unsigned long long ma = __ballot(ta), mb = __ballot(tb);
    if ((ma | mb) == 0ull) return;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int na = __popcll(ma), n = na + __popcll(mb);
    const int ia = __popcll(ma & below), ib = na + __popcll(mb & below);
    for (int b = 0; b < n; b += LANES) {
        // input j of packed face s at q[j * LANES + s]
        const bool pa = ta && ia >= b && ia < b + LANES, pb = tb && ib >= b && ib < b + LANES;
        if (pa) {
            const int sl = ia - b;
            q[0 * LANES + sl] = F[0].wl[0];
            q[1 * LANES + sl] = F[0].wl[3];
            q[2 * LANES + sl] = F[0].wr[0];
            q[3 * LANES + sl] = F[0].wr[3];
            q[4 * LANES + sl] = F[0].wr[1] - F[0].wl[1];
            q[5 * LANES + sl] = G.ppv[0];
        }
        if (pb) {
            const int sl = ib - b;
            q[0 * LANES + sl] = F[1].wl[0];
            q[1 * LANES + sl] = F[1].wl[3];
            q[2 * LANES + sl] = F[1].wr[0];
            q[3 * LANES + sl] = F[1].wr[3];
            q[4 * LANES + sl] = F[1].wr[1] - F[1].wl[1];
            q[5 * LANES + sl] = G.ppv[1];
        }
        __syncthreads();
        if (b + lane < n) {
            const double t = tsrs_guess(gamma, q[0 * LANES + lane], q[1 * LANES + lane], q[2 * LANES + lane],
                                        q[3 * LANES + lane], q[4 * LANES + lane], q[5 * LANES + lane]);
            q[lane] = t;  // this lane's own slot, its inputs already read
        }
        __syncthreads();
        if (pa) G.tsrs[0] = q[ia - b];
        if (pb) G.tsrs[1] = q[ib - b];
        __syncthreads();
    }
}

// The HLLC fluxes of a lane's two faces in one row: the top face of its
// cell (states yt below, ylo above, (rho, u, v, p) order) and its left face
// (hl from the left neighbour, xlo its own); fyt and fxl in (mass, mx, my,
// E) order.  The wavefront's two-shock guesses and two-rarefaction powers
// are taken together (tsrs_compact, trrs_powers), so every lane must call
// this.  trrs_q: 6 * 64 doubles of LDS; pow_lds: the staged pow tables.
template <bool BRANCHY>
__device__ __forceinline__ void hllc_row_faces(double gamma, double tol, const double* yt, const double* ylo,
                                               const double* hl, const double* xlo, double* fyt, double* fxl,
                                               int* ey, int* ex, double* trrs_q, int lane, const double* pow_lds)
{
    HllcFace F[2];  // y face (normal velocity first), x face
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const int t = (f == 1) ? 2 : (f == 2) ? 1 : f;
        F[0].wl[f] = yt[t];
        F[0].wr[f] = ylo[t];
        F[1].wl[f] = hl[f];
        F[1].wr[f] = xlo[f];
    }
    double o[2][4];
    if constexpr (BRANCHY) {
        hllc_begin_branchy(gamma, tol, F[0]);
        hllc_begin_branchy(gamma, tol, F[1]);
        trrs_powers(gamma, tol, F[0], F[1], trrs_q, lane, pow_lds);
        *ey = hllc_end_branchy(gamma, F[0], o[0]);
        *ex = hllc_end_branchy(gamma, F[1], o[1]);
    } else {
        HllcGuess<2> Gs;
        hllc_guess_begin<2>(gamma, F, Gs);
        tsrs_compact(gamma, F, Gs, trrs_q, lane);
        hllc_guess_end<2>(tol, F, Gs);
        trrs_powers(gamma, tol, F[0], F[1], trrs_q, lane, pow_lds);
        // both faces' regular fluxes together, then the rare vacuum faces
        int e2[2];
        hllc_end_regular_faces<2>(gamma, F, o, e2);
        *ey = e2[0];
        *ex = e2[1];
        if (F[0].kind == HLLC_VACUUM) {
            hllc_vacuum_flux<4>(gamma, F[0].wl, F[0].al, F[0].wr, F[0].ar, o[0]);
            *ey = ERR_NONE;
        }
        if (F[1].kind == HLLC_VACUUM) {
            hllc_vacuum_flux<4>(gamma, F[1].wl, F[1].al, F[1].wr, F[1].ar, o[1]);
            *ex = ERR_NONE;
        }
    }
    fyt[0] = o[0][0];
    fyt[1] = o[0][2];
    fyt[2] = o[0][1];
    fyt[3] = o[0][3];
#pragma unroll
    for (int f = 0; f < 4; f++) fxl[f] = o[1][f];
}

// Neighbour exchange across the wavefront by the GFX9 whole-wavefront DPP
// rotations (v_mov_b32_dpp wave_ror:1 / wave_rol:1), register to register,
// one 32-bit move per half of a double: lane i receives lane i-1's
// (from_left) or lane i+1's (from_right) value, and the edge lane the value
// of the lane at the other edge.  A rotation writes every lane, so the
// destination needs no copy of the old value first (a shift would, for the
// edge lane it leaves alone).  The edge lanes are outer halo columns of a
// strip (lane 0: column x0 - 2, lane 63: x0 + 61); what they receive feeds
// only their own slopes, the faces left and right of x0 - 2 and left of
// x0 + 61, and the right-face fluxes of lanes 62 and 63.  No cell that a
// strip updates and no face that it reports, queues or stores uses these
// (strip_step's lane roles); a wavefront-wide range test that one of them
// fails sends its group down the library form, whose bits are the same.
constexpr int DPP_WAVE_ROL1 = 0x134;
constexpr int DPP_WAVE_ROR1 = 0x13C;

template <int CTRL>
__device__ __forceinline__ double dpp_rotate(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = (int)b, hi = (int)(b >> 32);
    const int lo2 = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
    const int hi2 = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi2 << 32) | (unsigned int)lo2);
}

__device__ __forceinline__ double shfl_from_left(double v) { return dpp_rotate<DPP_WAVE_ROR1>(v); }
__device__ __forceinline__ double shfl_from_right(double v) { return dpp_rotate<DPP_WAVE_ROL1>(v); }

// DEFER (exact solver): a face whose two states differ (its Riemann problem
// needs the pressure iteration, or the vacuum sampler) is not solved by the
// lane that reconstructs it but queued, with
// its two states, for k_solve_faces, which spreads the queue over the whole
// chip; a wavefront of a blast region would otherwise solve its rows' faces
// one after another, with its lanes diverging, while the rest of the chip
// idles.  A cell next to a queued face is deferred too: its other faces'
// fluxes go to fx/fy and k_update_deferred applies the update, in the
// reference's order, once every flux is known.  Every other face and cell
// is done in the strip as usual.
struct Deferred {
    double* fx[4];              // flux through the face left of cell (q, c), at q * pitch + c
    double* fy[4];              // flux through the face above cell (q, c), at q * pitch + c
    double* qw[8];              // queued faces: left then right state, in the solver's order
    long long* qid;             // queued faces: (q * pitch + c) * 2 + (1: face above, 0: face left)
    long long* cells;           // deferred cells: (q * pitch + c) * 8 + DeferKind
    // queued faces still iterating after k_solve_faces' Newton cap: the
    // face's queue index and its NewtonState (p0, p_hi, f_hi, p_lo, f_lo;
    // it; bracketed), finished by k_solve_slow
    long long* sidx;
    double* sst[5];
    int* sit;
    int* sbr;
    // per set-aside face: the bracket scan and bisection that finish its
    // iteration if the Newton loop ends without converging and without a new
    // bracket (k_solve_spec): p* and the failure kind
    double* spp;
    int* sprc;
    unsigned long long* count;  // [0] queued faces, [1] deferred cells, [2] set-aside faces
    long long face_cap, cell_cap, slow_cap;
    // faces queued per strip tile (sy * nsx + sx) by the step's strip passes,
    // which the next step's split of the tiles reads (mh2d.hip)
    unsigned long long* tcount;
};

enum DeferKind : int { DEF_INTERIOR = 0, DEF_GHOST_XLO = 1, DEF_GHOST_XHI = 2, DEF_GHOST_YLO = 3, DEF_GHOST_YHI = 4 };

// Wave-aggregated append: one atomic per wavefront; each lane with `take`
// gets its own slot (-1 for the others).  Every lane must call this.
__device__ __forceinline__ long long wave_append(unsigned long long* counter, bool take, int lane)
{
    const unsigned long long m = __ballot(take);
    if (m == 0ull) return -1;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0ull;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, LANES);
    return take ? (long long)(base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull))) : -1;
}

// The exact flux of a face in (mass, mx, my, E) order when its states are
// equal (exact_sample's shortcut); true (f zero) when it must be queued.
// YFACE: the solver sees (rho, v, u, p), as in yflux.
template <bool YFACE>
__device__ __forceinline__ bool direct_face(double gamma, double tol, const double* l, const double* r, double* f)
{
    double ls[4], rs[4], o[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int t = YFACE ? ((q == 1) ? 2 : (q == 2) ? 1 : q) : q;
        ls[q] = l[t];
        rs[q] = r[t];
    }
    if (flux_exact<4, true>(gamma, tol, ls, rs, o) == ERR_PENDING) {
#pragma unroll
        for (int q = 0; q < 4; q++) f[q] = 0.0;
        return true;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) f[q] = o[YFACE ? ((q == 1) ? 2 : (q == 2) ? 1 : q) : q];
    return false;
}

// Queue the owned faces that need the iteration (states in the solver's order);
// *nq (wave-uniform) counts the wavefront's queued faces.
template <bool YFACE>
__device__ __forceinline__ void queue_face(const Deferred& D, bool take, long long k, const double* l,
                                           const double* r, int lane, unsigned long long* nq)
{
    *nq += (unsigned long long)__popcll(__ballot(take));
    const long long i = wave_append(&D.count[0], take, lane);
    if (i >= 0 && i < D.face_cap) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int t = YFACE ? ((q == 1) ? 2 : (q == 2) ? 1 : q) : q;
            D.qw[q][i] = l[t];
            D.qw[4 + q][i] = r[t];
        }
        D.qid[i] = k * 2 + (YFACE ? 1 : 0);
    }
}

__device__ __forceinline__ void defer_cell(const Deferred& D, bool take, long long k, int kind, int lane)
{
    const long long i = wave_append(&D.count[1], take, lane);
    if (i >= 0 && i < D.cell_cap) D.cells[i] = k * 8 + kind;
}

__device__ __forceinline__ void store_flux(double* const* F, long long k, const double* f)
{
#pragma unroll
    for (int q = 0; q < 4; q++) F[q][k] = f[q];
}

// Rows [y0, y1) of strip row sy, or of part sub of it when its rows are
// split in nsub runs of ceil(strip_rows / nsub) (empty when y0 >= y1).
__device__ __forceinline__ void strip_rows_of(const Grid2D& G, int sy, int sub, int nsub, int* y0, int* y1)
{
    const int base = G.g + sy * G.strip_rows;
    const int end = min(base + G.strip_rows, G.g + G.ny);
    const int run = (G.strip_rows + nsub - 1) / nsub;
    *y0 = base + sub * run;
    *y1 = min(*y0 + run, end);
}

struct StripResult {
    double smx, smy;     // CFL maxima of the strip's updated cells
    int err;             // ErrKind of a failed flux (0: none)
    long long err_cell;
};

// Rows of one state (the quiet part of a blast, a quadrant's interior): a
// row's results depend only on the cells of its stencil, rows r - AHEAD ..
// r + AHEAD of the wavefront's 64 columns as load_raw maps them.  When all
// of those hold, bit for bit, the same conserved state (and velocity flips)
// as lane 0's first cell, every cell and face of the row performs the
// operations of one cell on the same values.  uniform_rows computes those
// results once, from that state, by the calls the row loop of strip_step
// makes, in its order, with every neighbour equal to the cell itself (a
// neighbour exchange then returns the lane's own value, and every slope is
// limit(0, 0)); then it verifies the strip's rows 16 at a time and stores
// each row whose stencil is verified as soon as it is: the interior cells'
// new state, the ghost columns' one-face updates (each from its own stored
// state) and, at the grid's bottom and top, the ghost rows' one-face
// updates.  Returns the first row it has not stored (y0: none, when lane
// 0's first cell is not every lane's, or when a face would be queued
// (DEFER) or fails); the caller runs the rows from there as usual.  *smx,
// *smy: the stored cells' CFL maxima.  Every lane must call this
// (wavefront-wide guesses and powers in HLLC).
template <int LIM, int SOLVER, bool BRANCHY, bool DEFER>
__device__ __forceinline__ int uniform_rows(const Grid2D& G, double dt, const Fields4& U, const Fields4& V,
                                            int y0, int y1, int col, bool out_col, bool ghost_lo_col,
                                            bool ghost_hi_col, double cx, double cy, double rvol, bool vol_ok,
                                            double* trrs_q, int lane, const double* pow_lds, double* smx,
                                            double* smy)
{
    constexpr int AHEAD = (LIM == FIRST_ORDER) ? 1 : 2;
    const Fields4 none{};
    RawCell rc;
    load_raw<false>(G, U, none, y0 - AHEAD, col, rc);
    unsigned long long rb[4];
    bool same = true;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(rc.v[f]);
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
        rb[f] = ((unsigned long long)hi << 32) | lo;
        same &= b == rb[f];
    }
    const int fu0 = __builtin_amdgcn_readfirstlane(rc.fu), fv0 = __builtin_amdgcn_readfirstlane(rc.fv);
    same &= (rc.fu == fu0) & (rc.fv == fv0);
    if (!wave_all(same)) return y0;

    // the results of a cell whose stencil holds rc's state
    const double gamma = G.gamma, tol = G.tol;
    double w[4];
    raw_to_prim<false>(G, rc, w, rvol, vol_ok);
    double ylo[4], yhi[4], xlo[4], xhi[4];
    if constexpr (LIM == FIRST_ORDER) {
#pragma unroll
        for (int f = 0; f < 4; f++) ylo[f] = yhi[f] = xlo[f] = xhi[f] = w[f];
    } else {
        double cxr;
        const double a = ypredict<LIM>(w, w, w, cy, cx, gamma, ylo, yhi, &cxr);
        double d[4];
#pragma unroll
        for (int f = 0; f < 4; f++) d[f] = limit<LIM>(w[f] - w[f], w[f] - w[f]);
        predict_x_cr(w, d, cx, cxr, a, xlo, xhi);
    }
    // a row's top face (yhi below, ylo above) and a lane's left face (the
    // left neighbour's xhi, its own xlo); the bottom and right faces have the
    // same states
    double fy[4], fx[4];
    int ey, ex;
    if constexpr (SOLVER == HLLC) {
        hllc_row_faces<BRANCHY>(gamma, tol, yhi, ylo, xhi, xlo, fy, fx, &ey, &ex, trrs_q, lane, pow_lds);
    } else if constexpr (DEFER) {
        const bool qy = direct_face<true>(gamma, tol, yhi, ylo, fy);
        const bool qx = direct_face<false>(gamma, tol, xhi, xlo, fx);
        if (!wave_all(!qy && !qx)) return y0;
        ey = ex = ERR_NONE;
    } else {
        ey = yflux<SOLVER>(gamma, tol, yhi, ylo, fy);
        ex = flux<SOLVER, 4>(gamma, tol, xhi, xlo, fx);
    }
    if (!wave_all((ey == ERR_NONE) & (ex == ERR_NONE))) return y0;
    // an interior cell's update (integrator_muscl_hancock.c:806-814,
    // 1092-1100; source_term.c:213-240) and its CFL maxima (utils.c:75-92)
    double u[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        double x = rc.v[f];
        x += fx[f] * dt * G.area_x;
        x -= fx[f] * dt * G.area_x;
        x += fy[f] * dt * G.area_y;
        x -= fy[f] * dt * G.area_y;
        u[f] = x;
    }
    if (G.gravity != 0.0) gravity_source(dt, G.gravity, u[0], &u[2], &u[3]);
    double sx_ = -INFINITY, sy_ = -INFINITY;
    if (out_col) {
        double wn[4];
        c2p_2d_rn(gamma, G.volume, rvol, vol_ok, u[0], u[1], u[2], u[3], wn);
        const double a = sound_speed_rn(gamma, wn[0], wn[3]);
        sx_ = a + fabs(wn[1]);
        sy_ = a + fabs(wn[2]);
    }

    // verify the rows 16 at a time (the last row repeated past the end) and
    // store the rows whose stencils are verified
    constexpr int BATCH = 16;
    const int last = y1 + AHEAD - 1;
    int verified = y0 - AHEAD;  // rows y0 - AHEAD .. verified hold the state
    int stored = y0;            // rows y0 .. stored - 1 are stored
    for (int q = y0 - AHEAD + 1; q <= last && verified == q - 1; q += BATCH) {
        RawCell c[BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; j++) load_raw<false>(G, U, none, min(q + j, last), col, c[j]);
        int nok = 0;  // the batch's leading rows that hold the state
        bool run = true;
#pragma unroll
        for (int j = 0; j < BATCH; j++) {
            bool m = (c[j].fu == fu0) & (c[j].fv == fv0);
#pragma unroll
            for (int f = 0; f < 4; f++) m &= (unsigned long long)__double_as_longlong(c[j].v[f]) == rb[f];
            run = run && wave_all(m);
            nok += run ? 1 : 0;
        }
        verified = min(q + nok - 1, last);
        const int upto = min(y1, verified - AHEAD + 1);
        if (upto > stored && stored == G.g && out_col) {
            // the bottom ghost row's update from the y_min face (1092-1095)
            const long long k = (long long)(G.g - 1) * G.pitch + col;
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] - fy[f] * dt * G.area_y;
        }
        for (int r = stored; r < upto; r++) {
            const long long k = (long long)r * G.pitch + col;
            if (out_col) {
#pragma unroll
                for (int f = 0; f < 4; f++) V.f[f][k] = u[f];
            } else if (ghost_lo_col) {
#pragma unroll
                for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] - fx[f] * dt * G.area_x;
            } else if (ghost_hi_col) {
#pragma unroll
                for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] + fx[f] * dt * G.area_x;
            }
        }
        if (upto == y1 && stored < y1 && y1 == G.g + G.ny && out_col) {
            // the top ghost row's update from the y_max face (1097-1100)
            const long long k = (long long)(G.g + G.ny) * G.pitch + col;
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] + fy[f] * dt * G.area_y;
        }
        if (upto > stored) stored = upto;
    }
    if (stored > y0) {
        if (sx_ > *smx) *smx = sx_;
        if (sy_ > *smy) *smy = sy_;
    }
    return stored;
}

// One strip (sx, sy) of the step U -> V with time step dt, or (nsub > 1) part
// sub of its rows split in nsub runs of ceil(strip_rows / nsub), which the
// caller has checked to be non-empty (strip_rows_of).  BRANCHY: the
// HLLC faces by branches (hllc_begin_branchy) instead of selects; the same
// bits either way.  LIM == FIRST_ORDER: the first-order Godunov step of
// integrator_godunov_first_order.c:362-497, whose faces take the cells' own
// primitive states; its sweeps, update order and source term are those of
// the MUSCL-Hancock step, and it needs one ghost layer, not two.
template <int LIM, int SOLVER, bool WMODE, bool BRANCHY, bool DEFER = false, bool UNIFORM = true>
__device__ __forceinline__ StripResult strip_step(const Grid2D& G, double dt, const Fields4& U, const Fields4& V,
                                                  const Fields4& W, int sx, int sy, int lane,
                                                  const Deferred& D = Deferred(), int sub = 0, int nsub = 1)
{
    static_assert(!DEFER || SOLVER == EXACT, "only exact-solver faces are queued");
    const double cx = dt * G.area_x / G.volume;  // dt*A/V, integrator_muscl_hancock.c:636
    const double cy = dt * G.area_y / G.volume;  // :921
    const double rvol = rcp_core(G.volume);      // the c2p divisions by the cell volume (c2p_2d_rn)
    const bool vol_ok = div_core_ok(G.volume);

    const int g = G.g;
    const int x0 = g + sx * STRIP_COLS;        // first output column of the strip
    int y0, y1;                                // output rows [y0, y1)
    strip_rows_of(G, sy, sub, nsub, &y0, &y1);
    const int col = x0 - HALO + lane;
    const bool out_col = (lane >= HALO) && (lane < HALO + STRIP_COLS) && (col < g + G.nx);
    const bool ghost_lo_col = (sx == 0) && (col == g - 1);               // x_min face flux only
    const bool ghost_hi_col = (sx == G.nsx - 1) && (col == g + G.nx);    // x_max face flux only
    const bool edge_strip = (sx == 0) || (sx == G.nsx - 1);              // holds a ghost column's lane
    const double gamma = G.gamma, tol = G.tol;
    // DEFER: the lane's x face (left of col) is used by this wavefront when
    // its own cell or the cell on its left is updated here, and queued by
    // the lane whose own cell uses it (one owner per face)
    const bool left_used = (lane >= 1) && (((lane - 1 >= HALO) && (lane - 1 < HALO + STRIP_COLS) &&
                                            (col - 1 < g + G.nx)) || ((sx == 0) && (col == g)));
    const bool need_x = out_col || ghost_hi_col || left_used;
    const bool own_x = out_col || ghost_hi_col;
    bool hyb = false;  // DEFER: the flux below the current row is queued
    unsigned long long nq = 0ull;  // DEFER: faces this wavefront queued

    int err = ERR_NONE;
    long long err_cell = -1;
    double smx = -INFINITY, smy = -INFINITY;
    __shared__ double trrs_q[6 * LANES];
    __shared__ double pow_lds[POW_TABS_LEN];
    if constexpr (SOLVER == HLLC) stage_pow_tables(pow_lds, lane);
    constexpr int AHEAD = (LIM == FIRST_ORDER) ? 1 : 2;  // the row a step's iteration r reads
    // Leading rows of one state take the results of one cell (uniform_rows),
    // and the rows from the first it has not stored run as usual; the first
    // step of a run, which reads the primitive state, does not take this
    // form, nor does a kernel built without it (UNIFORM false: the same
    // results, fewer registers).
    if constexpr (!WMODE && UNIFORM) {
        if (!G.full_tiles) {
            y0 = uniform_rows<LIM, SOLVER, BRANCHY, DEFER>(G, dt, U, V, y0, y1, col, out_col, ghost_lo_col,
                                                           ghost_hi_col, cx, cy, rvol, vol_ok, trrs_q, lane, pow_lds,
                                                           &smx, &smy);
            if (y0 >= y1) {
                StripResult res;
                res.smx = smx;
                res.smy = smy;
                res.err = ERR_NONE;
                res.err_cell = -1;
                return res;
            }
        }
    }
    // The conserved states of the last AHEAD + 1 rows loaded, each lane its
    // own column (slot q % (AHEAD + 1) holds row q): a row's update takes its
    // conserved state from here, as loaded for its c2p AHEAD iterations
    // earlier, instead of holding it in registers or loading it again.  For an
    // updated cell (interior row and column) the loaded cell is the cell
    // itself.  A lane reads only what it wrote, so no barrier is needed.
    __shared__ double urows[(AHEAD + 1) * 4 * LANES];
    const auto stash_row = [&](int q, const RawCell& rc) {
        if constexpr (!WMODE) {
            double* u = urows + (q % (AHEAD + 1)) * 4 * LANES + lane;
#pragma unroll
            for (int f = 0; f < 4; f++) u[f * LANES] = rc.v[f];
        }
    };

    // ---- prologue: rows y0-2 .. y0+1 (first order: y0-1 .. y0) ----------
    double wa[4], wb[4], wc[4];
    double a_row = 0.0, a_next = 0.0;      // sound speeds of rows r and r+1 (MUSCL)
    double cxr_row = 0.0, cxr_next = 0.0;  // cx / rho of rows r and r+1
    double yt[4], fyb[4];
    // The next row's stored values, loaded an iteration ahead.  A row's update
    // takes the row's conserved state from urows (WMODE: loads it again, at
    // the top of its iteration and before that prefetch, since the
    // vector-memory counter completes in order), so no lane holds the
    // conserved states of the rows between their c2p and their update.
    RawCell next;
    if constexpr (LIM == FIRST_ORDER) {
        double wm1[4];
        load_prim<WMODE>(G, U, W, y0 - 1, col, wm1, rvol, vol_ok);
        {
            RawCell rc;
            load_raw<WMODE>(G, U, W, y0, col, rc);
            stash_row(y0, rc);
            raw_to_prim<WMODE>(G, rc, wa, rvol, vol_ok);
        }
#pragma unroll
        for (int f = 0; f < 4; f++) {
            wb[f] = wa[f];  // unused
            yt[f] = wa[f];
        }
        if constexpr (DEFER) {
            hyb = out_col && direct_face<true>(gamma, tol, wm1, wa, fyb);
            queue_face<true>(D, hyb && y0 == g, (long long)(y0 - 1) * G.pitch + col, wm1, wa, lane, &nq);
        } else {
            const int e = yflux<SOLVER>(gamma, tol, wm1, wa, fyb);
            if (e && out_col) { err = e; err_cell = (long long)(y0 - 1) * G.tx + col; }
        }
    } else {
        double wm2[4], wm1[4];
        load_prim<WMODE>(G, U, W, y0 - 2, col, wm2, rvol, vol_ok);
        load_prim<WMODE>(G, U, W, y0 - 1, col, wm1, rvol, vol_ok);
        {
            RawCell rc;
            load_raw<WMODE>(G, U, W, y0, col, rc);
            stash_row(y0, rc);
            raw_to_prim<WMODE>(G, rc, wa, rvol, vol_ok);
            load_raw<WMODE>(G, U, W, y0 + 1, col, rc);
            stash_row(y0 + 1, rc);
            raw_to_prim<WMODE>(G, rc, wb, rvol, vol_ok);
        }
        double lo[4], hi_m1[4], lo0[4], unused;
        ypredict<LIM>(wm2, wm1, wa, cy, cx, gamma, lo, hi_m1, &unused);         // row y0-1: top face state
        a_row = ypredict<LIM>(wm1, wa, wb, cy, cx, gamma, lo0, yt, &cxr_row);   // row y0: bottom and top face states
        if constexpr (DEFER) {
            hyb = out_col && direct_face<true>(gamma, tol, hi_m1, lo0, fyb);
            queue_face<true>(D, hyb && y0 == g, (long long)(y0 - 1) * G.pitch + col, hi_m1, lo0, lane, &nq);
        } else {
            const int e = yflux<SOLVER>(gamma, tol, hi_m1, lo0, fyb);
            if (e && out_col) { err = e; err_cell = (long long)(y0 - 1) * G.tx + col; }
        }
    }
    if constexpr (DEFER) {
        // the bottom ghost row's update waits for a queued y_min face
        defer_cell(D, y0 == g && out_col && hyb, (long long)(g - 1) * G.pitch + col, DEF_GHOST_YLO, lane);
    }
    if (y0 == g && out_col && !hyb) {
        // bottom ghost row receives the y_min face flux (1092-1095)
        const long long k = (long long)(g - 1) * G.pitch + col;
        V.f[0][k] = U.f[0][k] - fyb[0] * dt * G.area_y;
        V.f[1][k] = U.f[1][k] - fyb[1] * dt * G.area_y;
        V.f[2][k] = U.f[2][k] - fyb[2] * dt * G.area_y;
        V.f[3][k] = U.f[3][k] - fyb[3] * dt * G.area_y;
    }

    load_raw<WMODE>(G, U, W, y0 + AHEAD, col, next);
    for (int r = y0; r < y1; r++) {
        // y: row r+1's face states (the flux through the top face of row r
        // follows below); x: the face states of row r, with the lane's
        // neighbours taken by shuffles
        double ylo[4], yhi[4], xlo[4], xhi[4];
        raw_to_prim<WMODE>(G, next, wc, rvol, vol_ok);  // row r + AHEAD
        stash_row(r + AHEAD, next);
        const long long k = (long long)r * G.pitch + col;
        double uk[4];  // this row's conserved state, for its update (WMODE: loaded again)
        if constexpr (WMODE) {
            if (out_col) {
#pragma unroll
                for (int f = 0; f < 4; f++) uk[f] = U.f[f][k];
            }
        }
        if (r + 1 < y1) load_raw<WMODE>(G, U, W, r + 1 + AHEAD, col, next);
        if constexpr (LIM == FIRST_ORDER) {
#pragma unroll
            for (int f = 0; f < 4; f++) {
                ylo[f] = yhi[f] = wc[f];
                xlo[f] = xhi[f] = wa[f];
            }
        } else {
            a_next = ypredict<LIM>(wa, wb, wc, cy, cx, gamma, ylo, yhi, &cxr_next);
            double wl[4], wr[4];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                wl[f] = shfl_from_left(wa[f]);
                wr[f] = shfl_from_right(wa[f]);
            }
            double d[4];
#pragma unroll
            for (int f = 0; f < 4; f++) d[f] = limit<LIM>(wa[f] - wl[f], wr[f] - wa[f]);
            predict_x_cr(wa, d, cx, cxr_row, a_row, xlo, xhi);
        }
        double hl[4];
#pragma unroll
        for (int f = 0; f < 4; f++) hl[f] = shfl_from_left(xhi[f]);
        double fyt[4], fxl[4];
        int ey, ex;
        bool hy = false, hx = false;  // DEFER: this row's top face / the lane's x face is queued
        if constexpr (SOLVER == HLLC) {
            hllc_row_faces<BRANCHY>(gamma, tol, yt, ylo, hl, xlo, fyt, fxl, &ey, &ex, trrs_q, lane, pow_lds);
        } else if constexpr (DEFER) {
#pragma unroll
            for (int f = 0; f < 4; f++) fyt[f] = fxl[f] = 0.0;
            hy = out_col && direct_face<true>(gamma, tol, yt, ylo, fyt);
            hx = need_x && direct_face<false>(gamma, tol, hl, xlo, fxl);
            queue_face<true>(D, hy, (long long)r * G.pitch + col, yt, ylo, lane, &nq);
            queue_face<false>(D, hx && own_x, (long long)r * G.pitch + col, hl, xlo, lane, &nq);
            ey = ex = ERR_NONE;
        } else {
            ey = yflux<SOLVER>(gamma, tol, yt, ylo, fyt);
            ex = flux<SOLVER, 4>(gamma, tol, hl, xlo, fxl);
        }
        if (__ballot((ey | ex) != 0) != 0ull) {  // a failed face: rare, tested for the wavefront at once
            if (ey && out_col) { err = ey; err_cell = (long long)r * G.tx + col; }
            if (ex && (out_col || ghost_hi_col) && lane >= 1) { err = ex; err_cell = (long long)r * G.tx + col - 1; }
        }
        double fxr[4];
#pragma unroll
        for (int f = 0; f < 4; f++) fxr[f] = shfl_from_right(fxl[f]);

        bool deferred = false;
        if constexpr (DEFER) {
            const bool hxr = __shfl_down((int)hx, 1, LANES) != 0;
            int kind = DEF_INTERIOR;
            if (out_col) {
                deferred = hx || hxr || hyb || hy;
                if (deferred) {
                    // the cell's direct fluxes, for k_update_deferred
                    if (!hx) store_flux(D.fx, k, fxl);
                    if (!hxr) store_flux(D.fx, k + 1, fxr);
                    if (!hyb) store_flux(D.fy, k - G.pitch, fyb);
                    if (!hy) store_flux(D.fy, k, fyt);
                }
            } else if (ghost_lo_col) {
                deferred = hxr;
                kind = DEF_GHOST_XLO;
            } else if (ghost_hi_col) {
                deferred = hx;
                kind = DEF_GHOST_XHI;
            }
            defer_cell(D, deferred, k, kind, lane);
        }
        double u[4];  // the updated state of an updated cell
        if (deferred) {
            // updated by k_update_deferred
        } else if (out_col) {
            if constexpr (!WMODE) {
                const double* us = urows + (r % (AHEAD + 1)) * 4 * LANES + lane;
#pragma unroll
                for (int f = 0; f < 4; f++) uk[f] = us[f * LANES];
            }
#pragma unroll
            for (int f = 0; f < 4; f++) {
                double x = uk[f];
                x += fxl[f] * dt * G.area_x;
                x -= fxr[f] * dt * G.area_x;
                x += fyb[f] * dt * G.area_y;
                x -= fyt[f] * dt * G.area_y;
                u[f] = x;
            }
            // source_term.c:213-240 on the updated state
            if (G.gravity != 0.0) gravity_source(dt, G.gravity, u[0], &u[2], &u[3]);
#pragma unroll
            for (int f = 0; f < 4; f++) V.f[f][k] = u[f];
            // CFL maxima of the new state (utils.c:75-92)
            double w[4];
            c2p_2d_rn(gamma, G.volume, rvol, vol_ok, u[0], u[1], u[2], u[3], w);
            const double a = sound_speed_rn(gamma, w[0], w[3]);
            const double sx_ = a + fabs(w[1]);
            const double sy_ = a + fabs(w[2]);
            if (sx_ > smx) smx = sx_;
            if (sy_ > smy) smy = sy_;
        }
        // the x ghost columns next to the grid's edges (none of a strip away
        // from them: the test is the wavefront's, not a lane's)
        if (edge_strip) {
            if (!deferred && ghost_lo_col) {
#pragma unroll
                for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] - fxr[f] * dt * G.area_x;
            } else if (!deferred && ghost_hi_col) {
#pragma unroll
                for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] + fxl[f] * dt * G.area_x;
            }
        }

#pragma unroll
        for (int f = 0; f < 4; f++) {
            wa[f] = (LIM == FIRST_ORDER) ? wc[f] : wb[f];
            wb[f] = wc[f];
            yt[f] = yhi[f];
            fyb[f] = fyt[f];
        }
        a_row = a_next;
        cxr_row = cxr_next;
        hyb = hy;
    }
    if constexpr (DEFER) {
        // the top ghost row's update waits for a queued y_max face
        defer_cell(D, y1 == g + G.ny && out_col && hyb, (long long)(g + G.ny) * G.pitch + col, DEF_GHOST_YHI, lane);
        if (lane == 0 && nq != 0ull) atomicAdd(&D.tcount[sy * G.nsx + sx], nq);
    }
    if (y1 == g + G.ny && out_col && !hyb) {
        // top ghost row receives the y_max face flux (1097-1100)
        const long long k = (long long)(g + G.ny) * G.pitch + col;
#pragma unroll
        for (int f = 0; f < 4; f++) V.f[f][k] = U.f[f][k] + fyb[f] * dt * G.area_y;
    }

    StripResult res;
    res.smx = smx;
    res.smy = smy;
    res.err = err;
    res.err_cell = err_cell;
    return res;
}

// Fold a strip's CFL maxima and failure into the next controller slot.
__device__ __forceinline__ void strip_finish(StripResult r, Ctrl* out, int lane)
{
    double smx = wave_max(r.smx);
    double smy = wave_max(r.smy);
    const int err = r.err;
    const long long err_cell = r.err_cell;
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

// Consecutive workgroup ids are dealt round-robin to the 8 XCDs; remap them
// so each XCD gets a contiguous band of tiles (neighbours share halos in L2).
__device__ __forceinline__ int xcd_remap(int b, int ntiles)
{
    const int xcd = b % 8, idx = b / 8;
    const int base = ntiles / 8, extra = ntiles % 8;
    return xcd * base + min(xcd, extra) + idx;
}

// c2p of every cell of U into W followed by the ghost fill, or with fill = 0
// without it (mh2d.hip)
void prim_from_u(const Grid2D& G, const Fields4& U, const Fields4& W, hipStream_t s, int fill = 1,
                 const Ctrl* ctrl = nullptr, int slot = 0);

}  // namespace hfv
