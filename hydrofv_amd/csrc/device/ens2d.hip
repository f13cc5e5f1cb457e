// ens2d.hip -- batched ensemble of independent 2D MUSCL-Hancock problems.
//
// Hundreds of grids of uneven size advance together: one launch per step
// covers every strip of every instance.  A tile table lists (instance, strip)
// pairs, largest instances first, and workgroups are dealt to it XCD-aware.
// Each instance keeps its own time control (three controller slots, see
// hfv_ctrl.hpp), so every instance takes its own CFL dt exactly as a
// single-grid run would; the block holding an instance's first strip
// publishes that instance's next state.  The per-strip work is the single-grid
// code (mh2d_strip.hpp), so each instance's result is bitwise the one it
// would get alone.  A step is captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "hfv_ctrl.hpp"
#include "hfv_gpu_internal.hpp"
#include "hfv_math.hpp"
#include "mh2d_strip.hpp"

namespace hfv {

struct EnsInstance {
    Grid2D G;
    RunParams rp;
    Fields4 Ua, Ub, W;
    Ctrl* ctrl;  // 3 slots
};

template <int LIM, int SOLVER, bool WMODE, bool BRANCHY>
__global__ void __launch_bounds__(LANES) __attribute__((amdgpu_waves_per_eu(SOLVER == 2 ? 2 : 1))) k_ens_step(const EnsInstance* __restrict__ insts, const int2* __restrict__ tiles,
                                                    int ntiles, int slot)
{
    const int b = xcd_remap(blockIdx.x, ntiles);
    const int2 t = tiles[b];
    const EnsInstance& E = insts[t.x];
    const Grid2D G = E.G;
    const Ctrl* in = E.ctrl + slot;
    Ctrl* out = E.ctrl + (slot + 1) % 3;
    Ctrl* clear = E.ctrl + (slot + 2) % 3;
    int dt_zero;
    const StepPlan plan = plan_step(in, E.rp, &dt_zero);
    const int lane = threadIdx.x;
    if (t.y == 0 && lane == 0) publish(in, out, clear, plan, dt_zero);
    if (!plan.run) return;
    const Fields4 U = plan.cur ? E.Ub : E.Ua;
    const Fields4 V = plan.cur ? E.Ua : E.Ub;
    const StripResult r = strip_step<LIM, SOLVER, WMODE, BRANCHY>(G, plan.dt, U, V, E.W, t.y % G.nsx, t.y / G.nsx,
                                                                     lane);
    strip_finish(r, out, lane);
}

// per-instance controller reset (three slots each)
__global__ void k_ens_reset(const EnsInstance* __restrict__ insts, int count)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * count) return;
    Ctrl& c = insts[k / 3].ctrl[k % 3];
    c.t = 0.0;
    c.dt = 0.0;
    c.t_prev = 0.0;
    c.num_steps = 0;
    c.key[0] = 0ull;
    c.key[1] = 0ull;
    c.cur = 0;
    c.err_kind = ERR_NONE;
    c.err_cell = -1;
}

// first-dt CFL maxima from every instance's W (utils.c:55-95): workgroup
// (instance, b) strides over rows b, b + ENS_SMAX_ROW_BLOCKS, ... of that
// instance and adds one atomic per key
constexpr int ENS_SMAX_THREADS = 256;
constexpr int ENS_SMAX_ROW_BLOCKS = 16;

__global__ void __launch_bounds__(ENS_SMAX_THREADS) k_ens_smax(const EnsInstance* __restrict__ insts, int count)
{
    const EnsInstance& E = insts[blockIdx.x];
    const Grid2D& G = E.G;
    double sx = -INFINITY, sy = -INFINITY;
    for (int j = blockIdx.y; j < G.ny; j += gridDim.y) {
        const long long row = (long long)(j + G.g) * G.pitch + G.g;
        for (int i = threadIdx.x; i < G.nx; i += ENS_SMAX_THREADS) {
            const long long q = row + i;
            const double a = sound_speed(G.gamma, E.W.f[0][q], E.W.f[3][q]);
            const double vx = a + fabs(E.W.f[1][q]);
            const double vy = a + fabs(E.W.f[2][q]);
            if (vx > sx) sx = vx;
            if (vy > sy) sy = vy;
        }
    }
    block_smax_to<ENS_SMAX_THREADS>(E.ctrl, sx, sy);
}

}  // namespace hfv

using namespace hfv;

struct hfv_ens {
    int count;
    int limiter, solver, form;
    std::vector<hfv_mh2d_config> cfg;
    std::vector<EnsInstance> host;  // descriptors (device pointers inside)
    EnsInstance* dinst;
    int2* dtiles;
    int ntiles;
    double* slab;
    size_t slab_bytes;
    Ctrl* ctrl;
    long long launches;
    double* saved;           // hfv_ens_checkpoint's copy of the slab (allocated at first use)
    long long saved_launches;
    hipGraphExec_t graph;
    int graph_phase;
    hipStream_t graph_stream;
    hipStream_t own;  // used when the caller passes no stream
};

namespace {

void free_tiles(hfv_ens* e)
{
    if (e->dtiles) (void)hipFree(e->dtiles);
    e->dtiles = nullptr;
}

int build_tiles(hfv_ens* e)
{
    // largest instances first so the long strips start early
    std::vector<int> order(e->count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (long long)e->host[a].G.nx * e->host[a].G.ny > (long long)e->host[b].G.nx * e->host[b].G.ny;
    });
    std::vector<int2> tiles;
    for (int i : order) {
        const Grid2D& G = e->host[i].G;
        for (int t = 0; t < G.nsx * G.nsy; t++) tiles.push_back(make_int2(i, t));
    }
    // no launch may still read the old tables
    if (hipDeviceSynchronize() != hipSuccess) return HFV_EDEVICE;
    free_tiles(e);
    e->ntiles = (int)tiles.size();
    if (hipMalloc(&e->dtiles, tiles.size() * sizeof(int2)) != hipSuccess) return HFV_ENOMEM;
    if (hipMemcpy(e->dtiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
        return HFV_EDEVICE;
    if (hipMemcpy(e->dinst, e->host.data(), e->host.size() * sizeof(EnsInstance), hipMemcpyHostToDevice) !=
        hipSuccess)
        return HFV_EDEVICE;
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
    return HFV_OK;
}

template <int LIM, int SOLVER, bool BRANCHY>
void ens_launch_t(hfv_ens* e, bool wmode, hipStream_t s)
{
    const int slot = (int)(e->launches % 3);
    if (wmode)
        hipLaunchKernelGGL((k_ens_step<LIM, SOLVER, true, BRANCHY>), dim3(e->ntiles), dim3(LANES), 0, s, e->dinst,
                           e->dtiles, e->ntiles, slot);
    else
        hipLaunchKernelGGL((k_ens_step<LIM, SOLVER, false, BRANCHY>), dim3(e->ntiles), dim3(LANES), 0, s, e->dinst,
                           e->dtiles, e->ntiles, slot);
    e->launches++;
}

template <int LIM>
void ens_launch_hllc(hfv_ens* e, bool wmode, hipStream_t s)
{
    if (e->form == HFV_HLLC_BRANCHES) ens_launch_t<LIM, HLLC, true>(e, wmode, s);
    else ens_launch_t<LIM, HLLC, false>(e, wmode, s);
}

void ens_launch(hfv_ens* e, bool wmode, hipStream_t s)
{
    if (e->solver == HLLC) {
        if (e->limiter == MINMOD) ens_launch_hllc<MINMOD>(e, wmode, s);
        else if (e->limiter == VAN_LEER) ens_launch_hllc<VAN_LEER>(e, wmode, s);
        else ens_launch_hllc<MC>(e, wmode, s);
    } else {
        if (e->limiter == MINMOD) ens_launch_t<MINMOD, EXACT, false>(e, wmode, s);
        else if (e->limiter == VAN_LEER) ens_launch_t<VAN_LEER, EXACT, false>(e, wmode, s);
        else ens_launch_t<MC, EXACT, false>(e, wmode, s);
    }
}

}  // namespace

extern "C" {

int hfv_ens_create(int count, const hfv_mh2d_config* cfgs, hfv_ens** out)
{
    *out = nullptr;
    if (count <= 0) return HFV_EINVAL;
    for (int i = 0; i < count; i++) {
        const hfv_mh2d_config& c = cfgs[i];
        if (c.num_cells_x < c.num_ghost_cells_side || c.num_cells_y < c.num_ghost_cells_side ||
            c.num_ghost_cells_side < 2)
            return HFV_EINVAL;
        if (c.slope_limiter != cfgs[0].slope_limiter || c.riemann_solver != cfgs[0].riemann_solver) return HFV_EINVAL;
        if (c.slope_limiter < 1 || c.slope_limiter > 3) return HFV_EINVAL;
        if (c.riemann_solver != HLLC && c.riemann_solver != EXACT) return HFV_EINVAL;
        if (c.hllc_form != HFV_HLLC_SELECTS && c.hllc_form != HFV_HLLC_BRANCHES) return HFV_EINVAL;
        if (c.first_order) return HFV_EINVAL;  // the batched engine runs MUSCL-Hancock only
        for (int s = 0; s < 4; s++)
            if (c.bc[s] < BC_NONE || c.bc[s] > BC_PERIODIC) return HFV_EINVAL;
    }
    hfv_ens* e = new (std::nothrow) hfv_ens();
    if (!e) return HFV_ENOMEM;
    e->count = count;
    e->limiter = cfgs[0].slope_limiter;
    e->solver = cfgs[0].riemann_solver;
    e->form = cfgs[0].hllc_form;
    e->cfg.assign(cfgs, cfgs + count);
    e->host.resize(count);
    size_t total = 0;
    std::vector<size_t> offs(count);
    for (int i = 0; i < count; i++) {
        const hfv_mh2d_config& c = cfgs[i];
        Grid2D& G = e->host[i].G;
        G.nx = c.num_cells_x;
        G.ny = c.num_cells_y;
        G.g = c.num_ghost_cells_side;
        G.tx = G.nx + 2 * G.g;
        G.ty = G.ny + 2 * G.g;
        G.pitch = ((long long)G.tx + 31) / 32 * 32;
        G.gamma = c.gamma;
        G.gravity = c.gravity;
        G.tol = c.tol;
        G.area_x = c.area_x;
        G.area_y = c.area_y;
        G.volume = c.volume;
        for (int s = 0; s < 4; s++) G.bc[s] = c.bc[s];
        G.strip_rows = c.strip_rows > 0 ? c.strip_rows : 32;
        G.full_tiles = c.full_tiles != 0;
        G.nsx = (G.nx + STRIP_COLS - 1) / STRIP_COLS;
        G.nsy = (G.ny + G.strip_rows - 1) / G.strip_rows;
        RunParams& rp = e->host[i].rp;
        rp.cfl = c.cfl;
        rp.shrink_factor = c.cfl_initial_shrink_factor;
        rp.shrink_steps = c.cfl_initial_shrink_num_steps;
        rp.dim = 2;
        rp.tf = c.tf;
        rp.max_steps = c.max_steps;
        rp.fixed_dt = c.fixed_dt;
        rp.dx = c.dx;
        rp.dy = c.dy;
        offs[i] = total;
        total += (size_t)G.pitch * G.ty * 12;
    }
    const size_t bytes = total * sizeof(double) + (size_t)count * 3 * sizeof(Ctrl) + 256;
    e->slab_bytes = bytes;
    e->saved = nullptr;
    e->saved_launches = -1;
    if (hipMalloc(&e->slab, bytes) != hipSuccess) {
        delete e;
        return HFV_ENOMEM;
    }
    e->ctrl = reinterpret_cast<Ctrl*>(e->slab + total);
    for (int i = 0; i < count; i++) {
        EnsInstance& I = e->host[i];
        const size_t field = (size_t)I.G.pitch * I.G.ty;
        for (int f = 0; f < 4; f++) {
            I.Ua.f[f] = e->slab + offs[i] + field * f;
            I.Ub.f[f] = e->slab + offs[i] + field * (4 + f);
            I.W.f[f] = e->slab + offs[i] + field * (8 + f);
        }
        I.ctrl = e->ctrl + 3 * i;
    }
    if (hipMalloc(&e->dinst, (size_t)count * sizeof(EnsInstance)) != hipSuccess) {
        (void)hipFree(e->slab);
        delete e;
        return HFV_ENOMEM;
    }
    e->dtiles = nullptr;
    e->graph = nullptr;
    e->launches = 0;
    if (hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(e->dinst);
        (void)hipFree(e->slab);
        delete e;
        return HFV_EDEVICE;
    }
    const int rc = build_tiles(e);
    if (rc != HFV_OK) {
        free_tiles(e);
        (void)hipStreamDestroy(e->own);
        (void)hipFree(e->dinst);
        (void)hipFree(e->slab);
        delete e;
        return rc;
    }
    *out = e;
    return HFV_OK;
}

void hfv_ens_destroy(hfv_ens* e)
{
    if (!e) return;
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    (void)hipStreamSynchronize(e->own);
    (void)hipStreamDestroy(e->own);
    free_tiles(e);
    if (e->saved) (void)hipFree(e->saved);
    (void)hipFree(e->dinst);
    (void)hipFree(e->slab);
    delete e;
}

// Save (restore = 0) or restore (1) the whole device state: every
// instance's fields and time control, and the step count that selects the
// controller slot.
int hfv_ens_checkpoint(hfv_ens* e, int restore, void* stream)
{
    hipStream_t s = stream ? (hipStream_t)stream : e->own;
    if (restore) {
        if (!e->saved || e->saved_launches < 0) return HFV_EINVAL;
        if (hipMemcpyAsync(e->slab, e->saved, e->slab_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return HFV_EDEVICE;
        e->launches = e->saved_launches;
        return HFV_OK;
    }
    if (!e->saved && hipMalloc(&e->saved, e->slab_bytes) != hipSuccess) return HFV_ENOMEM;
    if (hipMemcpyAsync(e->saved, e->slab, e->slab_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return HFV_EDEVICE;
    e->saved_launches = e->launches;
    return HFV_OK;
}

int hfv_ens_set_strip_rows(hfv_ens* e, const int* rows)
{
    for (int i = 0; i < e->count; i++) {
        if (rows[i] <= 0) return HFV_EINVAL;
        Grid2D& G = e->host[i].G;
        G.strip_rows = rows[i];
        G.nsy = (G.ny + G.strip_rows - 1) / G.strip_rows;
    }
    return build_tiles(e);
}

int hfv_ens_set_hllc_form(hfv_ens* e, int form)
{
    if (form != HFV_HLLC_SELECTS && form != HFV_HLLC_BRANCHES) return HFV_EINVAL;
    if (form != e->form) {
        e->form = form;
        if (e->graph) (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    return HFV_OK;
}

int hfv_ens_num_tiles(hfv_ens* e) { return e->ntiles; }

int hfv_ens_upload(hfv_ens* e, int inst, const double* const W[4], const double* const U[4], void* stream)
{
    if (inst < 0 || inst >= e->count) return HFV_EINVAL;
    hipStream_t s = stream ? (hipStream_t)stream : e->own;
    const EnsInstance& I = e->host[inst];
    const size_t hp = (size_t)I.G.tx * sizeof(double), dp = (size_t)I.G.pitch * sizeof(double);
    for (int f = 0; f < 4; f++) {
        if (hipMemcpy2DAsync(I.W.f[f], dp, W[f], hp, hp, I.G.ty, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpy2DAsync(I.Ua.f[f], dp, U[f], hp, hp, I.G.ty, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpy2DAsync(I.Ub.f[f], dp, U[f], hp, hp, I.G.ty, hipMemcpyHostToDevice, s) != hipSuccess)
            return HFV_EDEVICE;
    }
    return HFV_OK;
}

int hfv_ens_start(hfv_ens* e, void* stream)
{
    hipStream_t s = stream ? (hipStream_t)stream : e->own;
    e->launches = 0;
    hipLaunchKernelGGL(k_ens_reset, dim3((3 * e->count + 255) / 256), dim3(256), 0, s, e->dinst, e->count);
    hipLaunchKernelGGL(k_ens_smax, dim3(e->count, ENS_SMAX_ROW_BLOCKS), dim3(ENS_SMAX_THREADS), 0, s, e->dinst,
                       e->count);
    ens_launch(e, true, s);
    return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

int hfv_ens_advance(hfv_ens* e, long long launches, int use_graph, void* stream)
{
    hipStream_t s = stream ? (hipStream_t)stream : e->own;
    const int GRAPH_LEN = 48;
    while (launches > 0) {
        if (!use_graph || launches < GRAPH_LEN) {
            ens_launch(e, false, s);
            launches--;
            continue;
        }
        const int phase = (int)(e->launches % 3);
        if (!e->graph || e->graph_phase != phase || e->graph_stream != s) {
            if (e->graph) (void)hipGraphExecDestroy(e->graph);
            e->graph = nullptr;
            hipGraph_t graph;
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return HFV_EDEVICE;
            const long long saved = e->launches;
            for (int i = 0; i < GRAPH_LEN; i++) ens_launch(e, false, s);
            e->launches = saved;
            if (hipStreamEndCapture(s, &graph) != hipSuccess) return HFV_EDEVICE;
            if (hipGraphInstantiate(&e->graph, graph, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return HFV_EDEVICE;
            }
            (void)hipGraphDestroy(graph);
            e->graph_phase = phase;
            e->graph_stream = s;
        }
        if (hipGraphLaunch(e->graph, s) != hipSuccess) return HFV_EDEVICE;
        e->launches += GRAPH_LEN;
        launches -= GRAPH_LEN;
    }
    return hipGetLastError() == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

// Status of every instance (st[count]).
int hfv_ens_status(hfv_ens* e, hfv_run_status* st, void* stream)
{
    hipStream_t s = stream ? (hipStream_t)stream : e->own;
    std::vector<Ctrl> c((size_t)e->count * 3);
    if (hipMemcpyAsync(c.data(), e->ctrl, c.size() * sizeof(Ctrl), hipMemcpyDeviceToHost, s) != hipSuccess)
        return HFV_EDEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return HFV_EDEVICE;
    const int slot = (int)(e->launches % 3);
    for (int i = 0; i < e->count; i++) {
        const Ctrl& x = c[(size_t)3 * i + slot];
        const RunParams& rp = e->host[i].rp;
        st[i].t = x.t;
        st[i].dt = x.dt;
        st[i].t_prev = x.t_prev;
        st[i].num_steps = x.num_steps;
        st[i].err_kind = x.err_kind;
        st[i].err_cell = x.err_cell;
        st[i].cur = x.cur;
        st[i].smax[0] = smax_from_key(x.key[0]);
        st[i].smax[1] = smax_from_key(x.key[1]);
        st[i].finished = (x.err_kind != ERR_NONE || !(x.t < rp.tf) || (rp.max_steps >= 0 && x.num_steps >= rp.max_steps));
        st[i].bisect_fails = 0;
    }
    return HFV_OK;
}

// Conserved state of one instance (and its primitive state: c2p + fill).
int hfv_ens_download(hfv_ens* e, int inst, double* const W[4], double* const U[4], void* stream)
{
    if (inst < 0 || inst >= e->count) return HFV_EINVAL;
    hipStream_t s = stream ? (hipStream_t)stream : e->own;
    std::vector<hfv_run_status> st(e->count);
    int rc = hfv_ens_status(e, st.data(), stream);
    if (rc) return rc;
    const EnsInstance& I = e->host[inst];
    const Fields4 cur = st[inst].cur ? I.Ub : I.Ua;
    prim_from_u(I.G, cur, I.W, s);
    const size_t hp = (size_t)I.G.tx * sizeof(double), dp = (size_t)I.G.pitch * sizeof(double);
    for (int f = 0; f < 4; f++) {
        if (W && hipMemcpy2DAsync(W[f], hp, I.W.f[f], dp, hp, I.G.ty, hipMemcpyDeviceToHost, s) != hipSuccess)
            return HFV_EDEVICE;
        if (U && hipMemcpy2DAsync(U[f], hp, cur.f[f], dp, hp, I.G.ty, hipMemcpyDeviceToHost, s) != hipSuccess)
            return HFV_EDEVICE;
    }
    return hipStreamSynchronize(s) == hipSuccess ? HFV_OK : HFV_EDEVICE;
}

}  // extern "C"
