// hfv_ctrl.hpp -- device-resident time control shared by the step kernels.
//
// The reference's time loop (integrator_muscl_hancock.c:530-550, 1143-1144)
// computes dt from the CFL reduction, clamps it to tf, checks dt == 0 and then
// advances t.  Here that logic runs inside every step launch, so steps can be
// queued back to back (and captured in a hipGraph) with no host round trip.
//
// Three controller slots rotate: launch k reads slot k%3, publishes the next
// state into slot (k+1)%3 (block 0 writes t/dt/count, every block adds its
// CFL maxima and any error with atomics) and clears slot (k+2)%3 for launch
// k+1.  A launch whose input says "finished" or "failed" only forwards the
// state.
#pragma once

#include <hip/hip_runtime.h>

#include "hfv_math.hpp"

namespace hfv {

struct alignas(64) Ctrl {
    double t;                  // simulation time at the start of the next step
    double dt;                 // dt of the last executed step
    double t_prev;             // time before the last executed step
    long long num_steps;       // executed steps
    unsigned long long key[2]; // CFL maxima of the current state (smax_key encoding)
    int cur;                   // which conserved buffer holds the current state
    int err_kind;              // ErrKind of the first failure (0: none)
    long long err_cell;        // flat index of a failing cell
};
static_assert(sizeof(Ctrl) == 64, "Ctrl must be one 64-byte line");

struct RunParams {
    double cfl;
    double shrink_factor;
    int shrink_steps;
    int dim;          // 1 or 2: which CFL formula (utils.c:52 / :94)
    double tf;
    long long max_steps;  // < 0: unlimited
    double fixed_dt;      // > 0: constant dt instead of the CFL dt
    double dx, dy;
    // Pause for a snapshot: no step starts once a step after step resume_n
    // has ended at t >= t_pause (the storing test of
    // integrator_muscl_hancock.c:1147-1155, evaluated after every step).
    double t_pause = __builtin_inf();
    long long resume_n = 0;
};

struct StepPlan {
    bool run;   // execute this step
    double dt;
    double t;
    long long n;
    int cur;
};

// Decide whether this launch executes a step and with which dt.  Every block
// evaluates the same expression on the same inputs, so all agree bitwise.
__device__ __forceinline__ StepPlan plan_step(const Ctrl* in, const RunParams& rp, int* dt_zero)
{
    StepPlan s;
    s.t = in->t;
    s.n = in->num_steps;
    s.cur = in->cur;
    s.dt = 0.0;
    *dt_zero = 0;
    const bool finished = (in->err_kind != ERR_NONE) || !(s.t < rp.tf) ||
                          (rp.max_steps >= 0 && s.n >= rp.max_steps) || (s.n > rp.resume_n && s.t >= rp.t_pause);
    if (finished) {
        s.run = false;
        return s;
    }
    double dt;
    if (rp.fixed_dt > 0.0) {
        dt = rp.fixed_dt;
    } else {
        const double cfl = (s.n < rp.shrink_steps) ? rp.cfl * rp.shrink_factor : rp.cfl;
        const double sx = smax_from_key(in->key[0]);
        if (rp.dim == 2) {
            const double sy = smax_from_key(in->key[1]);
            dt = cfl / (sx / rp.dx + sy / rp.dy);  // utils.c:94
        } else {
            dt = cfl * rp.dx / sx;                 // utils.c:52
        }
    }
    if (s.t + dt > rp.tf) dt = rp.tf - s.t;   // integrator_muscl_hancock.c:541-544
    if (dt == 0.0) {
        *dt_zero = 1;
        s.run = false;
        return s;
    }
    s.dt = dt;
    s.run = true;
    return s;
}

// Block 0 / thread 0 publishes the next controller state.
__device__ __forceinline__ void publish(const Ctrl* in, Ctrl* out, Ctrl* clear, const StepPlan& s, int dt_zero)
{
    clear->key[0] = 0ull;
    clear->key[1] = 0ull;
    clear->err_kind = ERR_NONE;
    clear->err_cell = -1;
    if (!s.run) {
        out->t = in->t;
        out->dt = in->dt;
        out->t_prev = in->t_prev;
        out->num_steps = in->num_steps;
        out->key[0] = in->key[0];
        out->key[1] = in->key[1];
        out->cur = in->cur;
        out->err_kind = dt_zero ? ERR_DT_ZERO : in->err_kind;
        out->err_cell = in->err_cell;
        return;
    }
    out->t = s.t + s.dt;
    out->dt = s.dt;
    out->t_prev = s.t;
    out->num_steps = s.n + 1;
    out->cur = s.cur ^ 1;
    // key[] and err_kind of *out were cleared by the previous launch and are
    // filled by this launch's atomics.
}

// wave64 reductions ---------------------------------------------------------
__device__ __forceinline__ double wave_max(double v);

// Workgroup maxima of (sx, sy), folded into c's CFL keys by one atomic per key
// (THREADS a multiple of 64; every thread of the workgroup calls this).
template <int THREADS>
__device__ __forceinline__ void block_smax_to(Ctrl* c, double sx, double sy)
{
    sx = wave_max(sx);
    sy = wave_max(sy);
    __shared__ double red[2][THREADS / 64];
    const int w = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0) {
        red[0][w] = sx;
        red[1][w] = sy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < THREADS / 64; q++) {
            if (red[0][q] > sx) sx = red[0][q];
            if (red[1][q] > sy) sy = red[1][q];
        }
        const unsigned long long kx = smax_key(sx), ky = smax_key(sy);
        if (kx) atomicMax(&c->key[0], kx);
        if (ky) atomicMax(&c->key[1], ky);
    }
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v) ? o : v;
    }
    return v;
}

__device__ __forceinline__ void report_error(Ctrl* out, int kind, long long cell)
{
    if (atomicCAS(&out->err_kind, 0, kind) == 0) {
        atomicExch((unsigned long long*)&out->err_cell, (unsigned long long)cell);
    }
}

}  // namespace hfv
