// devhooks.hip -- device forms of the custom boundary hooks the tests use
// (their host forms are lid_hook, inflow_hook_2d and copy_hook_2d in
// tests/_cases.py), written the way a user of hfv_register_device_bc_2d
// would: the hook enqueues one kernel on the run's stream.  Rows are `pitch`
// doubles apart; `user` points at host memory read when the hook is called.
#include <hip/hip_runtime.h>

namespace {

// every ghost row above the top wall, corner columns included: density and
// pressure mirrored from the rows below, the lid's velocity
__global__ void k_lid(double* rho, double* u, double* v, double* p, int g, int nx, int ny, long long pitch,
                      double lid_u, double lid_v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= nx + 2 * g || j >= g) return;
    const long long dst = (long long)(g + ny + j) * pitch + i;
    const long long src = (long long)(g + ny - 1 - j) * pitch + i;
    rho[dst] = rho[src];
    p[dst] = p[src];
    u[dst] = lid_u;
    v[dst] = lid_v;
}

// a fixed state in the x_min ghost columns of the interior rows
__global__ void k_inflow(double* rho, double* u, double* v, double* p, int g, int ny, long long pitch, double r0,
                         double u0, double v0, double p0)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= ny || i >= g) return;
    const long long k = (long long)(g + j) * pitch + i;
    rho[k] = r0;
    u[k] = u0;
    v[k] = v0;
    p[k] = p0;
}

// the nearest interior layer copied outwards, on side 1 (x_max), 2 (y_min)
// or 3 (y_max)
__global__ void k_copy(double* rho, double* u, double* v, double* p, int g, int nx, int ny, long long pitch,
                       int side)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int layer = blockIdx.y;
    const int span = side == 1 ? ny : nx;
    if (a >= span || layer >= g) return;
    long long dst, src;
    if (side == 1) {
        dst = (long long)(g + a) * pitch + g + nx + layer;
        src = (long long)(g + a) * pitch + g + nx - 1;
    } else if (side == 2) {
        dst = (long long)(g - 1 - layer) * pitch + g + a;
        src = (long long)g * pitch + g + a;
    } else {
        dst = (long long)(g + ny + layer) * pitch + g + a;
        src = (long long)(g + ny - 1) * pitch + g + a;
    }
    rho[dst] = rho[src];
    u[dst] = u[src];
    v[dst] = v[src];
    p[dst] = p[src];
}

}  // namespace

extern "C" {

// user: double[2] = {lid_u, lid_v}
void hfv_test_lid_hook(double* rho, double* u, double* v, double* p, int g, int nx, int ny, long long pitch,
                       void* stream, void* user)
{
    const double* par = static_cast<const double*>(user);
    const dim3 grid((unsigned)((nx + 2 * g + 63) / 64), (unsigned)g);
    hipLaunchKernelGGL(k_lid, grid, dim3(64), 0, (hipStream_t)stream, rho, u, v, p, g, nx, ny, pitch, par[0],
                       par[1]);
}

// user: double[4] = the inflow state (density, velocity_x, velocity_y, pressure)
void hfv_test_inflow_hook(double* rho, double* u, double* v, double* p, int g, int nx, int ny, long long pitch,
                          void* stream, void* user)
{
    const double* s = static_cast<const double*>(user);
    const dim3 grid((unsigned)((ny + 63) / 64), (unsigned)g);
    hipLaunchKernelGGL(k_inflow, grid, dim3(64), 0, (hipStream_t)stream, rho, u, v, p, g, ny, pitch, s[0], s[1],
                       s[2], s[3]);
}

// user: int[1] = the side (1, 2 or 3)
void hfv_test_copy_hook(double* rho, double* u, double* v, double* p, int g, int nx, int ny, long long pitch,
                        void* stream, void* user)
{
    const int side = *static_cast<const int*>(user);
    const int span = side == 1 ? ny : nx;
    const dim3 grid((unsigned)((span + 63) / 64), (unsigned)g);
    hipLaunchKernelGGL(k_copy, grid, dim3(64), 0, (hipStream_t)stream, rho, u, v, p, g, nx, ny, pitch, side);
}

}  // extern "C"
