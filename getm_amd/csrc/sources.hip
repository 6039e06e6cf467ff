// Vertical (sinking/floating) transfer into source terms: first-order upstream
// advection by a vertical velocity, written as sources of the layers it moves
// material between, python/src/wrapper.F90:447-486 (c_vertical_advection_to_sources),
// called by pygetm's FABM coupling (python/pygetm/fabm.py:368) for the sinking and
// floating of biogeochemical tracers.  FP64, bit-identical to the reference.
// Arrays are the Fortran (nx, ny, nz) = C (nz, ny, nx) with k = 0 at the surface.
#include "common.h"
#include "internal.h"

namespace getm {

__global__ void __launch_bounds__(64) k_vadv_flag_clear(int* flag)
{
    if (threadIdx.x == 0) *flag = 0;
}

// Activity test (wrapper.F90:461-467): is w nonzero anywhere in the interior of
// any layer?  Lanes that find one store 1 (the same value, so the race is benign).
__global__ void __launch_bounds__(256) k_vadv_active(int nx, int ny, int nz, int halox, int haloy,
                                                     const double* __restrict__ w, int* __restrict__ flag)
{
    const int i = halox + blockIdx.x * blockDim.x + threadIdx.x, j = haloy + blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= nx - halox || j >= ny - haloy) return;
    const size_t n2 = (size_t)nx * ny, p = (size_t)j * nx + i;
    for (int k = 0; k < nz; k++)
        if (w[(size_t)k * n2 + p] != 0.0) {
            *flag = 1;
            return;
        }
}

// One lane per interior column, layers in increasing k as the reference's k loop:
// s_{k+1} carries the transfer from the pair (k, k+1) into that of (k+1, k+2).
__global__ void __launch_bounds__(256) k_vadv_sources(int nx, int ny, int nz, int halox, int haloy,
                                                      const int* __restrict__ mask, const double* __restrict__ c,
                                                      const double* __restrict__ w, const double* __restrict__ h,
                                                      double* __restrict__ s, const int* __restrict__ flag)
{
    if (*flag == 0) return;
    const int i = halox + blockIdx.x * blockDim.x + threadIdx.x, j = haloy + blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= nx - halox || j >= ny - haloy) return;
    const size_t n2 = (size_t)nx * ny, p = (size_t)j * nx + i;
    const double upward = -1.0;  // surface-to-bottom ordering
    double sk = s[p];
    for (int k = 0; k < nz - 1; k++) {
        const size_t p0 = (size_t)k * n2 + p, p1 = p0 + n2;
        double sk1 = s[p1];
        if (mask[p0] == 1 && mask[p1] == 1) {
            const double lw = upward * 0.5 * (w[p0] + w[p1]);
            const double flux = lw > 0.0 ? lw * c[p0] : lw * c[p1];
            sk = sk - flux / h[p0];
            sk1 = sk1 + flux / h[p1];
        }
        s[p0] = sk;  // level k is final: later pairs do not touch it (unchanged values are stored back as they were)
        sk = sk1;
    }
    s[(size_t)(nz - 1) * n2 + p] = sk;
}

}  // namespace getm

using namespace getm;

extern "C" int c_vertical_advection_to_sources(int nx, int ny, int nz, int halox, int haloy, const int* mask,
                                               const double* c, const double* w, const double* h, double* s,
                                               void* stream)
{
    if (nx < 1 || ny < 1 || nz < 1 || halox < 0 || haloy < 0) return GETM_ERR_ARG;
    if (!mask || !c || !w || !h || !s) return GETM_ERR_ARG;
    if (nx - 2 * halox < 1 || ny - 2 * haloy < 1) return GETM_OK;  // no interior point
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // the call's own activity flag, in stream-ordered memory: calls on different
    // streams never share one, and nothing outlives the call
    int* flag = nullptr;
    GETM_HIP_OK(hipMallocAsync((void**)&flag, sizeof(int), st));
    dim3 block(64, 4);
    dim3 grid((nx - 2 * halox + 63) / 64, (ny - 2 * haloy + 3) / 4);
    hipLaunchKernelGGL(k_vadv_flag_clear, dim3(1), dim3(64), 0, st, flag);
    hipLaunchKernelGGL(k_vadv_active, grid, block, 0, st, nx, ny, nz, halox, haloy, w, flag);
    hipLaunchKernelGGL(k_vadv_sources, grid, block, 0, st, nx, ny, nz, halox, haloy, mask, c, w, h, s, flag);
    const hipError_t launched = hipGetLastError();
    GETM_HIP_OK(hipFreeAsync(flag, st));
    GETM_HIP_OK(launched);
    return GETM_OK;
}
