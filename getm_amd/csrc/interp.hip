// Grid-to-grid averages of python/src/interpolation.F90 (grid_interp_x/_y/_z/_xy,
// bound at python/_pygetm.pyx:180-190): they move velocities and vertical
// velocities between the staggered grids, e.g. onto the advection grids of the
// momentum self-advection (python/pygetm/momentum.py:928-961).  FP64, same
// operation order, bit-identical.  Arrays are i-fastest with extents
// (nx, ny, nz); target points outside the written range keep their values.
//
// Launch shape: blockIdx.x walks rows (one (j, k) pair or (k, j) pair), blockIdx.y
// 256-wide chunks of i, so the row index needs no division per lane and a grid
// of ny * nz rows fits (gridDim.x allows 2^31 - 1).  Each kernel streams its
// source once and writes its target once: HBM-bound at 16-40 B per point.
#include "common.h"
#include "internal.h"

namespace getm {

__global__ void __launch_bounds__(256) k_interp_x(int nx, int ny, const double* __restrict__ s,
                                                  double* __restrict__ t, int off)
{
    const int i = blockIdx.y * 256 + threadIdx.x;
    if (i >= nx - 1) return;
    const size_t p = (size_t)blockIdx.x * nx + i;  // row = j + ny * k
    t[p + off] = 0.5 * (s[p] + s[p + 1]);
}

__global__ void __launch_bounds__(256) k_interp_y(int nx, int ny, const double* __restrict__ s,
                                                  double* __restrict__ t, int off)
{
    const int i = blockIdx.y * 256 + threadIdx.x;
    if (i >= nx) return;
    const unsigned r = blockIdx.x, k = r / (unsigned)(ny - 1), j = r - k * (unsigned)(ny - 1);
    const size_t p = ((size_t)k * ny + j) * nx + i;
    t[p + (size_t)off * nx] = 0.5 * (s[p] + s[p + nx]);
}

__global__ void __launch_bounds__(256) k_interp_z(int nx, int ny, const double* __restrict__ s,
                                                  double* __restrict__ t, int off)
{
    const int i = blockIdx.y * 256 + threadIdx.x;
    if (i >= nx) return;
    const size_t n2 = (size_t)nx * ny, p = (size_t)blockIdx.x * nx + i;  // row = j + ny * k
    t[p + (size_t)off * n2] = 0.5 * (s[p] + s[p + n2]);
}

__global__ void __launch_bounds__(256) k_interp_xy(int nx1, int ny1, int nx2, int ny2, const double* __restrict__ s,
                                                   double* __restrict__ t, int ioff, int joff)
{
    const int i = blockIdx.y * 256 + threadIdx.x;
    if (i >= nx1 - 1) return;
    const unsigned r = blockIdx.x, k = r / (unsigned)(ny1 - 1), j = r - k * (unsigned)(ny1 - 1);
    const size_t p = ((size_t)k * ny1 + j) * nx1 + i;
    const double v = 0.25 * (s[p] + s[p + 1] + s[p + nx1] + s[p + nx1 + 1]);
    t[((size_t)k * ny2 + j + joff) * nx2 + i + ioff] = v;
}

static bool rows_ok(long long rows) { return rows > 0 && rows <= 0x7fffffffLL; }

}  // namespace getm

using namespace getm;

// The reference loops write target(i + offset) from source(i), source(i + 1) in
// increasing i; the lanes here are independent, which equals that only when the
// two arrays do not overlap, so the same pointer is rejected.
extern "C" int grid_interp_x(int nx, int ny, int nz, const double* source, double* target, int ioffset,
                             void* stream)
{
    if (nx < 1 || ny < 1 || nz < 1 || !source || !target || source == target) return GETM_ERR_ARG;
    if (ioffset < 0 || ioffset > 1) return GETM_ERR_ARG;
    if (nx < 2 || !rows_ok((long long)ny * nz)) return nx < 2 ? GETM_OK : GETM_ERR_ARG;
    hipLaunchKernelGGL(k_interp_x, dim3(ny * nz, (nx - 1 + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), nx, ny, source, target, ioffset);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

extern "C" int grid_interp_y(int nx, int ny, int nz, const double* source, double* target, int joffset,
                             void* stream)
{
    if (nx < 1 || ny < 1 || nz < 1 || !source || !target || source == target) return GETM_ERR_ARG;
    if (joffset < 0 || joffset > 1) return GETM_ERR_ARG;
    if (ny < 2) return GETM_OK;
    if (!rows_ok((long long)(ny - 1) * nz)) return GETM_ERR_ARG;
    hipLaunchKernelGGL(k_interp_y, dim3((ny - 1) * nz, (nx + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), nx, ny, source, target, joffset);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

extern "C" int grid_interp_z(int nx, int ny, int nz1, int nz2, const double* source, double* target, int koffset,
                             void* stream)
{
    if (nx < 1 || ny < 1 || nz1 < 1 || nz2 < 1 || !source || !target || source == target) return GETM_ERR_ARG;
    // target level k + koffset for k = 1 .. nz1 - 1 (Fortran) must exist
    if (koffset < 0 || nz1 - 1 + koffset > nz2) return GETM_ERR_ARG;
    if (nz1 < 2) return GETM_OK;
    if (!rows_ok((long long)ny * (nz1 - 1))) return GETM_ERR_ARG;
    hipLaunchKernelGGL(k_interp_z, dim3(ny * (nz1 - 1), (nx + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), nx, ny, source, target, koffset);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

extern "C" int grid_interp_xy(int nx1, int ny1, int nx2, int ny2, int nz, const double* source, double* target,
                              int ioffset, int joffset, void* stream)
{
    if (nx1 < 1 || ny1 < 1 || nx2 < 1 || ny2 < 1 || nz < 1 || !source || !target || source == target)
        return GETM_ERR_ARG;
    if (ioffset < 0 || joffset < 0 || nx1 - 1 + ioffset > nx2 || ny1 - 1 + joffset > ny2) return GETM_ERR_ARG;
    if (nx1 < 2 || ny1 < 2) return GETM_OK;
    if (!rows_ok((long long)(ny1 - 1) * nz)) return GETM_ERR_ARG;
    hipLaunchKernelGGL(k_interp_xy, dim3((ny1 - 1) * nz, (nx1 - 1 + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), nx1, ny1, nx2, ny2, source, target, ioffset, joffset);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}
