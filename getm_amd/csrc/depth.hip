// Depth update before the tracer step (Domain.update_depth,
// python/pygetm/domain.py / simulation.py:1582-1600): water depth from elevation
// and the vertical coordinates of the new layer thicknesses, the reference's
// c_elevation2depth and c_thickness2vertical_coordinates (python/src/wrapper.F90),
// FP64 and bit-identical (flang's MAX is the compare-select a > b ? a : b).
#include "common.h"
#include "internal.h"

namespace getm {

__global__ void __launch_bounds__(256) k_elevation2depth(int n, const double* __restrict__ z,
                                                         const double* __restrict__ H, double Dmin,
                                                         const int* __restrict__ mask, double* __restrict__ D)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || mask[p] == 0) return;
    D[p] = fmax2(H[p] + z[p], Dmin);
}

// One lane per (i, j): the reference's per-k whole-array statements are
// independent between points, so each lane runs its column's recurrences.
__global__ void __launch_bounds__(256) k_thickness2vc(int nx, int ny, int nz, const int* __restrict__ mask,
                                                      const double* __restrict__ H, const double* __restrict__ h,
                                                      double* __restrict__ zc, double* __restrict__ zf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const size_t n2 = (size_t)nx * ny, p = (size_t)j * nx + i;
    if (mask[p] == 0) return;
    const double mH = -H[p];
    double hk = h[p];
    double c = mH + 0.5 * hk;
    zc[p] = c;
    double f = mH;
    zf[p] = f;
    for (int k = 1; k < nz; k++) {
        const double hk1 = h[(size_t)k * n2 + p];
        c = c + 0.5 * (hk + hk1);
        zc[(size_t)k * n2 + p] = c;
        f = f + hk;
        zf[(size_t)k * n2 + p] = f;
        hk = hk1;
    }
    zf[(size_t)nz * n2 + p] = f + hk;
}

}  // namespace getm

using namespace getm;

extern "C" int c_elevation2depth(int n, const double* z, const double* H, double Dmin, const int* mask, double* D,
                                 void* stream)
{
    if (n < 0 || (n > 0 && (!z || !H || !mask || !D))) return GETM_ERR_ARG;
    if (n == 0) return GETM_OK;
    hipLaunchKernelGGL(k_elevation2depth, dim3((n + 255) / 256), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       n, z, H, Dmin, mask, D);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

extern "C" int c_thickness2vertical_coordinates(int nx, int ny, int nz, const int* mask, const double* bottom_depth,
                                                const double* h, double* zc, double* zf, void* stream)
{
    if (nx < 1 || ny < 1 || nz < 1 || !mask || !bottom_depth || !h || !zc || !zf) return GETM_ERR_ARG;
    dim3 block(64, 4);
    dim3 grid((nx + 63) / 64, (ny + 3) / 4);
    hipLaunchKernelGGL(k_thickness2vc, grid, block, 0, reinterpret_cast<hipStream_t>(stream), nx, ny, nz, mask,
                       bottom_depth, h, zc, zf);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}
