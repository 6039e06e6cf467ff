// Depth update before the tracer step (Simulation.update_depth,
// python/pygetm/simulation.py:1498-1600): water depth from elevation, the
// shallow-water damping factor, the layer thicknesses of the vertical coordinates
// and the depths of layer centres and interfaces -- the reference's
// c_elevation2depth, c_alpha, c_update_gvc and c_thickness2vertical_coordinates
// (python/src/wrapper.F90) and the Sigma product of vertical_coordinates.py.
// FP64 and bit-identical (flang's MAX and MIN are the compare-selects
// a > b ? a : b and a < b ? a : b, as fmax2/fmin2).  All are HBM-bound streams.
#include "common.h"
#include "internal.h"

namespace getm {

// z and D may be the same array (update_depth turns its half-step elevation into
// a depth in place), so neither is declared restrict.
__global__ void __launch_bounds__(256) k_elevation2depth(int n, const double* z, const double* __restrict__ H,
                                                         double Dmin, const int* __restrict__ mask, double* D)
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

__global__ void __launch_bounds__(256) k_alpha(int n, const double* __restrict__ D, double Dmin, double Dcrit,
                                               const int* __restrict__ mask, double* __restrict__ alpha)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || mask[p] == 0) return;
    alpha[p] = fmax2(0.0, fmin2(1.0, (D[p] - Dmin) / (Dcrit - Dmin)));
}

// One lane per point; the nz coefficients are the same for every lane.
__global__ void __launch_bounds__(256) k_update_gvc(int n2, int nz, double dsigma, const double* __restrict__ dbeta,
                                                    double Dgamma, int kref, const double* __restrict__ D,
                                                    const int* __restrict__ mask, double* __restrict__ h)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n2 || mask[p] == 0) return;
    const double d = D[p], bref = dbeta[kref];
    const double alpha = (fmin2(Dgamma / d, 1.0) * dsigma - bref) / (dsigma - bref);
    for (int k = 0; k < nz; k++) h[(size_t)k * n2 + p] = d * (dbeta[k] + alpha * (dsigma - dbeta[k]));
}

__global__ void __launch_bounds__(256) k_update_sigma(int n2, int nz, const double* __restrict__ dga,
                                                      const double* __restrict__ D, const int* __restrict__ mask,
                                                      double* __restrict__ h)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n2 || mask[p] == 0) return;
    const double d = D[p];
    for (int k = 0; k < nz; k++) h[(size_t)k * n2 + p] = dga[k] * d;
}

}  // namespace getm

using namespace getm;

extern "C" int c_alpha(int n, const double* D, double Dmin, double Dcrit, const int* mask, double* alpha,
                       void* stream)
{
    if (n < 0 || (n > 0 && (!D || !mask || !alpha))) return GETM_ERR_ARG;
    if (n == 0) return GETM_OK;
    hipLaunchKernelGGL(k_alpha, dim3((n + 255) / 256), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), n, D,
                       Dmin, Dcrit, mask, alpha);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

extern "C" int c_update_gvc(int nx, int ny, int nz, double dsigma, const double* dbeta, double Dgamma, int kk,
                            const double* D, const int* mask, double* h, void* stream)
{
    if (nx < 1 || ny < 1 || nz < 1 || kk < 1 || kk > nz || !dbeta || !D || !mask || !h) return GETM_ERR_ARG;
    const long long n2 = (long long)nx * ny;
    if (n2 > 0x7fffffffLL) return GETM_ERR_ARG;
    hipLaunchKernelGGL(k_update_gvc, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), (int)n2, nz, dsigma, dbeta, Dgamma, kk - 1, D, mask, h);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

extern "C" int c_update_sigma(int nx, int ny, int nz, const double* dga, const double* D, const int* mask, double* h,
                              void* stream)
{
    if (nx < 1 || ny < 1 || nz < 1 || !dga || !D || !mask || !h) return GETM_ERR_ARG;
    const long long n2 = (long long)nx * ny;
    if (n2 > 0x7fffffffLL) return GETM_ERR_ARG;
    hipLaunchKernelGGL(k_update_sigma, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), (int)n2, nz, dga, D, mask, h);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

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
