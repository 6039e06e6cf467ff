// Input producer of the transport step: vertical velocity at layer interfaces
// from the layer transports and the layer-thickness change,
// python/src/momentum.F90:191-216 (c_w_momentum_3d), called by
// Momentum.update_diagnostics (python/pygetm/momentum.py:879-883) before the
// tracers are advected with it.  FP64, bit-identical to the reference: the
// expression keeps the Fortran's left-to-right order and idt = 1/dt once.
#include "common.h"
#include "internal.h"

namespace getm {

// One lane per (i, j) of [i0, i1] x [j0, j1] (0-based, inclusive); the interfaces
// are summed upward in the lane, so w_k reuses w_{k-1} from a register (the
// reference reads back the value it has just written).
__global__ void __launch_bounds__(256) k_w_momentum_3d(int nx, int ny, int nz, int i0, int i1, int j0, int j1,
                                                       const int* __restrict__ mask, const double* __restrict__ dyu,
                                                       const double* __restrict__ dxv,
                                                       const double* __restrict__ iarea,
                                                       const double* __restrict__ ho, const double* __restrict__ hn,
                                                       const double* __restrict__ pk, const double* __restrict__ qk,
                                                       double idt, double* __restrict__ w)
{
    const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x, j = j0 + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > i1 || j > j1) return;
    const size_t n2 = (size_t)nx * ny, p = (size_t)j * nx + i;
    if (mask[p] != 1) return;
    const double my_dyu = dyu[p], w_dyu = dyu[p - 1], my_dxv = dxv[p], s_dxv = dxv[p - nx], ia = iarea[p];
    double wk = w[p];
    for (int k = 1; k <= nz - 1; k++) {
        const size_t L = (size_t)(k - 1) * n2 + p;  // layer below interface k
        wk = wk - (hn[L] - ho[L]) * idt -
             (pk[L] * my_dyu - pk[L - 1] * w_dyu + qk[L] * my_dxv - qk[L - nx] * s_dxv) * ia;
        w[(size_t)k * n2 + p] = wk;
    }
}

// uk = pk / hn over every point, numpy's divide in Momentum.update_diagnostics
// (python/pygetm/momentum.py:885-892); IEEE division, so the same bits as numpy's,
// zero thicknesses giving inf or NaN alike.
__global__ void __launch_bounds__(256) k_divide(long long n, const double* __restrict__ a,
                                                const double* __restrict__ b, double* __restrict__ out)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = a[p] / b[p];
}

// un = (un * hn - Uo) / dt as python/src/momentum.F90:323-340 computes it (idt
// = 1/dt once, then a product), over all i and rows j0 .. j0 + nj - 1 of every
// layer; blockIdx.x = (layer, row) pair, blockIdx.y a 256-wide chunk of i.
__global__ void __launch_bounds__(256) k_reconstruct_transport_change(int nx, int ny, int j0, int nj,
                                                                      double* __restrict__ un,
                                                                      const double* __restrict__ hn,
                                                                      const double* __restrict__ Uo, double idt)
{
    const int i = blockIdx.y * 256 + threadIdx.x;
    if (i >= nx) return;
    const unsigned r = blockIdx.x, k = r / (unsigned)nj, j = j0 + (r - k * (unsigned)nj);
    const size_t p = ((size_t)k * ny + j) * nx + i;
    un[p] = (un[p] * hn[p] - Uo[p]) * idt;
}

}  // namespace getm

using namespace getm;

extern "C" int c_w_momentum_3d(int nx, int ny, int nz, int imin, int imax, int jmin, int jmax, const int* mask,
                               const double* dyu, const double* dxv, const double* iarea, const double* ho,
                               const double* hn, const double* pk, const double* qk, double timestep, double* w,
                               void* stream)
{
    // imin..jmax are the Fortran (1-based) bounds the reference wrapper passes
    const int i0 = imin - 1, i1 = imax - 1, j0 = jmin - 1, j1 = jmax - 1;
    if (nx < 1 || ny < 1 || nz < 1 || i0 < 1 || j0 < 1 || i1 >= nx || j1 >= ny || i1 < i0 || j1 < j0)
        return GETM_ERR_ARG;
    if (!mask || !dyu || !dxv || !iarea || !ho || !hn || !pk || !qk || !w) return GETM_ERR_ARG;
    if (nz < 2) return GETM_OK;  // no interior interface
    const double idt = 1.0 / timestep;
    dim3 block(64, 4);
    dim3 grid((i1 - i0 + 1 + 63) / 64, (j1 - j0 + 1 + 3) / 4);
    hipLaunchKernelGGL(k_w_momentum_3d, grid, block, 0, reinterpret_cast<hipStream_t>(stream), nx, ny, nz, i0, i1,
                       j0, j1, mask, dyu, dxv, iarea, ho, hn, pk, qk, idt, w);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

extern "C" int c_reconstruct_transport_change(int nx, int ny, int nz, int imin, int imax, int jmin, int jmax,
                                              double* un, const double* hn, const double* Uo, double timestep,
                                              void* stream)
{
    // the reference takes imin/imax but loops i over 1 .. nx (momentum.F90:335); so does this
    (void)imin;
    (void)imax;
    const int j0 = jmin - 1, j1 = jmax - 1;
    if (nx < 1 || ny < 1 || nz < 1 || j0 < 0 || j1 >= ny || !un || !hn || !Uo) return GETM_ERR_ARG;
    if (j1 < j0) return GETM_OK;
    const long long rows = (long long)nz * (j1 - j0 + 1);
    if (rows > 0x7fffffffLL) return GETM_ERR_ARG;
    const double idt = 1.0 / timestep;
    hipLaunchKernelGGL(k_reconstruct_transport_change, dim3((unsigned)rows, (nx + 255) / 256), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), nx, ny, j0, j1 - j0 + 1, un, hn, Uo, idt);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

extern "C" int c_velocities_from_transports(long long n, const double* pk, const double* hn, double* uk, void* stream)
{
    if (n < 0 || (n > 0 && (!pk || !hn || !uk))) return GETM_ERR_ARG;
    if (n == 0) return GETM_OK;
    const long long nblk = (n + 255) / 256;
    if (nblk > 0x7fffffffLL) return GETM_ERR_ARG;
    hipLaunchKernelGGL(k_divide, dim3((unsigned)nblk), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), n, pk, hn,
                       uk);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}
