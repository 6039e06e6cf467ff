// Non-finite detection in the tracers, the transport path's part of the
// reference's Simulation.check_finite (python/pygetm/simulation.py:348-410):
// per tracer, the number of non-finite values over the interior cells whose
// T mask is 1 or 2, and the first such cell in memory order.  Device-side, one
// pass over the tracers; the result buffer is initialised by a kernel of its
// own (no memset), so the check may be captured in a graph with the step.
#include <stdint.h>

#include "common.h"
#include "internal.h"

namespace getm {

__global__ void __launch_bounds__(64) k_finite_init(int ntr, long long* out)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= ntr) return;
    out[2 * t] = 0;
    out[2 * t + 1] = 0x7fffffffffffffffLL;
}

// one lane per column of the interior, blockIdx.z = tracer
__global__ void __launch_bounds__(256) k_finite(int nx_, int ny_, int nz, int imin, int imax, int jmin, int jmax,
                                                const int* __restrict__ mask, const float* __restrict__ f32,
                                                const double* __restrict__ f64, long long* __restrict__ out)
{
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, t = blockIdx.z;
    if (i < imin || i > imax || j < jmin || j > jmax) return;
    const size_t n2 = (size_t)nx_ * ny_, p = (size_t)j * nx_ + i;
    const int m = mask[p];
    if (m != 1 && m != 2) return;
    const size_t base = (size_t)t * n2 * nz + p;
    long long bad = 0, first = 0x7fffffffffffffffLL;
    for (int k = 0; k < nz; k++) {
        const size_t q = base + (size_t)k * n2;
        const bool finite = f32 ? isfinite(f32[q]) : isfinite(f64[q]);
        if (!finite) {
            if (!bad) first = (long long)(q - (size_t)t * n2 * nz);
            bad++;
        }
    }
    if (bad) {
        atomicAdd((unsigned long long*)&out[2 * t], (unsigned long long)bad);
        atomicMin(&out[2 * t + 1], first);
    }
}

int launch_check_finite(const Grid& g, int ntr, const float* f32, const double* f64, long long* out, hipStream_t s)
{
    if (ntr < 1) return GETM_OK;
    if (ntr > 65535) return GETM_ERR_CAPACITY;
    hipLaunchKernelGGL(k_finite_init, dim3((ntr + 63) / 64), dim3(64), 0, s, ntr, out);
    dim3 grid((g.nx_ + 63) / 64, (g.ny_ + 3) / 4, ntr);
    hipLaunchKernelGGL(k_finite, grid, dim3(64, 4), 0, s, g.nx_, g.ny_, g.nz, g.imin, g.imax, g.jmin, g.jmax, g.mask,
                       f32, f64, out);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

}  // namespace getm
