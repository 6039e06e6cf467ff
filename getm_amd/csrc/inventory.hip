// Tracer inventories and the FP32 inventory fixer.
//
// Inventory of tracer t: sum over wet cells of area * h * c
// (the reference's domain integral, python/pygetm/simulation.py:1408-1466),
// accumulated in FP64 with Neumaier compensation in a fixed order (lane
// column sums, then a sequential block sum, then a sequential sum over
// blocks), so it is deterministic.
//
// Fixer (FP32 tracers): FP32 storage rounds every cell once per stage, which
// makes the inventory drift by ~1e-11 relative per step.  The fixer restores
// a target inventory M exactly up to one cell's rounding: the defect
// D = M - I is shared in proportion to each cell's mass, which in units of
// the cell's ulp is q_i = (D/I) * c_i / ulp(c_i); with S_i the prefix sum of
// c/ulp in a fixed cell order, cell i changes by
//   n_i = rint((D/I) S_i) - rint((D/I) S_{i-1})  ulps   (n_i in {-1, 0, 1} in practice),
// which is exact in FP32 and touches ~|D/I| * 2^23 cells per tracer.  Flipped
// cells carry different masses, so a residual of a few cell-ulps remains; the
// fix kernel accumulates the fixed inventory on the way, and a sequential
// greedy pass over the first wet cells of each tracer removes the residual to
// within half of one cell's ulp mass (relative ~1e-7 / wet cells).
#include <type_traits>

#include "common.h"
#include "internal.h"

namespace getm {

constexpr int IBX = 64, IBY = 4, INV_THREADS = IBX * IBY;

__device__ __forceinline__ void neumaier(double& s, double& c, double x)
{
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

// c / ulp(c) for an FP32 value (0 for zero and subnormals).
__device__ __forceinline__ double q_of(float c)
{
    const int e = ilogbf(c);
    if (c == 0.f || e < -126) return 0.0;
    return (double)c * exp2((double)(23 - e));
}

// Partial sums per (tracer, block): mass (hi, lo) and the sum of c/ulp(c) (FP32 only).
template <typename R>
__global__ void __launch_bounds__(INV_THREADS) k_inventory(const StepArgs A, const double* __restrict__ h,
                                                           const double* __restrict__ area, int ntr,
                                                           const R* __restrict__ f, double* __restrict__ part)
{
    __shared__ double s_m[INV_THREADS], s_c[INV_THREADS], s_q[INV_THREADS];
    const int tid = threadIdx.y * IBX + threadIdx.x;
    const int i = blockIdx.x * IBX + threadIdx.x, j = blockIdx.y * IBY + threadIdx.y;
    const int t = blockIdx.z;
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    double m = 0.0, mc = 0.0, q = 0.0;
    if (i < A.nx_ && j < A.ny_) {
        const size_t p = (size_t)j * A.nx_ + i;
        const bool wet = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax && A.tmask[p] == 1;
        if (wet) {
            const R* fc = f + (size_t)t * A.n3 + p;
            const double a = area[p];
            for (int k = 0; k < A.nz; k++) {
                const R c = fc[(size_t)k * A.n2];
                neumaier(m, mc, (a * h[(size_t)k * A.n2 + p]) * (double)c);
                if constexpr (std::is_same<R, float>::value) q += q_of(c);
            }
        }
    }
    s_m[tid] = m;
    s_c[tid] = mc;
    s_q[tid] = q;
    __syncthreads();
    if (tid == 0) {
        double S = 0.0, C = 0.0, Q = 0.0;
        for (int l = 0; l < INV_THREADS; l++) {
            neumaier(S, C, s_m[l]);
            C += s_c[l];
            Q += s_q[l];
        }
        double* pt = part + (size_t)t * 3 * nblk;
        pt[blk] = S;
        pt[nblk + blk] = C;
        pt[2 * nblk + blk] = Q;
    }
}

// One lane per tracer: inventory = compensated sum over blocks; for the fixer
// also the exclusive prefix of the per-block c/ulp sums and the ratio D / I.
__global__ void k_inventory_reduce(int ntr, int nblk, double* __restrict__ part, double* __restrict__ inv,
                                   const double* __restrict__ target, double* __restrict__ ratio)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntr) return;
    double* pt = part + (size_t)t * 3 * nblk;
    double S = 0.0, C = 0.0, Q = 0.0;
    for (int b = 0; b < nblk; b++) {
        neumaier(S, C, pt[b]);
        C += pt[nblk + b];
        const double qb = pt[2 * nblk + b];
        pt[2 * nblk + b] = Q;  // exclusive prefix, in place
        Q += qb;
    }
    const double I = S + C;
    if (inv) inv[t] = I;
    if (target && ratio) ratio[t] = (I != 0.0) ? (target[t] - I) / I : 0.0;
}

// Same block decomposition and summation order as k_inventory, so the block
// prefixes from k_inventory_reduce continue exactly into each block.
__global__ void __launch_bounds__(INV_THREADS) k_fix_f32(const StepArgs A, int ntr, float* __restrict__ f,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ ratio,
                                                         const double* __restrict__ area, double* __restrict__ part2)
{
    __shared__ double s_q[INV_THREADS], s_m[INV_THREADS], s_c[INV_THREADS];
    const int tid = threadIdx.y * IBX + threadIdx.x;
    const int i = blockIdx.x * IBX + threadIdx.x, j = blockIdx.y * IBY + threadIdx.y;
    const int t = blockIdx.z;
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    const double r = ratio[t];
    bool wet = false;
    size_t p = 0;
    double q = 0.0;
    float* fc = nullptr;
    if (i < A.nx_ && j < A.ny_) {
        p = (size_t)j * A.nx_ + i;
        wet = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax && A.tmask[p] == 1;
        if (wet) {
            fc = f + (size_t)t * A.n3 + p;
            for (int k = 0; k < A.nz; k++) q += q_of(fc[(size_t)k * A.n2]);
        }
    }
    s_q[tid] = q;
    __syncthreads();
    if (tid == 0) {
        double Q = part[(size_t)t * 3 * nblk + 2 * nblk + blk];
        for (int l = 0; l < INV_THREADS; l++) {
            const double ql = s_q[l];
            s_q[l] = Q;  // exclusive prefix of this lane
            Q += ql;
        }
    }
    __syncthreads();
    double m = 0.0, mc = 0.0;
    if (wet) {
        const double a = area[p];
        double S0 = s_q[tid];
        for (int k = 0; k < A.nz; k++) {
            float c = fc[(size_t)k * A.n2];
            const double S1 = S0 + q_of(c);
            const double n = rint(r * S1) - rint(r * S0);
            if (n != 0.0) {
                const int e = ilogbf(c);
                c = c + (float)(n * exp2((double)(e - 23)));
                fc[(size_t)k * A.n2] = c;
            }
            neumaier(m, mc, (a * A.hn[(size_t)k * A.n2 + p]) * (double)c);
            S0 = S1;
        }
    }
    s_m[tid] = m;
    s_c[tid] = mc;
    __syncthreads();
    if (tid == 0) {
        double S = 0.0, C = 0.0;
        for (int l = 0; l < INV_THREADS; l++) {
            neumaier(S, C, s_m[l]);
            C += s_c[l];
        }
        double* pt = part2 + (size_t)t * 2 * nblk;
        pt[blk] = S;
        pt[nblk + blk] = C;
    }
}

// One lane per tracer: remaining defect after the dithering pass.
__global__ void k_residual(int ntr, int nblk, const double* __restrict__ part2, const double* __restrict__ target,
                           double* __restrict__ resid)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntr) return;
    const double* pt = part2 + (size_t)t * 2 * nblk;
    double S = 0.0, C = 0.0;
    for (int b = 0; b < nblk; b++) {
        neumaier(S, C, pt[b]);
        C += pt[nblk + b];
    }
    // target - (S + C), compensated
    double R = 0.0, RC = 0.0;
    neumaier(R, RC, target[t]);
    neumaier(R, RC, -S);
    neumaier(R, RC, -C);
    resid[t] = R + RC;
}

// One lane per tracer: greedy one-ulp changes over the wet cells in memory
// order until the residual is below half of the next cell's ulp mass.
__global__ void k_greedy_f32(const StepArgs A, int ntr, float* __restrict__ f, const double* __restrict__ area,
                             const double* __restrict__ resid, const double* __restrict__ target, int max_cells)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntr) return;
    double R = resid[t];
    const double tol = 1e-14 * fabs(target[t]);  // two orders below the conservation target
    if (fabs(R) <= tol) return;
    float* ft = f + (size_t)t * A.n3;
    int visited = 0;
    for (int j = A.jmin; j <= A.jmax && visited < max_cells; j++) {
        for (int i = A.imin; i <= A.imax && visited < max_cells; i++) {
            const size_t p = (size_t)j * A.nx_ + i;
            if (A.tmask[p] != 1) continue;
            const double a = area[p];
            for (int k = 0; k < A.nz && visited < max_cells; k++, visited++) {
                const size_t pk = (size_t)k * A.n2 + p;
                const float c = ft[pk];
                const int e = ilogbf(c);
                if (c == 0.f || e < -126) continue;
                const double ulp = exp2((double)(e - 23));
                const double wm = (a * A.hn[pk]) * ulp;  // mass of one ulp in this cell
                if (!(wm > 0.0)) continue;
                const double n = fmin(fmax(rint(R / wm), -1.0), 1.0);
                if (n != 0.0) {
                    const float cn = c + (float)(n * ulp);
                    R -= (a * A.hn[pk]) * ((double)cn - (double)c);
                    ft[pk] = cn;
                    if (fabs(R) <= tol) return;
                }
            }
        }
    }
}

template <typename R>
int launch_inventory(const StepArgs& A, const double* h, const double* area, int ntr, const R* f, double* part,
                     double* inv, const double* target, double* ratio, hipStream_t s)
{
    dim3 block(IBX, IBY, 1);
    dim3 grid((A.nx_ + IBX - 1) / IBX, (A.ny_ + IBY - 1) / IBY, ntr);
    const int nblk = grid.x * grid.y;
    ProfileScope ps(GETM_K_INVENTORY, s);
    hipLaunchKernelGGL((k_inventory<R>), grid, block, 0, s, A, h, area, ntr, f, part);
    hipLaunchKernelGGL(k_inventory_reduce, dim3((ntr + 63) / 64), dim3(64), 0, s, ntr, nblk, part, inv, target, ratio);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

int launch_fix_f32(const StepArgs& A, int ntr, float* f, const double* part, const double* ratio, const double* area,
                   double* part2, const double* target, double* resid, hipStream_t s)
{
    dim3 block(IBX, IBY, 1);
    dim3 grid((A.nx_ + IBX - 1) / IBX, (A.ny_ + IBY - 1) / IBY, ntr);
    const int nblk = grid.x * grid.y;
    ProfileScope ps(GETM_K_FIX, s);
    hipLaunchKernelGGL(k_fix_f32, grid, block, 0, s, A, ntr, f, part, ratio, area, part2);
    hipLaunchKernelGGL(k_residual, dim3((ntr + 63) / 64), dim3(64), 0, s, ntr, nblk, part2, target, resid);
    hipLaunchKernelGGL(k_greedy_f32, dim3((ntr + 63) / 64), dim3(64), 0, s, A, ntr, f, area, resid, target, 1 << 16);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

size_t inventory_part_doubles(const StepArgs& A, int ntr)
{
    const size_t nblk = (size_t)((A.nx_ + IBX - 1) / IBX) * ((A.ny_ + IBY - 1) / IBY);
    return nblk * 3 * (size_t)ntr;
}

template int launch_inventory<float>(const StepArgs&, const double*, const double*, int, const float*, double*,
                                     double*, const double*, double*, hipStream_t);
template int launch_inventory<double>(const StepArgs&, const double*, const double*, int, const double*, double*,
                                      double*, const double*, double*, hipStream_t);

}  // namespace getm
