// Launchers of the batched tracer-transport kernels (transport_kernels.h) for
// the FP64 and FP32 forms, the ensembles, and the stochastic-rounding step
// counters; the compensated FP32 launchers are in transport_comp.hip.
#include "transport_kernels.h"

namespace getm {


// Step counters of the stochastic rounding, advanced after the step's last
// kernel (a kernel of its own, so every block of the step read the old count).
__global__ void __launch_bounds__(64) k_count_step(unsigned long long* counter)
{
    if (threadIdx.x == 0) *counter += 1ull;
}
__global__ void __launch_bounds__(64) k_count_steps_multi(const EnsDom<float>* __restrict__ doms, int ndom)
{
    const int d = blockIdx.x * 64 + threadIdx.x;
    if (d < ndom && doms[d].rstep) *doms[d].rstep += 1ull;
}
int launch_count_step(unsigned long long* counter, hipStream_t s)
{
    if (!counter) return GETM_OK;
    hipLaunchKernelGGL(k_count_step, dim3(1), dim3(64), 0, s, counter);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}
int launch_count_steps_multi(const EnsDom<float>* doms, int ndom, hipStream_t s)
{
    if (ndom <= 0) return GETM_OK;
    hipLaunchKernelGGL(k_count_steps_multi, dim3((ndom + 63) / 64), dim3(64), 0, s, doms, ndom);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

// Thickness of the halo ring (the cells outside [imin, imax] x [jmin, jmax]),
// which no stage changes: ho into ha and, if given, hb, every layer.  The stage
// pairs of the FP32 forms cover the interior only and leave those cells to this.
__global__ void __launch_bounds__(256) k_h_ring(const StepArgs A, const double* __restrict__ ho,
                                                double* __restrict__ ha, double* __restrict__ hb)
{
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)A.n2) return;
    const int i = (int)(p % (size_t)A.nx_), j = (int)(p / (size_t)A.nx_);
    if (i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax) return;
    for (int k = 0; k < A.nz; k++) {
        const size_t o = (size_t)k * A.n2 + p;
        ha[o] = ho[o];
        if (hb) hb[o] = ho[o];
    }
}
int launch_h_ring(const StepArgs& A, const double* ho, double* ha, double* hb, hipStream_t s)
{
    const size_t nblk = ((size_t)A.n2 + 255) / 256;
    if (nblk == 0) return GETM_OK;
    if (nblk > 0x7fffffff) return GETM_ERR_CAPACITY;
    hipLaunchKernelGGL(k_h_ring, dim3((unsigned)nblk), dim3(256), 0, s, A, ho, ha, hb);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

template <typename R>
int launch_wcol(int scheme, const StepArgs& A, double dt, int ntr, int tpb, const R* fin, R* fout, const double* hin,
                double* hout, const double* w, const double* wv, int* stats, hipStream_t s, bool copy_dry)
{
    if (tpb <= 0 || tpb > ntr) tpb = ntr;
    const int nb = nbatch(ntr, tpb);
    if (nb > 65535) return GETM_ERR_CAPACITY;
    dim3 block(64, 4, 1);
    dim3 grid(xcd_grid((size_t)((A.nx_ + 63) / 64) * ((A.ny_ + 3) / 4) * nb));
    ProfileScope ps(GETM_K_WCOL, s);
    // tracers per sweep: the thickness and w loads of a sweep are shared by TB
    // tracers; 8 costs occupancy (125 against 96 VGPRs), so only for larger batches
    if constexpr (std::is_same<R, double>::value) {
        GETM_SCHEME_SWITCH(scheme, {
            hipLaunchKernelGGL((k_wcol<S_, R, 2>), grid, block, 0, s, A, dt, ntr, tpb, fin, fout, hin, hout, w, wv,
                               stats, (int)copy_dry, nullptr, nullptr);
        });
    } else if (tpb > 4) {
        GETM_SCHEME_SWITCH(scheme, {
            hipLaunchKernelGGL((k_wcol<S_, R, 8>), grid, block, 0, s, A, dt, ntr, tpb, fin, fout, hin, hout, w, wv,
                               stats, (int)copy_dry, nullptr, nullptr);
        });
    } else {
        GETM_SCHEME_SWITCH(scheme, {
            hipLaunchKernelGGL((k_wcol<S_, R, 4>), grid, block, 0, s, A, dt, ntr, tpb, fin, fout, hin, hout, w, wv,
                               stats, (int)copy_dry, nullptr, nullptr);
        });
    }
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

int launch_vdiff_f32(const StepArgs& A, double dt, double cnpar, double molecular, const double* nuh, int ntr, int tpb,
                     const float* fin, float* fout, hipStream_t s, const double* ea4)
{
    if (A.nz > 64 || A.nz < 1) return GETM_ERR_CAPACITY;
    if (tpb <= 0 || tpb > ntr) tpb = ntr;
    const int nb = nbatch(ntr, tpb);
    if (nb > 65535) return GETM_ERR_CAPACITY;
    const bool cn1 = cnpar == 1.0;
    const size_t shm = (size_t)(cn1 ? 2 : 4) * A.nz * 64 * sizeof(float);
    dim3 grid((A.nx_ + 63) / 64, A.ny_, nb);
    ProfileScope ps(GETM_K_VDIFF_F32, s);
#define GETM_VD(NZM)                                                                                             \
    do {                                                                                                         \
        if (ea4 && cn1)                                                                                          \
            hipLaunchKernelGGL((k_vdiff_f32<NZM, true, true>), grid, dim3(64), shm, s, A, dt, cnpar, molecular, nuh, \
                               ntr, tpb, fin, fout, ea4);                                                        \
        else if (ea4)                                                                                            \
            hipLaunchKernelGGL((k_vdiff_f32<NZM, false, true>), grid, dim3(64), shm, s, A, dt, cnpar, molecular,    \
                               nuh, ntr, tpb, fin, fout, ea4);                                                   \
        else if (cn1)                                                                                            \
            hipLaunchKernelGGL((k_vdiff_f32<NZM, true, false>), grid, dim3(64), shm, s, A, dt, cnpar, molecular, nuh, \
                               ntr, tpb, fin, fout, nullptr);                                                    \
        else                                                                                                     \
            hipLaunchKernelGGL((k_vdiff_f32<NZM, false, false>), grid, dim3(64), shm, s, A, dt, cnpar, molecular,   \
                               nuh, ntr, tpb, fin, fout, nullptr);                                               \
    } while (0)
    // depth class: the kernel is unrolled for 8 (cls + 1) layers and needs nz > 8 cls
    switch ((A.nz - 1) / 8) {
        case 0: GETM_VD(8); break;
        case 1: GETM_VD(16); break;
        case 2: GETM_VD(24); break;
        case 3: GETM_VD(32); break;
        case 4: GETM_VD(40); break;
        case 5: GETM_VD(48); break;
        case 6: GETM_VD(56); break;
        default: GETM_VD(64); break;
    }
#undef GETM_VD
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

// PCR depth class: 0 -> 8 levels, 1 -> 16, 2 -> 32 (2 <= nz <= 32)
static inline int pcr_class(int nz) { return nz <= 8 ? 0 : nz <= 16 ? 1 : 2; }

int launch_vdiff_pcr(const StepArgs& A, double dt, double cnpar, double molecular, const double* nuh, int ntr, int tpb,
                     const float* fin, float* fout, hipStream_t s)
{
    if (A.nz > 32 || A.nz < 2) return GETM_ERR_CAPACITY;
    if (tpb <= 0 || tpb > ntr) tpb = ntr;
    const int nb = nbatch(ntr, tpb);
    const int cls = pcr_class(A.nz);
    const size_t cols = (size_t)1024 >> (3 + cls);
    const size_t ntile = ((size_t)A.n2 + cols - 1) / cols;
    if (nb > 65535 || ntile > 0x7fffffff) return GETM_ERR_CAPACITY;
    const bool cn1 = cnpar == 1.0;
    dim3 grid((unsigned)ntile, nb);
    ProfileScope ps(GETM_K_VDIFF_F32, s);
#define GETM_PC(NP)                                                                                              \
    do {                                                                                                         \
        if (cn1)                                                                                                 \
            hipLaunchKernelGGL((k_vdiff_pcr<NP, true>), grid, dim3(1024), 0, s, A, dt, cnpar, molecular, nuh, ntr, \
                               tpb, fin, fout);                                                                  \
        else                                                                                                     \
            hipLaunchKernelGGL((k_vdiff_pcr<NP, false>), grid, dim3(1024), 0, s, A, dt, cnpar, molecular, nuh, ntr, \
                               tpb, fin, fout);                                                                  \
    } while (0)
    switch (cls) {
        case 0: GETM_PC(8); break;
        case 1: GETM_PC(16); break;
        default: GETM_PC(32); break;
    }
#undef GETM_PC
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

template <typename R>
int launch_hpair(int d1, int scheme, int ty, const StepArgs& A, double dt, int ntr, int tpb, const R* fin, R* fout,
                 const double* hin, double* hout, unsigned sr_stage, hipStream_t s, const double* hin_dry,
                 bool write_dry)
{
    return launch_hpair_impl<R, false>(d1, scheme, ty, A, dt, ntr, tpb, fin, fout, nullptr, nullptr, hin, hout,
                                       sr_stage, s, hin_dry, write_dry);
}

template int launch_hpair<float>(int, int, int, const StepArgs&, double, int, int, const float*, float*,
                                 const double*, double*, unsigned, hipStream_t, const double*, bool);
template int launch_hpair<double>(int, int, int, const StepArgs&, double, int, int, const double*, double*,
                                  const double*, double*, unsigned, hipStream_t, const double*, bool);

// ---------------------------------------------------------------------------
// ensemble launchers: one launch per kernel kind covers every domain
// ---------------------------------------------------------------------------
int launch_hpair_multi(int stage, int d1, int scheme, int ty, int tx, int kg, const EnsDom<float>* doms,
                       const int4* work, int nwork, hipStream_t s)
{
    if (nwork <= 0) return GETM_OK;
    if (kg != 1) return GETM_ERR_ARG;  // ensembles take one layer per work item
    if (!ens_tile_ok(tx, ty)) return GETM_ERR_ARG;
    ProfileScope ps(stage == 1 ? GETM_K_HPAIR_VU : GETM_K_HPAIR_UV, s);
    // tiles of 1024 cells: two cells per lane (512-lane blocks, two per CU)
#define GETM_HPM(TXV, TYV)                                                                                    \
    GETM_SCHEME_SWITCH(scheme, {                                                                              \
        constexpr int L = TXV * TYV >= 1024 ? 2 : 1;                                                          \
        const dim3 blk(TXV * TYV / L);                                                                        \
        if (stage == 1 && d1 == 2)                                                                            \
            hipLaunchKernelGGL((k_hpair_multi<S_, float, 2, TYV, 1, 1, TXV, L>), dim3(nwork), blk, 0, s, doms, \
                               work);                                                                         \
        else if (stage == 1)                                                                                  \
            hipLaunchKernelGGL((k_hpair_multi<S_, float, 1, TYV, 1, 1, TXV, L>), dim3(nwork), blk, 0, s, doms, \
                               work);                                                                         \
        else if (d1 == 2)                                                                                     \
            hipLaunchKernelGGL((k_hpair_multi<S_, float, 2, TYV, 2, 1, TXV, L>), dim3(nwork), blk, 0, s, doms, \
                               work);                                                                         \
        else                                                                                                  \
            hipLaunchKernelGGL((k_hpair_multi<S_, float, 1, TYV, 2, 1, TXV, L>), dim3(nwork), blk, 0, s, doms, \
                               work);                                                                         \
    })
    if (tx == 64) {
        if (ty == 16) {
            GETM_HPM(64, 16);
        } else if (ty == 8) {
            GETM_HPM(64, 8);
        } else {
            GETM_HPM(64, 4);
        }
    } else if (tx == 32) {
        if (ty == 16) {
            GETM_HPM(32, 16);
        } else {
            GETM_HPM(32, 8);
        }
    } else {
        if (ty == 8) {
            GETM_HPM(128, 8);
        } else {
            GETM_HPM(128, 4);
        }
    }
#undef GETM_HPM
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

int launch_wcol_multi(int scheme, const EnsDom<float>* doms, const int4* work, int nwork, hipStream_t s)
{
    if (nwork <= 0) return GETM_OK;
    ProfileScope ps(GETM_K_WCOL, s);
    GETM_SCHEME_SWITCH(scheme, {
        hipLaunchKernelGGL((k_wcol_multi<S_, float, 4>), dim3(nwork), dim3(64, 4, 1), 0, s, doms, work);
    });
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

int launch_vdiff_multi(bool cn1, int cls, const EnsDom<float>* doms, const int4* work, int nwork, hipStream_t s)
{
    if (nwork <= 0) return GETM_OK;
    if (cls < 0 || cls > 7) return GETM_ERR_CAPACITY;
    const size_t shm = (size_t)(cn1 ? 2 : 4) * 8 * (cls + 1) * 64 * sizeof(float);
    ProfileScope ps(GETM_K_VDIFF_F32, s);
#define GETM_VDM(NZM)                                                                                            \
    do {                                                                                                        \
        if (cn1)                                                                                                \
            hipLaunchKernelGGL((k_vdiff_multi<NZM, true>), dim3(nwork), dim3(64), shm, s, doms, work);          \
        else                                                                                                    \
            hipLaunchKernelGGL((k_vdiff_multi<NZM, false>), dim3(nwork), dim3(64), shm, s, doms, work);         \
    } while (0)
    switch (cls) {
        case 0: GETM_VDM(8); break;
        case 1: GETM_VDM(16); break;
        case 2: GETM_VDM(24); break;
        case 3: GETM_VDM(32); break;
        case 4: GETM_VDM(40); break;
        case 5: GETM_VDM(48); break;
        case 6: GETM_VDM(56); break;
        default: GETM_VDM(64); break;
    }
#undef GETM_VDM
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}
template int launch_wcol<float>(int, const StepArgs&, double, int, int, const float*, float*, const double*, double*,
                                const double*, const double*, int*, hipStream_t, bool);
template int launch_wcol<double>(int, const StepArgs&, double, int, int, const double*, double*, const double*,
                                 double*, const double*, const double*, int*, hipStream_t, bool);

int launch_vdiff_pcr_multi(bool cn1, int cls, const EnsDom<float>* doms, const int4* work, int nwork, hipStream_t s)
{
    if (nwork <= 0) return GETM_OK;
    if (cls < 0 || cls > 2) return GETM_ERR_CAPACITY;
    ProfileScope ps(GETM_K_VDIFF_F32, s);
#define GETM_PCM(NP)                                                                                          \
    do {                                                                                                      \
        if (cn1)                                                                                              \
            hipLaunchKernelGGL((k_vdiff_pcr_multi<NP, true>), dim3(nwork), dim3(1024), 0, s, doms, work);     \
        else                                                                                                  \
            hipLaunchKernelGGL((k_vdiff_pcr_multi<NP, false>), dim3(nwork), dim3(1024), 0, s, doms, work);    \
    } while (0)
    switch (cls) {
        case 0: GETM_PCM(8); break;
        case 1: GETM_PCM(16); break;
        default: GETM_PCM(32); break;
    }
#undef GETM_PCM
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

}  // namespace getm
