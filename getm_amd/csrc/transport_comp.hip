// Launchers of the compensated FP32 form (GETM_F32C): the batched step's
// kernels (transport_kernels.h) instantiated with FP64 working values, FP32
// values plus int16 residuals in memory and the FP32 limiter.
#include "transport_kernels.h"

namespace getm {

int launch_hpair_c(int d1, int scheme, int ty, const StepArgs& A, double dt, int ntr, int tpb, const float* fin,
                   const short* lin, float* fout, short* lout, const double* hin, double* hout, hipStream_t s,
                   const double* hin_dry, bool write_dry)
{
    return launch_hpair_impl<double, true>(d1, scheme, ty, A, dt, ntr, tpb, fin, fout, lin, lout, hin, hout, 0u, s,
                                           hin_dry, write_dry);
}


int launch_wcol_c(int scheme, const StepArgs& A, double dt, int ntr, int tpb, const float* fin, const short* lin,
                  float* fout, short* lout, const double* hin, double* hout, const double* w, const double* wv,
                  int* stats, hipStream_t s, bool copy_dry)
{
    if (tpb <= 0 || tpb > ntr) tpb = ntr;
    const int nb = nbatch(ntr, tpb);
    if (nb > 65535) return GETM_ERR_CAPACITY;
    ProfileScope ps(GETM_K_WCOL, s);
    // at least 16 tracers in columns of up to 40 levels: per-level terms in LDS,
    // 16 waves per row of 64 columns, each with its own tracers (no dry-column
    // copies: the next stage reads those from hin)
    constexpr int NW = 16;
    if (!copy_dry && ntr >= NW && A.nz <= 40 && (size_t)A.n3 * 4 < ((size_t)1 << 32)) {
        const dim3 grid(xcd_grid((size_t)((A.nx_ + 63) / 64) * A.ny_ * nb)), block(64, NW);
#define GETM_WCC(NZM)                                                                                      \
    GETM_SCHEME_SWITCH(scheme, {                                                                           \
        hipLaunchKernelGGL((k_wcol_c<S_, NZM, NW>), grid, block, sizeof(WcolLds<S_, NZM, NW>), s, A, dt, ntr, \
                           tpb, fin, fout, hin, hout, w, wv, stats, lin, lout);                              \
    })
        switch ((A.nz - 1) / 8) {
            case 0: GETM_WCC(8); break;
            case 1: GETM_WCC(16); break;
            case 2: GETM_WCC(24); break;
            case 3: GETM_WCC(32); break;
            default: GETM_WCC(40); break;
        }
#undef GETM_WCC
        GETM_HIP_OK(hipGetLastError());
        return GETM_OK;
    }
    dim3 block(64, 4, 1);
    dim3 grid(xcd_grid((size_t)((A.nx_ + 63) / 64) * ((A.ny_ + 3) / 4) * nb));
    // two tracers per sweep: the FP64 column windows take twice the registers of FP32
    GETM_SCHEME_SWITCH(scheme, {
        hipLaunchKernelGGL((k_wcol<S_, double, 2, true>), grid, block, 0, s, A, dt, ntr, tpb, fin, fout, hin, hout, w,
                           wv, stats, (int)copy_dry, lin, lout);
    });
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}


int launch_vdiff_c(const StepArgs& A, double dt, double cnpar, double molecular, const double* nuh, int ntr, int tpb,
                   const float* fin, const short* lin, float* fout, short* lout, hipStream_t s, const double* ea4)
{
    if (A.nz > 64 || A.nz < 1) return GETM_ERR_CAPACITY;
    if (tpb <= 0 || tpb > ntr) tpb = ntr;
    const int nb = nbatch(ntr, tpb);
    if (nb > 65535) return GETM_ERR_CAPACITY;
    const bool cn1 = cnpar == 1.0;
    // FP32 factors e2, ru (and g3, g4 unless cnpar = 1), FP64 content weights, for
    // every level of the depth class
    const int nzc = (A.nz + 7) / 8 * 8;
    const size_t shm = (size_t)(cn1 ? 2 : 4) * nzc * 64 * sizeof(float) + (size_t)nzc * 64 * sizeof(double);
    int dev = 0, max_shm = 0;
    GETM_HIP_OK(hipGetDevice(&dev));
    GETM_HIP_OK(hipDeviceGetAttribute(&max_shm, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (shm > (size_t)max_shm) return GETM_ERR_CAPACITY;
    // 32-bit offsets within a tracer (FP64 sources and thickness included)
    if ((size_t)A.n3 * sizeof(double) >= ((size_t)1 << 32)) return GETM_ERR_CAPACITY;
    constexpr int NW = 4;  // waves per block, sharing the block's factorisation
    dim3 grid(xcd_grid((size_t)((A.nx_ + 63) / 64) * A.ny_ * nb)), block(64, NW);
    ProfileScope ps(GETM_K_VDIFF_F32, s);
#define GETM_VDC(NZM)                                                                                            \
    do {                                                                                                         \
        if (ea4 && cn1)                                                                                          \
            hipLaunchKernelGGL((k_vdiff_c<NZM, true, true, NW>), grid, block, shm, s, A, dt, cnpar, molecular,     \
                               nuh, ntr, tpb, fin, lin, fout, lout, ea4);                                        \
        else if (ea4)                                                                                            \
            hipLaunchKernelGGL((k_vdiff_c<NZM, false, true, NW>), grid, block, shm, s, A, dt, cnpar, molecular,    \
                               nuh, ntr, tpb, fin, lin, fout, lout, ea4);                                        \
        else if (cn1)                                                                                            \
            hipLaunchKernelGGL((k_vdiff_c<NZM, true, false, NW>), grid, block, shm, s, A, dt, cnpar, molecular,    \
                               nuh, ntr, tpb, fin, lin, fout, lout, nullptr);                                    \
        else                                                                                                     \
            hipLaunchKernelGGL((k_vdiff_c<NZM, false, false, NW>), grid, block, shm, s, A, dt, cnpar, molecular,   \
                               nuh, ntr, tpb, fin, lin, fout, lout, nullptr);                                    \
    } while (0)
    switch ((A.nz - 1) / 8) {
        case 0: GETM_VDC(8); break;
        case 1: GETM_VDC(16); break;
        case 2: GETM_VDC(24); break;
        case 3: GETM_VDC(32); break;
        case 4: GETM_VDC(40); break;
        case 5: GETM_VDC(48); break;
        case 6: GETM_VDC(56); break;
        default: GETM_VDC(64); break;
    }
#undef GETM_VDC
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

}  // namespace getm
