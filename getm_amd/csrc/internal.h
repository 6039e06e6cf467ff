// Internal object layouts and launcher declarations shared by the C-ABI
// (capi.hip) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/getm_amd.h"

namespace getm {

// RAII bracket of one kernel launch with two HIP events on its stream when
// profiling is enabled (profile.hip); a no-op otherwise.
class ProfileScope {
   public:
    ProfileScope(int id, hipStream_t s);
    ~ProfileScope();

   private:
    int id_;
    hipStream_t s_;
    void* a_ = nullptr;
    void* b_ = nullptr;
};

// A horizontal grid (T, U or V points) with halo, as create_grid builds it
// (python/src/wrapper.F90:48-59).  Field pointers are caller-owned device
// buffers registered by name with grid_set_array.
struct Grid {
    int imin, imax, jmin, jmax, kmin, kmax, halox, haloy;
    int nx_, ny_, nz;  // padded extents
    const int* mask = nullptr;
    const double* dx = nullptr;
    const double* dy = nullptr;
    const double* idx = nullptr;
    const double* idy = nullptr;
    const double* iarea = nullptr;
    const double* area = nullptr;
    const double* hn = nullptr;
    const double* ho = nullptr;
    size_t n2() const { return (size_t)nx_ * ny_; }
    size_t n3() const { return n2() * nz; }
};

// Growable device scratch owned by an operator object.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t n);
    void release();
};

struct Advection {
    int scheme;
    const Grid* tgrid;
    DeviceBuffer work;    // out-of-place target of the x/y sweeps (nk layers)
    DeviceBuffer work2;   // target of the w stage in advection_apply_3d
    DeviceBuffer wsh;     // z-sweep sub-iteration thickness
    DeviceBuffer wsf;     // z-sweep sub-iteration tracer
    DeviceBuffer stats;   // int32[2]: largest sub-iteration count, columns capped at 200
    DeviceBuffer hA, hB;  // thickness after stages 1-2 and after the w stage (advection_apply_3d)
};

struct VerticalDiffusion {
    const Grid* tgrid;
    DeviceBuffer auxo, auxn, ru;
    int prepared = 0;
    int nz_prepared = 0;
};

// Everything the batched step kernels read besides the tracers, by value.
struct StepArgs {
    int nx_, ny_, nz, imin, imax, jmin, jmax;
    size_t n2, n3;
    const int* tmask;
    const int* umask;
    const int* vmask;
    const double* idxu;
    const double* dyu;
    const double* dxv;
    const double* idyv;
    const double* iarea;
    const double* u;
    const double* v;
    const double* hu;
    const double* hv;
    const double* hn;
    const double* ah_u;  // both NULL or both set
    const double* ah_v;
    // FP32 tracer stores: NULL rounds to nearest; else stochastic rounding seeded
    // by this device step counter (advanced by one after every step)
    const unsigned long long* rstep;
};

// Layers per block of the horizontal stage pair: a group of consecutive layers
// shares the block's layer-independent face and cell terms, which pays when the
// tracer-independent set-up dominates (few tracers per block); with many tracers
// one layer per block keeps the tracer loop's registers free.  FP32 only.
constexpr int HPAIR_KG = 8;
inline int hpair_group(int tracers_per_block, bool f32) { return f32 && tracers_per_block <= 8 ? HPAIR_KG : 1; }

// Rounding sites of one step (stage ids of the stochastic-rounding seeds):
// stage pair 1 -> 1, 2; stage pair 2 -> 3, 4; column solve -> 5; w sub-step it -> 16 + it.
enum { SR_PAIR1 = 1, SR_PAIR2 = 3, SR_VDIFF = 5, SR_WCOL = 16 };

// One domain of an ensemble as the ensemble kernels read it (device table).
template <typename R>
struct EnsDom {
    StepArgs A;
    R* tracers;      // ntr x nz x ny_ x nx_
    R* scratch;      // same size, stage ping-pong
    const double* ho;
    double* hA;      // thickness after stages 1-2, and the final thickness
    double* hB;      // thickness after stage 3
    const double* w;
    const double* nuh;
    double dt, cnpar, molecular;
    int ntr, tpb_h, tpb_w, tpb_d;
    int* stats;
    unsigned long long* rstep;  // = A.rstep, advanced after every step (NULL: round to nearest)
};

// Batched T-tracer step object (transport_create).
struct Transport {
    int scheme, precision, ufirst;
    double cnpar;
    const Grid* tgrid;
    const Grid* ugrid;
    const Grid* vgrid;
    DeviceBuffer scratch;  // tracer ping-pong buffer
    DeviceBuffer hA, hB;   // FP64 thickness between stages
    DeviceBuffer stats;    // int32[2] as in Advection
    DeviceBuffer auxo, auxn, ru;  // FP64 diffusion (precision GETM_F64)
    int tpb_h = 0, tpb_w = 16, tpb_d = 16;  // tracers per block of the three kernel kinds (0 = all; w: 64 for GETM_F32C)
    int tile_y = 0;                         // height of the horizontal-stage tiles (4, 8, 12, 16; 0 = by tracer count)
    int pcr_max_nz = 0;                     // FP32 column solve by wave PCR for 2 <= nz <= this (<= 32; 0: Thomas only)
    DeviceBuffer part, part2, ratio, resid;  // inventory partial sums, fixer ratios and residuals
    int rounding = 1;                        // FP32 stores: 1 stochastic (default), 0 nearest
    DeviceBuffer rstep_own;                  // the object's own step counter (zeroed at creation)
    unsigned long long* rstep_ext = nullptr; // caller's counter (transport_set_rounding), else rstep_own
    const double* ea4 = nullptr;             // FP32 tracer sources (transport_set_sources), NULL: none
};

// Ensemble of independent domains stepped by one launch per kernel kind
// (ensemble_create); FP32 tracers with FP64 thickness.
struct Ensemble {
    int ndom = 0, scheme = 5, ufirst = 0;
    double cnpar = 1.0;
    int tile_y = 0, tpb_h = 0, tpb_w = 16, tpb_d = 16;  // tile_y 0 = 16 rows (8 at 128 columns)
    int tile_y_used = 16;                               // the height of the committed work table
    int tile_x = 64;                                    // width of the horizontal-stage tiles (32, 64, 128)
    int tile_x_used = 64;                               // the width of the committed work table
    int hpair_kg = 1;                                   // layers per horizontal-stage work item (hpair_group)
    int nwork_h = 0, nwork_w = 0, nwork_d = 0, nzmax = 1;
    int nwork_d_class[8] = {};  // diffusion work items per depth class (8 c-7 .. 8 c + 8 layers), in this order
    int pcr_max_nz = 0;         // domains with 2 <= nz <= this use the wave-PCR solve
    int nwork_p_class[3] = {};  // PCR work items per class (8, 16, 32 levels), after the Thomas items
    int committed = 0;
    EnsDom<float>* host = nullptr;  // host copy of the domain table
    const Grid** tg = nullptr;
    const Grid** ug = nullptr;
    const Grid** vg = nullptr;
    DeviceBuffer doms, work_h, work_w, work_d, scratch, hbuf, stats;
    int rounding = 1;                       // as Transport::rounding
    DeviceBuffer rstep_own;                 // ndom step counters of the ensemble's own (zeroed at creation)
    unsigned long long** rstep_ext = nullptr;  // per-domain caller counters (NULL entries: own)
};

// Tile height of the horizontal stage pair for few tracers per block (up to 8
// FP32 tracers, which take layer groups) and for the FP64 and compensated forms
// with horizontal diffusion: 8 rows.  Three 512-lane blocks per CU overlap each other's barrier
// phases; 16-row tiles of one cell per lane have a smaller halo share but fit
// one or two 1024-lane blocks per CU, and were slower at 2 tracers on the
// coastal grid.  (With more FP32 tracers per block, and for the FP64 and
// compensated forms without horizontal diffusion, the pairs take 16-row tiles
// of two cells per lane: 512-lane blocks, two per CU.)
inline int auto_tile_y(int tracers_per_domain)
{
    (void)tracers_per_domain;
    return 8;
}

// ---- batched kernels (transport_batched.hip) ----
// thickness of the halo ring (outside [imin, imax] x [jmin, jmax]): ho -> ha (and hb if not NULL)
int launch_h_ring(const StepArgs& A, const double* ho, double* ha, double* hb, hipStream_t s);
// tpb = tracers per block (tracer batches go into the grid); <= 0 means all
template <typename R>
// hin_dry: thickness of the cells the stages do not update (NULL: hin); write_dry:
// also store those cells' thickness into hout (false only where hout already holds it)
int launch_hpair(int d1, int scheme, int ty, const StepArgs& A, double dt, int ntr, int tpb, const R* fin, R* fout,
                 const double* hin, double* hout, unsigned sr_stage, hipStream_t s, const double* hin_dry = nullptr,
                 bool write_dry = true);
int launch_hpair_multi(int stage, int d1, int scheme, int ty, int tx, int kg, const EnsDom<float>* doms,
                       const int4* work, int nwork, hipStream_t s);
// tile shapes of the ensembles' horizontal stages: 32 x {8, 16}, 64 x {4, 8, 16}, 128 x {4, 8}
inline bool ens_tile_ok(int tx, int ty)
{
    return (tx == 32 && (ty == 8 || ty == 16)) || (tx == 64 && (ty == 4 || ty == 8 || ty == 16)) ||
           (tx == 128 && (ty == 4 || ty == 8));
}
int launch_wcol_multi(int scheme, const EnsDom<float>* doms, const int4* work, int nwork, hipStream_t s);
// one depth class: every domain of the work items has 8 (cls + 1) - 7 .. 8 (cls + 1) layers
int launch_vdiff_multi(bool cn1, int cls, const EnsDom<float>* doms, const int4* work, int nwork, hipStream_t s);
// wave-PCR column solve (2 <= nz <= 32): single domain, and one PCR depth class (0..2: 8, 16, 32 levels) of an ensemble
int launch_vdiff_pcr(const StepArgs& A, double dt, double cnpar, double molecular, const double* nuh, int ntr, int tpb,
                     const float* fin, float* fout, hipStream_t s);
int launch_vdiff_pcr_multi(bool cn1, int cls, const EnsDom<float>* doms, const int4* work, int nwork, hipStream_t s);
template <typename R>
// copy_dry: copy the thickness of columns that are not updated from hin to hout
int launch_wcol(int scheme, const StepArgs& A, double dt, int ntr, int tpb, const R* fin, R* fout, const double* hin,
                double* hout, const double* w, const double* wv, int* stats, hipStream_t s, bool copy_dry = true);
int launch_vdiff_f32(const StepArgs& A, double dt, double cnpar, double molecular, const double* nuh, int ntr, int tpb,
                     const float* fin, float* fout, hipStream_t s,
                     const double* ea4 = nullptr);
// compensated FP32 (GETM_F32C): FP32 values plus int16 residuals (common.h cencode / cdecode)
int launch_hpair_c(int d1, int scheme, int ty, const StepArgs& A, double dt, int ntr, int tpb, const float* fin,
                   const short* lin, float* fout, short* lout, const double* hin, double* hout, hipStream_t s,
                   const double* hin_dry = nullptr, bool write_dry = true);
int launch_wcol_c(int scheme, const StepArgs& A, double dt, int ntr, int tpb, const float* fin, const short* lin,
                  float* fout, short* lout, const double* hin, double* hout, const double* w, const double* wv,
                  int* stats, hipStream_t s, bool copy_dry);
int launch_vdiff_c(const StepArgs& A, double dt, double cnpar, double molecular, const double* nuh, int ntr, int tpb,
                   const float* fin, const short* lin, float* fout, short* lout, hipStream_t s,
                   const double* ea4 = nullptr);
// advance the step counter of the stochastic rounding (single counter, or every domain's)
int launch_count_step(unsigned long long* counter, hipStream_t s);
int launch_count_steps_multi(const EnsDom<float>* doms, int ndom, hipStream_t s);

// ---- inventories (inventory.hip) ----
template <typename R>
int launch_inventory(const StepArgs& A, const double* h, const double* area, int ntr, const R* f, double* part,
                     double* inv, const double* target, double* ratio, hipStream_t s);
int launch_fix_f32(const StepArgs& A, int ntr, float* f, const double* part, const double* ratio, const double* area,
                   double* part2, const double* target, double* resid, hipStream_t s);
int launch_inventory_c(const StepArgs& A, const double* h, const double* area, int ntr, const float* f,
                       const short* lo, double* part, double* inv, const double* target, double* ratio, hipStream_t s);
int launch_fix_c(const StepArgs& A, int ntr, float* f, short* lo, const double* ratio, hipStream_t s);
size_t fix_part_doubles(const StepArgs& A, int ntr);

// ---- non-finite detection (check.hip) ----
// out: 2 ntr int64 -- per tracer the count of non-finite interior values with T
// mask 1 or 2, and the in-tracer index of the first (INT64_MAX: none)
int launch_check_finite(const Grid& g, int ntr, const float* f32, const double* f64, long long* out, hipStream_t s);
size_t inventory_part_doubles(const StepArgs& A, int ntr);

// ---- FP64 kernels (transport_fp64.hip) ----
// in-place x half-sweep (rows of up to 2048 cells; GETM_ERR_CAPACITY beyond)
int launch_adv_x_inplace_f64(int scheme, int nk, const Grid& t, const Grid& uv, const double* vel, const double* ah,
                             double dt, double* h, const double* huv, double* f, hipStream_t s);
// in-place y half-sweep; snap holds the old rows around the 32-row tile boundaries
size_t adv_y_snap_doubles(const Grid& t, int nk);
int launch_adv_y_inplace_f64(int scheme, int nk, const Grid& t, const Grid& uv, const double* vel, const double* ah,
                             double dt, double* h, const double* huv, double* f, double* snap, hipStream_t s);
int launch_adv_uv_f64(int direction, int scheme, int nk, const Grid& t, const Grid& uv,
                      const double* vel, const double* ah, double dt, double* h, const double* huv,
                      const double* f_in, double* f_out, hipStream_t s);
int launch_adv_w_f64(int scheme, const Grid& t, const double* w, const double* wv, double dt, double* h,
                     double* f, double* wsh, double* wsf, int* stats, hipStream_t s);
int launch_vdiff_prepare_f64(const Grid& t, int nz, double molecular, const double* nuh, double dt,
                             double cnpar, const double* dzo, const double* dzn, double* auxo, double* auxn,
                             hipStream_t s);
int launch_vdiff_apply_f64(const Grid& t, int nz, const int* mask, const double* dzo, const double* dzn,
                           const double* auxo, const double* auxn, double* var, const double* ea2,
                           const double* ea4, double* ru, hipStream_t s);
// the transport step's FP64 diffusion of all tracers, fin -> fout (dzo = dzn = hn);
// fin is scratch: the solve overwrites it
int launch_vdiff_batch_f64(const Grid& t, int nz, const int* mask, const double* hn, const double* nuh, double dt,
                           double cnpar, double molecular, int ntr, double* fin, double* fout, hipStream_t s);

}  // namespace getm

#define GETM_HIP_OK(expr)                                  \
    do {                                                   \
        if ((expr) != hipSuccess) return GETM_ERR_HIP;     \
    } while (0)
