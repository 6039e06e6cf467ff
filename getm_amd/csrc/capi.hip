// C-ABI entry points (include/getm_amd.h): object lifetime, argument checks
// and the stage sequencing of the reference's Python/Cython layer.
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "internal.h"

using namespace getm;

int getm::DeviceBuffer::ensure(size_t n)
{
    if (n <= bytes) return GETM_OK;
    release();
    if (hipMalloc(&ptr, n) != hipSuccess) {
        ptr = nullptr;
        return GETM_ERR_HIP;
    }
    bytes = n;
    return GETM_OK;
}

void getm::DeviceBuffer::release()
{
    if (ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
}

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
static StepArgs grid_args(const Grid* t, const Grid* gu, const Grid* gv, const double* u, const double* v,
                          const double* Ah_u, const double* Ah_v);

extern "C" {

void* create_grid(int imin, int imax, int jmin, int jmax, int kmin, int kmax, int halox, int haloy)
{
    if (imax < imin || jmax < jmin || kmax < kmin || halox < 0 || haloy < 0) return nullptr;
    // in-layer byte offsets of FP64 fields are 32-bit in the kernels, and row and
    // column indices take 24-bit multiplies
    const long long nx_ = imax - imin + 1 + 2LL * halox, ny_ = jmax - jmin + 1 + 2LL * haloy;
    if (nx_ * ny_ * 8 >= (1LL << 32) || nx_ >= (1LL << 24) || ny_ >= (1LL << 24)) return nullptr;
    Grid* g = new (std::nothrow) Grid();
    if (!g) return nullptr;
    g->imin = imin;
    g->imax = imax;
    g->jmin = jmin;
    g->jmax = jmax;
    g->kmin = kmin;
    g->kmax = kmax;
    g->halox = halox;
    g->haloy = haloy;
    g->nx_ = imax - imin + 1 + 2 * halox;
    g->ny_ = jmax - jmin + 1 + 2 * haloy;
    g->nz = kmax - kmin + 1;
    // kernels index the padded block from 0: interior starts at the halo width
    g->imax = halox + (imax - imin);
    g->imin = halox;
    g->jmax = haloy + (jmax - jmin);
    g->jmin = haloy;
    return g;
}

void grid_finalize(void* grid) { delete static_cast<Grid*>(grid); }

int grid_set_array(void* grid, const char* name, void* p)
{
    Grid* g = static_cast<Grid*>(grid);
    if (!g || !name) return GETM_ERR_ARG;
    if (!strcmp(name, "mask")) g->mask = static_cast<const int*>(p);
    else if (!strcmp(name, "dx")) g->dx = static_cast<const double*>(p);
    else if (!strcmp(name, "dy")) g->dy = static_cast<const double*>(p);
    else if (!strcmp(name, "idx")) g->idx = static_cast<const double*>(p);
    else if (!strcmp(name, "idy")) g->idy = static_cast<const double*>(p);
    else if (!strcmp(name, "iarea")) g->iarea = static_cast<const double*>(p);
    else if (!strcmp(name, "area")) g->area = static_cast<const double*>(p);
    else if (!strcmp(name, "hn")) g->hn = static_cast<const double*>(p);
    else if (!strcmp(name, "ho")) g->ho = static_cast<const double*>(p);
    else return GETM_ERR_ARG;
    return GETM_OK;
}

void* advection_create(int scheme, void* tgrid)
{
    if (scheme < GETM_HSIMT || scheme > GETM_UPSTREAM || !tgrid) return nullptr;
    Advection* a = new (std::nothrow) Advection();
    if (!a) return nullptr;
    a->scheme = scheme;
    a->tgrid = static_cast<const Grid*>(tgrid);
    if (a->stats.ensure(2 * sizeof(int)) != GETM_OK || hipMemset(a->stats.ptr, 0, 2 * sizeof(int)) != hipSuccess) {
        delete a;
        return nullptr;
    }
    return a;
}

void advection_finalize(void* adv)
{
    Advection* a = static_cast<Advection*>(adv);
    if (!a) return;
    a->work.release();
    a->wsh.release();
    a->wsf.release();
    a->stats.release();
    a->hA.release();
    a->hB.release();
    delete a;
}

static int check_grids(const Grid* t, const Grid* uv)
{
    if (!t || !uv) return GETM_ERR_ARG;
    if (t->nx_ != uv->nx_ || t->ny_ != uv->ny_ || t->halox < 2 || t->haloy < 2) return GETM_ERR_ARG;
    return GETM_OK;
}

static int uv_sweep(Advection* a, int direction, int nk, const Grid* t, const Grid* uv, const double* vel,
                    const double* ah, double dt, double* h, const double* huv, const double* f_in, double* f_out,
                    hipStream_t s)
{
    return launch_adv_uv_f64(direction, a->scheme, nk, *t, *uv, vel, ah, dt, h, huv, f_in, f_out, s);
}

int advection_uv_calculate(int direction, int nk, void* adv, void* tgrid, void* uvgrid, const double* uv,
                           const double* Ah, double timestep, double* h, const double* huv, double* var,
                           void* stream)
{
    Advection* a = static_cast<Advection*>(adv);
    const Grid* t = static_cast<const Grid*>(tgrid);
    const Grid* g = static_cast<const Grid*>(uvgrid);
    if (!a) return GETM_ERR_SCHEME;
    if ((direction != 1 && direction != 2) || nk < 1 || !uv || !h || !huv || !var) return GETM_ERR_ARG;
    int rc = check_grids(t, g);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    // in place, without a copy back: x with each row's old values staged in LDS;
    // y in tiles of 32 rows, with the old rows around the tile boundaries saved first
    if (direction == 1 && t->nx_ <= 2048)
        return launch_adv_x_inplace_f64(a->scheme, nk, *t, *g, uv, Ah, timestep, h, huv, var, s);
    if (direction == 2) {
        if ((rc = a->work.ensure(std::max<size_t>(adv_y_snap_doubles(*t, nk), 1) * sizeof(double)))) return rc;
        return launch_adv_y_inplace_f64(a->scheme, nk, *t, *g, uv, Ah, timestep, h, huv, var,
                                        static_cast<double*>(a->work.ptr), s);
    }
    const size_t bytes = t->n2() * (size_t)nk * sizeof(double);
    if ((rc = a->work.ensure(bytes))) return rc;
    double* tmp = static_cast<double*>(a->work.ptr);
    if ((rc = uv_sweep(a, direction, nk, t, g, uv, Ah, timestep, h, huv, var, tmp, s))) return rc;
    GETM_HIP_OK(hipMemcpyAsync(var, tmp, bytes, hipMemcpyDeviceToDevice, s));
    return GETM_OK;
}

static int w_sweep(Advection* a, const Grid* t, const double* w, const double* wv, double dt, double* h, double* var,
                   hipStream_t s)
{
    const size_t bytes = t->n3() * sizeof(double);
    int rc;
    if ((rc = a->wsh.ensure(bytes)) || (rc = a->wsf.ensure(bytes))) return rc;
    return launch_adv_w_f64(a->scheme, *t, w, wv, dt, h, var, static_cast<double*>(a->wsh.ptr),
                            static_cast<double*>(a->wsf.ptr), static_cast<int*>(a->stats.ptr), s);
}

int advection_w_calculate(void* adv, void* tgrid, const double* w, const double* w_var, double timestep, double* h,
                          double* var, void* stream)
{
    Advection* a = static_cast<Advection*>(adv);
    const Grid* t = static_cast<const Grid*>(tgrid);
    if (!a) return GETM_ERR_SCHEME;
    if (!t || !w || !w_var || !h || !var) return GETM_ERR_ARG;
    return w_sweep(a, t, w, w_var, timestep, h, var, as_stream(stream));
}

int advection_apply_3d(void* adv, void* tgrid, void* ugrid, void* vgrid, const double* u, const double* v,
                       const double* w, const double* w_var, const double* Ah_u, const double* Ah_v, double timestep,
                       double* h, double* var, int ufirst, void* stream)
{
    Advection* a = static_cast<Advection*>(adv);
    const Grid* t = static_cast<const Grid*>(tgrid);
    const Grid* gu = static_cast<const Grid*>(ugrid);
    const Grid* gv = static_cast<const Grid*>(vgrid);
    if (!a) return GETM_ERR_SCHEME;
    if (!u || !v || !w || !h || !var) return GETM_ERR_ARG;
    int rc;
    if ((rc = check_grids(t, gu)) || (rc = check_grids(t, gv))) return rc;
    if (!gu->hn || !gv->hn) return GETM_ERR_ARG;
    if (!w_var) w_var = w;
    const size_t n3 = t->n3();
    if ((rc = a->work.ensure(n3 * sizeof(double))) || (rc = a->hA.ensure(n3 * sizeof(double))) ||
        (rc = a->hB.ensure(n3 * sizeof(double))))
        return rc;
    const bool pairs = (Ah_u == nullptr) == (Ah_v == nullptr);
    if (pairs && (rc = a->work2.ensure(n3 * sizeof(double)))) return rc;
    hipStream_t s = as_stream(stream);
    double* tmp = static_cast<double*>(a->work.ptr);
    double* hA = static_cast<double*>(a->hA.ptr);
    double* hB = static_cast<double*>(a->hB.ptr);
    // adv1/adv2 as in operators.py:118-121: v first unless ufirst.  The batched
    // FP64 kernels with one tracer: each pair of x/y sweeps (v/2 then u/2, and
    // u/2 then v/2) in one kernel through an LDS tile, bit-identical to the single
    // sweeps, with the thickness of each stage in hA / hB and the last in h;
    // var -> tmp -> tmp2 (w stage) -> var (each writes every cell of its target).
    const int d1 = ufirst ? 1 : 2, d2 = 3 - d1;
    const double half = 0.5 * timestep;
    if (!pairs) {
        // horizontal diffusion in one direction only: the single sweeps, each with
        // its own Ah (the pair kernel takes both or neither), var -> tmp -> var
        // (w in place) -> tmp -> var
        const Grid* g1 = d1 == 1 ? gu : gv;
        const Grid* g2 = d1 == 1 ? gv : gu;
        const double* vel1 = d1 == 1 ? u : v;
        const double* vel2 = d1 == 1 ? v : u;
        const double* ah1 = d1 == 1 ? Ah_u : Ah_v;
        const double* ah2 = d1 == 1 ? Ah_v : Ah_u;
        if ((rc = uv_sweep(a, d1, t->nz, t, g1, vel1, ah1, half, h, g1->hn, var, tmp, s))) return rc;
        if ((rc = uv_sweep(a, d2, t->nz, t, g2, vel2, ah2, half, h, g2->hn, tmp, var, s))) return rc;
        if ((rc = w_sweep(a, t, w, w_var, timestep, h, var, s))) return rc;
        if ((rc = uv_sweep(a, d2, t->nz, t, g2, vel2, ah2, half, h, g2->hn, var, tmp, s))) return rc;
        return uv_sweep(a, d1, t->nz, t, g1, vel1, ah1, half, h, g1->hn, tmp, var, s);
    }
    const StepArgs A = grid_args(t, gu, gv, u, v, Ah_u, Ah_v);
    // 16-row tiles of two cells per lane (8 rows with horizontal diffusion)
    const int ty = Ah_u ? auto_tile_y(1) : 16;
    if ((rc = launch_hpair<double>(d1, a->scheme, ty, A, half, 1, 1, var, tmp, h, hA, 0u, s))) return rc;
    double* tmp2 = static_cast<double*>(a->work2.ptr);
    if ((rc = launch_wcol<double>(a->scheme, A, timestep, 1, 1, tmp, tmp2, hA, hB, w, w_var,
                                  static_cast<int*>(a->stats.ptr), s, true)))
        return rc;
    return launch_hpair<double>(d2, a->scheme, ty, A, half, 1, 1, tmp2, var, hB, h, 0u, s);
}

int advection_get_stats(void* adv, int* max_iters, int* capped_columns, int reset, void* stream)
{
    Advection* a = static_cast<Advection*>(adv);
    if (!a) return GETM_ERR_ARG;
    int host[2] = {0, 0};
    hipStream_t s = as_stream(stream);
    GETM_HIP_OK(hipMemcpyAsync(host, a->stats.ptr, sizeof(host), hipMemcpyDeviceToHost, s));
    GETM_HIP_OK(hipStreamSynchronize(s));
    if (max_iters) *max_iters = host[0] > 1 ? host[0] : 1;
    if (capped_columns) *capped_columns = host[1];
    if (reset) GETM_HIP_OK(hipMemsetAsync(a->stats.ptr, 0, sizeof(host), s));
    return GETM_OK;
}

void* vertical_diffusion_create(void* tgrid)
{
    if (!tgrid) return nullptr;
    VerticalDiffusion* d = new (std::nothrow) VerticalDiffusion();
    if (!d) return nullptr;
    d->tgrid = static_cast<const Grid*>(tgrid);
    return d;
}

void vertical_diffusion_finalize(void* vd)
{
    VerticalDiffusion* d = static_cast<VerticalDiffusion*>(vd);
    if (!d) return;
    d->auxo.release();
    d->auxn.release();
    d->ru.release();
    delete d;
}

int c_vertical_diffusion_prepare(void* vd, int nx, int ny, int nz, double molecular, const double* nuh,
                                 double timestep, double cnpar, const int* mask, const double* ho, const double* hn,
                                 void* stream)
{
    VerticalDiffusion* d = static_cast<VerticalDiffusion*>(vd);
    if (!d || !nuh || !ho || !hn) return GETM_ERR_ARG;
    (void)mask;  // the reference takes it but the set-up is not masked (diffusion.F90:163-198)
    const Grid* t = d->tgrid;
    if (nx != t->nx_ || ny != t->ny_ || nz < 1) return GETM_ERR_ARG;
    const size_t bytes = t->n2() * (size_t)nz * sizeof(double);
    int rc;
    if ((rc = d->auxo.ensure(bytes)) || (rc = d->auxn.ensure(bytes))) return rc;
    rc = launch_vdiff_prepare_f64(*t, nz, molecular, nuh, timestep, cnpar, ho, hn, static_cast<double*>(d->auxo.ptr),
                                  static_cast<double*>(d->auxn.ptr), as_stream(stream));
    if (rc) return rc;
    d->prepared = 1;
    d->nz_prepared = nz;
    return GETM_OK;
}

int c_vertical_diffusion_apply(void* vd, int nx, int ny, int nz, const int* mask, const double* ho, const double* hn,
                               double* var, const double* ea2, const double* ea4, void* stream)
{
    VerticalDiffusion* d = static_cast<VerticalDiffusion*>(vd);
    if (!d || !mask || !ho || !hn || !var) return GETM_ERR_ARG;
    const Grid* t = d->tgrid;
    if (nx != t->nx_ || ny != t->ny_ || nz < 1) return GETM_ERR_ARG;
    if (nz > 1 && (!d->prepared || d->nz_prepared != nz)) return GETM_ERR_NOT_PREPARED;
    int rc;
    if (nz > 1 && (rc = d->ru.ensure(t->n2() * (size_t)nz * sizeof(double)))) return rc;
    return launch_vdiff_apply_f64(*t, nz, mask, ho, hn, static_cast<const double*>(d->auxo.ptr),
                                  static_cast<const double*>(d->auxn.ptr), var, ea2, ea4,
                                  static_cast<double*>(d->ru.ptr), as_stream(stream));
}

void* transport_create(void* tgrid, void* ugrid, void* vgrid, int scheme, double cnpar, int precision, int ufirst)
{
    if (scheme < GETM_HSIMT || scheme > GETM_UPSTREAM || !tgrid || !ugrid || !vgrid) return nullptr;
    if (precision != GETM_F64 && precision != GETM_F32 && precision != GETM_F32C) return nullptr;
    const Grid* t = static_cast<const Grid*>(tgrid);
    const Grid* gu = static_cast<const Grid*>(ugrid);
    const Grid* gv = static_cast<const Grid*>(vgrid);
    if (check_grids(t, gu) || check_grids(t, gv)) return nullptr;
    Transport* tr = new (std::nothrow) Transport();
    if (!tr) return nullptr;
    tr->scheme = scheme;
    tr->precision = precision;
    // compensated w stage: 64 tracers per block (its per-level set-up in LDS is
    // shared by four tracers per wave; 256 tracers: 32.8 -> 30.5 ms against 16)
    if (precision == GETM_F32C) tr->tpb_w = 64;
    tr->ufirst = ufirst ? 1 : 0;
    tr->cnpar = cnpar;
    tr->tgrid = t;
    tr->ugrid = gu;
    tr->vgrid = gv;
    if (tr->stats.ensure(2 * sizeof(int)) != GETM_OK || hipMemset(tr->stats.ptr, 0, 2 * sizeof(int)) != hipSuccess ||
        tr->rstep_own.ensure(sizeof(unsigned long long)) != GETM_OK ||
        hipMemset(tr->rstep_own.ptr, 0, sizeof(unsigned long long)) != hipSuccess) {
        delete tr;
        return nullptr;
    }
    return tr;
}

void transport_finalize(void* p)
{
    Transport* tr = static_cast<Transport*>(p);
    if (!tr) return;
    for (DeviceBuffer* b : {&tr->scratch, &tr->hA, &tr->hB, &tr->stats, &tr->auxo, &tr->auxn, &tr->ru, &tr->part,
                            &tr->part2, &tr->ratio, &tr->resid, &tr->rstep_own})
        b->release();
    delete tr;
}

}  // extern "C"

static StepArgs grid_args(const Grid* t, const Grid* gu, const Grid* gv, const double* u, const double* v,
                          const double* Ah_u, const double* Ah_v)
{
    StepArgs A;
    A.nx_ = t->nx_;
    A.ny_ = t->ny_;
    A.nz = t->nz;
    A.imin = t->imin;
    A.imax = t->imax;
    A.jmin = t->jmin;
    A.jmax = t->jmax;
    A.n2 = t->n2();
    A.n3 = t->n3();
    A.tmask = t->mask;
    A.umask = gu->mask;
    A.vmask = gv->mask;
    A.idxu = gu->idx;
    A.dyu = gu->dy;
    A.dxv = gv->dx;
    A.idyv = gv->idy;
    A.iarea = t->iarea;
    A.u = u;
    A.v = v;
    A.hu = gu->hn;
    A.hv = gv->hn;
    A.hn = t->hn;
    A.ah_u = Ah_u;
    A.ah_v = Ah_v;
    A.rstep = nullptr;
    return A;
}

static StepArgs step_args(const Transport* tr, const double* u, const double* v, const double* Ah_u, const double* Ah_v)
{
    return grid_args(tr->tgrid, tr->ugrid, tr->vgrid, u, v, Ah_u, Ah_v);
}

// Every grid-dimension limit of the step's launches, checked before the first
// of them, so that a call that cannot run leaves the tracers untouched: the
// tracer batches of the w stage and of the diffusion (grid.z <= 65535; the FP64
// diffusion takes one tracer per z slice), the work items of the horizontal
// stage pairs (a 1-D grid of at most 0x7fffff00 blocks, at least 4-row tiles of
// 32 columns and one layer each) and the column solve's depth.
static int step_capacity(const Transport* tr, int ntr)
{
    const Grid* t = tr->tgrid;
    auto batches = [&](int tpb) -> long { return tpb <= 0 || tpb >= ntr ? 1 : (ntr + tpb - 1) / tpb; };
    if (batches(tr->tpb_w) > 65535) return GETM_ERR_CAPACITY;
    if (tr->precision == GETM_F64 ? ntr > 65535 : batches(tr->tpb_d) > 65535) return GETM_ERR_CAPACITY;
    if (tr->precision != GETM_F64 && (t->nz > 64 || t->nz < 1)) return GETM_ERR_CAPACITY;
    const long tiles = (long)((t->nx_ + 31) / 32) * ((t->ny_ + 3) / 4);
    if ((double)batches(tr->tpb_h) * t->nz * tiles > (double)0x7fffff00L) return GETM_ERR_CAPACITY;
    return GETM_OK;
}

// Consecutive tracers [t0, t1) that share a value: the step's per-tracer
// w stages and diffusion run once per such run of tracers.
template <typename V, typename F>
static int for_runs(int ntr, const V* vals, V fallback, F&& fn)
{
    int t0 = 0;
    while (t0 < ntr) {
        const V v0 = vals ? vals[t0] : fallback;
        int t1 = t0 + 1;
        while (t1 < ntr && (vals ? vals[t1] : fallback) == v0) t1++;
        const int rc = fn(t0, t1, v0);
        if (rc) return rc;
        t0 = t1;
    }
    return GETM_OK;
}

// One step of ntr tracers, tracer t advected in the w stage by wv[t] (NULL: w)
// and diffused with molecular diffusivity mol[t].  Stages 1-2 and 4-5 take all
// tracers at once; the w stage and the diffusion run once per run of tracers
// with equal w_var / molecular value (tracer.py:223-231, 273-275).
//   R = double: the FP64 form; R = float: FP32 (C = false) or compensated FP32
//   (C = true: f holds ntr FP32 planes and then ntr int16 residual planes, and
//   so does the scratch buffer).
template <typename R, bool C>
static int transport_step_t(Transport* tr, const StepArgs& A, int ntr, R* f, const double* w,
                            const double* const* wv, const double* nuh, double dt, const double* mol,
                            double* h_out, hipStream_t s)
{
    const Grid* t = tr->tgrid;
    const size_t n3 = t->n3(), nn = n3 * (size_t)ntr;
    int rc;
    const size_t per = C ? sizeof(float) + sizeof(short) : sizeof(R);
    if ((rc = tr->scratch.ensure(nn * per)) || (rc = tr->hA.ensure(n3 * sizeof(double))) ||
        (rc = tr->hB.ensure(n3 * sizeof(double))))
        return rc;
    R* S = static_cast<R*>(tr->scratch.ptr);
    short* fl = C ? reinterpret_cast<short*>(f + nn) : nullptr;
    short* SL = C ? reinterpret_cast<short*>(S + nn) : nullptr;
    double* hA = static_cast<double*>(tr->hA.ptr);
    double* hB = static_cast<double*>(tr->hB.ptr);
    double* hfinal = h_out ? h_out : hA;
    const double half = 0.5 * dt;
    const int d1 = tr->ufirst ? 1 : 2;  // first direction of stages 1-2 (operators.py:118-121)
    const int d2 = 3 - d1;
    const int tpb_h = tr->tpb_h <= 0 || tr->tpb_h > ntr ? ntr : tr->tpb_h;
    // FP32 and compensated forms with many tracers per block: 16-row tiles, two
    // cells per lane (less halo to load and sweep than 8-row tiles); and the FP64
    // and compensated forms (FP64 working values) without horizontal diffusion with
    // any number of tracers: the same tiles (coastal, 2 tracers: FP64 pairs 1.02 /
    // 1.09 -> 0.99 / 1.01 ms, compensated 1.22 / 1.27 -> 1.06 / 1.06 ms against
    // 8-row tiles)
    const bool tall = ((C || std::is_same<R, float>::value) && ntr >= 16 && hpair_group(tpb_h, !C) == 1) ||
                      ((std::is_same<R, double>::value || C) && !A.ah_u);
    const int ty = tr->tile_y ? tr->tile_y : (tall ? 16 : auto_tile_y(ntr));
    // compensated form with many tracers: the y-first pair on 12-row tiles, two
    // 768-lane blocks per CU (its first stage adds no rows, so the shorter tile
    // costs it no halo work); the x-first pair on 16-row tiles, two cells per lane
    auto ty_of = [&](int dfirst) {
        return (C && !tr->tile_y && ntr >= 16 && dfirst == 2 && !A.ah_u) ? 12 : ty;
    };
    int* stats = static_cast<int*>(tr->stats.ptr);
    // FP32 forms: the w stage leaves the thickness of the columns it does not update
    // in hA (stages 1-2 wrote every cell there), and stages 4-5 read those cells from
    // hA instead of a copy in hB; into hA itself they store updated cells only.
    constexpr bool f32 = std::is_same<R, float>::value;
    // the FP32 forms' stage pairs cover the interior only: the halo ring's
    // thickness (which no stage changes) goes to hA and to h_out first
    if constexpr (f32 || C)
        if ((rc = launch_h_ring(A, t->ho, hA, hfinal != hA ? hfinal : nullptr, s))) return rc;
    if constexpr (C)
        rc = launch_hpair_c(d1, tr->scheme, ty_of(d1), A, half, ntr, tr->tpb_h, f, fl, S, SL, t->ho, hA, s);
    else
        rc = launch_hpair<R>(d1, tr->scheme, ty, A, half, ntr, tr->tpb_h, f, S, t->ho, hA, SR_PAIR1, s);
    if (rc) return rc;
    rc = for_runs<const double*>(ntr, wv, nullptr, [&](int t0, int t1, const double* wvar) -> int {
        const double* wu = wvar ? wvar : w;
        const size_t o = (size_t)t0 * n3;
        if constexpr (C)
            return launch_wcol_c(tr->scheme, A, dt, t1 - t0, tr->tpb_w, S + o, SL + o, f + o, fl + o, hA, hB, w, wu,
                                 stats, s, false);
        else
            return launch_wcol<R>(tr->scheme, A, dt, t1 - t0, tr->tpb_w, S + o, f + o, hA, hB, w, wu, stats, s, !f32);
    });
    if (rc) return rc;
    if constexpr (C)
        rc = launch_hpair_c(d2, tr->scheme, ty_of(d2), A, half, ntr, tr->tpb_h, f, fl, S, SL, hB, hfinal, s, hA,
                            hfinal != hA);
    else
        rc = launch_hpair<R>(d2, tr->scheme, ty, A, half, ntr, tr->tpb_h, f, S, hB, hfinal, SR_PAIR2, s,
                             f32 ? hA : hB, !(f32 && hfinal == hA));
    if (rc) return rc;
    rc = for_runs<double>(ntr, mol, 0.0, [&](int t0, int t1, double molecular) -> int {
        const size_t o = (size_t)t0 * n3;
        const int nr = t1 - t0;
        const double* ea4 = tr->ea4 ? tr->ea4 + o : nullptr;
        if constexpr (C) {
            return launch_vdiff_c(A, dt, tr->cnpar, molecular, nuh, nr, tr->tpb_d, S + o, SL + o, f + o, fl + o, s,
                                  ea4);
        } else if constexpr (f32) {
            if (t->nz >= 2 && t->nz <= tr->pcr_max_nz && !ea4)
                return launch_vdiff_pcr(A, dt, tr->cnpar, molecular, nuh, nr, tr->tpb_d, S + o, f + o, s);
            return launch_vdiff_f32(A, dt, tr->cnpar, molecular, nuh, nr, tr->tpb_d, S + o, f + o, s, ea4);
        } else {
            // FP64: the reference's prepare + per-tracer apply (tracer.py:273-276) in one
            // launch, scratch -> tracers, bit for bit
            return launch_vdiff_batch_f64(*t, t->nz, t->mask, t->hn, nuh, dt, tr->cnpar, molecular, nr, S + o, f + o,
                                          s);
        }
    });
    if (rc) return rc;
    if constexpr (f32 && !C) return launch_count_step(const_cast<unsigned long long*>(A.rstep), s);
    return GETM_OK;
}

// Checks shared by transport_step and transport_step_mixed, then the step.
static int conserve_check(const Transport* tr, int ntracers, const void* tracers, const double* target)
{
    if (!tr || ntracers < 1 || !tracers || !target) return GETM_ERR_ARG;
    if (tr->precision == GETM_F64) return GETM_OK;
    const Grid* t = tr->tgrid;
    if (!t->area || !t->mask || !t->hn) return GETM_ERR_ARG;
    if (ntracers > 65535) return GETM_ERR_CAPACITY;
    return GETM_OK;
}

// The fixer of the FP32 forms with its working memory given: part (inventory
// partial sums), ratio (per tracer), and for GETM_F32 part2 (fixer partial sums)
// and resid (2 per tracer).
static int conserve_run(Transport* tr, const StepArgs& A, int ntracers, void* tracers, const double* target,
                        double* part, double* part2, double* ratio, double* resid, hipStream_t s)
{
    const Grid* t = tr->tgrid;
    float* f = static_cast<float*>(tracers);
    int rc;
    if (tr->precision == GETM_F32C) {
        // x (1 + (target - I) / I) on every wet cell, exact to the storage rounding
        short* lo = reinterpret_cast<short*>(f + t->n3() * (size_t)ntracers);
        if ((rc = launch_inventory_c(A, t->hn, t->area, ntracers, f, lo, part, nullptr, target, ratio, s))) return rc;
        return launch_fix_c(A, ntracers, f, lo, ratio, s);
    }
    if ((rc = launch_inventory<float>(A, t->hn, t->area, ntracers, f, part, nullptr, target, ratio, s))) return rc;
    return launch_fix_f32(A, ntracers, f, part, ratio, t->area, part2, target, resid, s);
}

// Working memory of transport_conserve_noalloc: one array per GPU, in the
// library's code object, held until the process ends; every call on a GPU
// shares it.  Host code must not use the name `scratch` as an address (it is a
// host-side stand-in): conserve_scratch looks up each device's address of it
// once, with that device current, and callers use that.
constexpr size_t kConserveScratchDoubles = size_t(4) << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kConserveScratchDoubles];

static std::mutex scratch_lock;
static double* scratch_at[256] = {};  // device address of `scratch` per device ordinal

static int conserve_scratch(double** out)
{
    int dev = -1;
    GETM_HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) return GETM_ERR_ARG;
    std::lock_guard<std::mutex> guard(scratch_lock);
    if (!scratch_at[dev]) {
        void* p = nullptr;
        GETM_HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return GETM_ERR_HIP;
        scratch_at[dev] = static_cast<double*>(p);
    }
    *out = scratch_at[dev];
    return GETM_OK;
}

static int transport_step_any(Transport* tr, int ntracers, void* tracers, const double* u, const double* v,
                              const double* w, const double* const* w_var, const double* nuh, const double* Ah_u,
                              const double* Ah_v, double timestep, const double* molecular, double* h_out,
                              void* stream)
{
    if (!tr) return GETM_ERR_ARG;
    if (ntracers < 0 || (ntracers > 0 && !tracers) || !u || !v || !w || !nuh) return GETM_ERR_ARG;
    if ((Ah_u == nullptr) != (Ah_v == nullptr)) return GETM_ERR_ARG;
    const Grid* t = tr->tgrid;
    if (!t->ho || !t->hn || !t->mask || !t->iarea || !tr->ugrid->hn || !tr->vgrid->hn || !tr->ugrid->mask ||
        !tr->vgrid->mask || !tr->ugrid->idx || !tr->ugrid->dy || !tr->vgrid->dx || !tr->vgrid->idy)
        return GETM_ERR_ARG;
    if (t->nz > 64 && tr->precision != GETM_F64) return GETM_ERR_CAPACITY;
    hipStream_t s = as_stream(stream);
    if (ntracers == 0) {
        if (h_out) GETM_HIP_OK(hipMemcpyAsync(h_out, t->ho, t->n3() * sizeof(double), hipMemcpyDeviceToDevice, s));
        return GETM_OK;
    }
    int rc;
    if ((rc = step_capacity(tr, ntracers))) return rc;
    StepArgs A = step_args(tr, u, v, Ah_u, Ah_v);
    if (tr->precision == GETM_F32 && tr->rounding)
        A.rstep = tr->rstep_ext ? tr->rstep_ext : static_cast<unsigned long long*>(tr->rstep_own.ptr);
    if (tr->precision == GETM_F32C)
        return transport_step_t<float, true>(tr, A, ntracers, static_cast<float*>(tracers), w, w_var, nuh, timestep,
                                             molecular, h_out, s);
    if (tr->precision == GETM_F32)
        return transport_step_t<float, false>(tr, A, ntracers, static_cast<float*>(tracers), w, w_var, nuh,
                                              timestep, molecular, h_out, s);
    return transport_step_t<double, false>(tr, A, ntracers, static_cast<double*>(tracers), w, w_var, nuh, timestep,
                                           molecular, h_out, s);
}

extern "C" {

int transport_step(void* p, int ntracers, void* tracers, const double* u, const double* v, const double* w,
                   const double* nuh, const double* Ah_u, const double* Ah_v, double timestep, double molecular,
                   double* h_out, void* stream)
{
    Transport* tr = static_cast<Transport*>(p);
    if (ntracers < 0) return GETM_ERR_ARG;
    std::vector<double> mol((size_t)ntracers, molecular);
    return transport_step_any(tr, ntracers, tracers, u, v, w, nullptr, nuh, Ah_u, Ah_v, timestep, mol.data(), h_out,
                              stream);
}

int transport_step_mixed(void* p, int ntracers, void* tracers, const double* u, const double* v, const double* w,
                         const double* const* w_var, const double* nuh, const double* Ah_u, const double* Ah_v,
                         double timestep, const double* molecular, double* h_out, void* stream)
{
    if (ntracers > 0 && !molecular) return GETM_ERR_ARG;
    return transport_step_any(static_cast<Transport*>(p), ntracers, tracers, u, v, w, w_var, nuh, Ah_u, Ah_v,
                              timestep, molecular, h_out, stream);
}

int transport_inventory(void* p, int ntracers, const void* tracers, const double* h, double* inventory, void* stream)
{
    Transport* tr = static_cast<Transport*>(p);
    if (!tr || ntracers < 1 || !tracers || !inventory) return GETM_ERR_ARG;
    const Grid* t = tr->tgrid;
    if (!t->area || !t->mask || !(h || t->hn)) return GETM_ERR_ARG;
    if (ntracers > 65535) return GETM_ERR_CAPACITY;
    const StepArgs A = step_args(tr, nullptr, nullptr, nullptr, nullptr);
    int rc;
    if ((rc = tr->part.ensure(inventory_part_doubles(A, ntracers) * sizeof(double)))) return rc;
    double* part = static_cast<double*>(tr->part.ptr);
    const double* hh = h ? h : t->hn;
    hipStream_t s = as_stream(stream);
    if (tr->precision == GETM_F32C) {
        const float* f = static_cast<const float*>(tracers);
        return launch_inventory_c(A, hh, t->area, ntracers, f,
                                  reinterpret_cast<const short*>(f + t->n3() * (size_t)ntracers), part, inventory,
                                  nullptr, nullptr, s);
    }
    if (tr->precision == GETM_F32)
        return launch_inventory<float>(A, hh, t->area, ntracers, static_cast<const float*>(tracers), part, inventory,
                                       nullptr, nullptr, s);
    return launch_inventory<double>(A, hh, t->area, ntracers, static_cast<const double*>(tracers), part, inventory,
                                    nullptr, nullptr, s);
}

int transport_conserve(void* p, int ntracers, void* tracers, const double* target, void* stream)
{
    Transport* tr = static_cast<Transport*>(p);
    int rc = conserve_check(tr, ntracers, tracers, target);
    if (rc || tr->precision == GETM_F64) return rc;
    const StepArgs A = step_args(tr, nullptr, nullptr, nullptr, nullptr);
    if ((rc = tr->part.ensure(inventory_part_doubles(A, ntracers) * sizeof(double))) ||
        (tr->precision == GETM_F32 && (rc = tr->part2.ensure(fix_part_doubles(A, ntracers) * sizeof(double)))) ||
        (rc = tr->ratio.ensure((size_t)ntracers * sizeof(double))) ||
        (tr->precision == GETM_F32 && (rc = tr->resid.ensure(2 * (size_t)ntracers * sizeof(double)))))
        return rc;
    return conserve_run(tr, A, ntracers, tracers, target, static_cast<double*>(tr->part.ptr),
                        static_cast<double*>(tr->part2.ptr), static_cast<double*>(tr->ratio.ptr),
                        static_cast<double*>(tr->resid.ptr), as_stream(stream));
}

int transport_conserve_noalloc(void* p, int ntracers, void* tracers, const double* target, void* stream)
{
    Transport* tr = static_cast<Transport*>(p);
    int rc = conserve_check(tr, ntracers, tracers, target);
    if (rc || tr->precision == GETM_F64) return rc;
    const StepArgs A = step_args(tr, nullptr, nullptr, nullptr, nullptr);
    const size_t np = inventory_part_doubles(A, ntracers);
    const size_t n2 = tr->precision == GETM_F32 ? fix_part_doubles(A, ntracers) : 0;
    const size_t nr = tr->precision == GETM_F32 ? 2 * (size_t)ntracers : 0;
    if (np + n2 + (size_t)ntracers + nr > kConserveScratchDoubles) return GETM_ERR_CAPACITY;
    double* base = nullptr;
    if ((rc = conserve_scratch(&base))) return rc;
    // every kernel writes its part of the scratch before reading it: nothing to clear
    double* part = base;
    double* part2 = part + np;
    double* ratio = part2 + n2;
    double* resid = ratio + ntracers;
    return conserve_run(tr, A, ntracers, tracers, target, part, part2, ratio, resid, as_stream(stream));
}

int transport_check_finite(void* p, int ntracers, const void* tracers, long long* result, void* stream)
{
    Transport* tr = static_cast<Transport*>(p);
    if (!tr || ntracers < 0 || (ntracers > 0 && (!tracers || !result))) return GETM_ERR_ARG;
    const Grid* t = tr->tgrid;
    if (!t->mask) return GETM_ERR_ARG;
    const bool f32 = tr->precision != GETM_F64;  // GETM_F32C: the FP32 planes come first
    return launch_check_finite(*t, ntracers, f32 ? static_cast<const float*>(tracers) : nullptr,
                               f32 ? nullptr : static_cast<const double*>(tracers), result, as_stream(stream));
}

int transport_set_tuning(void* p, int tile_y, int tracers_per_block_h, int tracers_per_block_w, int tracers_per_block_d)
{
    Transport* tr = static_cast<Transport*>(p);
    if (!tr || tracers_per_block_h < 0 || tracers_per_block_w < 0 || tracers_per_block_d < 0) return GETM_ERR_ARG;
    // 12-row tiles: the compensated form only
    if (tile_y != 0 && tile_y != 4 && tile_y != 8 && tile_y != 16 && !(tile_y == 12 && tr->precision == GETM_F32C))
        return GETM_ERR_ARG;
    tr->tile_y = tile_y;
    tr->tpb_h = tracers_per_block_h;
    tr->tpb_w = tracers_per_block_w;
    tr->tpb_d = tracers_per_block_d;
    return GETM_OK;
}

int transport_set_solver(void* p, int pcr_max_nz)
{
    Transport* tr = static_cast<Transport*>(p);
    if (!tr || pcr_max_nz < 0 || pcr_max_nz > 32) return GETM_ERR_ARG;
    tr->pcr_max_nz = pcr_max_nz;
    return GETM_OK;
}

int transport_set_rounding(void* p, int mode, unsigned long long* step_counter)
{
    Transport* tr = static_cast<Transport*>(p);
    if (!tr || (mode != GETM_ROUND_NEAREST && mode != GETM_ROUND_STOCHASTIC)) return GETM_ERR_ARG;
    tr->rounding = mode;
    tr->rstep_ext = step_counter;
    return GETM_OK;
}

int transport_set_sources(void* p, const double* ea4)
{
    Transport* tr = static_cast<Transport*>(p);
    if (!tr) return GETM_ERR_ARG;
    if (ea4 && tr->precision == GETM_F64) return GETM_ERR_ARG;
    tr->ea4 = ea4;
    return GETM_OK;
}

int transport_get_stats(void* p, int* max_iters, int* capped_columns, int reset, void* stream)
{
    Transport* tr = static_cast<Transport*>(p);
    if (!tr) return GETM_ERR_ARG;
    int host[2] = {0, 0};
    hipStream_t s = as_stream(stream);
    GETM_HIP_OK(hipMemcpyAsync(host, tr->stats.ptr, sizeof(host), hipMemcpyDeviceToHost, s));
    GETM_HIP_OK(hipStreamSynchronize(s));
    if (max_iters) *max_iters = host[0] > 1 ? host[0] : 1;
    if (capped_columns) *capped_columns = host[1];
    if (reset) GETM_HIP_OK(hipMemsetAsync(tr->stats.ptr, 0, sizeof(host), s));
    return GETM_OK;
}


void* ensemble_create(int ndomains, void* const* tgrids, void* const* ugrids, void* const* vgrids,
                      const int* ntracers, int scheme, double cnpar, int ufirst)
{
    if (ndomains < 1 || !tgrids || !ugrids || !vgrids || !ntracers) return nullptr;
    if (scheme < GETM_HSIMT || scheme > GETM_UPSTREAM) return nullptr;
    for (int d = 0; d < ndomains; d++) {
        const Grid* t = static_cast<const Grid*>(tgrids[d]);
        if (check_grids(t, static_cast<const Grid*>(ugrids[d])) || check_grids(t, static_cast<const Grid*>(vgrids[d])))
            return nullptr;
        if (ntracers[d] < 1 || t->nz > 64) return nullptr;
    }
    Ensemble* e = new (std::nothrow) Ensemble();
    if (!e) return nullptr;
    e->ndom = ndomains;
    e->scheme = scheme;
    e->cnpar = cnpar;
    e->ufirst = ufirst ? 1 : 0;
    e->host = new (std::nothrow) EnsDom<float>[ndomains]();
    e->tg = new (std::nothrow) const Grid*[ndomains];
    e->ug = new (std::nothrow) const Grid*[ndomains];
    e->vg = new (std::nothrow) const Grid*[ndomains];
    e->rstep_ext = new (std::nothrow) unsigned long long*[ndomains]();
    if (!e->host || !e->tg || !e->ug || !e->vg || !e->rstep_ext || e->stats.ensure(2 * sizeof(int)) != GETM_OK ||
        hipMemset(e->stats.ptr, 0, 2 * sizeof(int)) != hipSuccess ||
        e->rstep_own.ensure((size_t)ndomains * sizeof(unsigned long long)) != GETM_OK ||
        hipMemset(e->rstep_own.ptr, 0, (size_t)ndomains * sizeof(unsigned long long)) != hipSuccess) {
        ensemble_finalize(e);
        return nullptr;
    }
    for (int d = 0; d < ndomains; d++) {
        e->tg[d] = static_cast<const Grid*>(tgrids[d]);
        e->ug[d] = static_cast<const Grid*>(ugrids[d]);
        e->vg[d] = static_cast<const Grid*>(vgrids[d]);
        e->host[d].ntr = ntracers[d];
    }
    return e;
}

void ensemble_finalize(void* p)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e) return;
    for (DeviceBuffer* b : {&e->doms, &e->work_h, &e->work_w, &e->work_d, &e->scratch, &e->hbuf, &e->stats,
                            &e->rstep_own})
        b->release();
    delete[] e->host;
    delete[] e->rstep_ext;
    delete[] e->tg;
    delete[] e->ug;
    delete[] e->vg;
    delete e;
}

int ensemble_bind(void* p, int d, float* tracers, const double* u, const double* v, const double* w, const double* nuh,
                  double timestep, double molecular)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e || d < 0 || d >= e->ndom || !tracers || !u || !v || !w || !nuh) return GETM_ERR_ARG;
    const Grid* t = e->tg[d];
    if (!t->ho || !t->hn || !t->mask || !t->iarea || !e->ug[d]->hn || !e->vg[d]->hn || !e->ug[d]->mask ||
        !e->vg[d]->mask || !e->ug[d]->idx || !e->ug[d]->dy || !e->vg[d]->dx || !e->vg[d]->idy)
        return GETM_ERR_ARG;
    Transport view;  // only to reuse step_args
    view.tgrid = t;
    view.ugrid = e->ug[d];
    view.vgrid = e->vg[d];
    EnsDom<float>& D = e->host[d];
    D.A = step_args(&view, u, v, nullptr, nullptr);
    D.tracers = tracers;
    D.ho = t->ho;
    D.w = w;
    D.nuh = nuh;
    D.dt = timestep;
    D.cnpar = e->cnpar;
    D.molecular = molecular;
    e->committed = 0;
    return GETM_OK;
}

int ensemble_set_tile_x(void* p, int tile_x)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e || (tile_x != 32 && tile_x != 64 && tile_x != 128)) return GETM_ERR_ARG;
    e->tile_x = tile_x;
    e->committed = 0;
    return GETM_OK;
}

int ensemble_set_tuning(void* p, int tile_y, int tpb_h, int tpb_w, int tpb_d)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e || tpb_h < 0 || tpb_w < 0 || tpb_d < 0 || (tile_y != 0 && tile_y != 4 && tile_y != 8 && tile_y != 16))
        return GETM_ERR_ARG;
    e->tile_y = tile_y;
    e->tpb_h = tpb_h;
    e->tpb_w = tpb_w;
    e->tpb_d = tpb_d;
    e->committed = 0;
    return GETM_OK;
}

int ensemble_set_solver(void* p, int pcr_max_nz)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e || pcr_max_nz < 0 || pcr_max_nz > 32) return GETM_ERR_ARG;
    e->pcr_max_nz = pcr_max_nz;
    e->committed = 0;
    return GETM_OK;
}

int ensemble_set_rounding(void* p, int mode, unsigned long long* const* step_counters)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e || (mode != GETM_ROUND_NEAREST && mode != GETM_ROUND_STOCHASTIC)) return GETM_ERR_ARG;
    e->rounding = mode;
    for (int d = 0; d < e->ndom; d++) e->rstep_ext[d] = step_counters ? step_counters[d] : nullptr;
    e->committed = 0;
    return GETM_OK;
}

static int upload_work(DeviceBuffer& buf, const std::vector<int4>& w)
{
    if (w.empty()) return GETM_OK;
    int rc = buf.ensure(w.size() * sizeof(int4));
    if (rc) return rc;
    GETM_HIP_OK(hipMemcpy(buf.ptr, w.data(), w.size() * sizeof(int4), hipMemcpyHostToDevice));
    return GETM_OK;
}

int ensemble_commit(void* p, void* stream)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e) return GETM_ERR_ARG;
    (void)stream;
    size_t nscr = 0, nh = 0;
    long ntr_all = 0;
    for (int d = 0; d < e->ndom; d++) {
        if (!e->host[d].tracers) return GETM_ERR_ARG;  // not bound
        nscr += e->tg[d]->n3() * (size_t)e->host[d].ntr;
        nh += 2 * e->tg[d]->n3();
        ntr_all += e->host[d].ntr;
    }
    // default tiles: 16 rows, two cells per lane at 64 wide (96 uneven domains of 1-32
    // tracers: 80.8 ms per step against 101.2 ms on 64 x 8 tiles); 8 rows at 128 wide
    e->tile_y_used = e->tile_y ? e->tile_y : (e->tile_x <= 64 ? 16 : 8);
    if (!ens_tile_ok(e->tile_x, e->tile_y_used)) return GETM_ERR_ARG;  // e.g. 128-wide tiles of 16 rows
    e->tile_x_used = e->tile_x;
    e->hpair_kg = 1;  // one layer per work item: measured faster than layer groups on uneven ensembles
    int rc;
    if ((rc = e->scratch.ensure(nscr * sizeof(float))) || (rc = e->hbuf.ensure(nh * sizeof(double)))) return rc;
    float* scr = static_cast<float*>(e->scratch.ptr);
    double* hb = static_cast<double*>(e->hbuf.ptr);
    std::vector<int4> wh, ww, wd_class[8], wp_class[3];
    e->nzmax = 1;
    for (int d = 0; d < e->ndom; d++) {
        EnsDom<float>& D = e->host[d];
        const Grid* t = e->tg[d];
        D.scratch = scr;
        scr += t->n3() * (size_t)D.ntr;
        D.hA = hb;
        D.hB = hb + t->n3();
        hb += 2 * t->n3();
        D.stats = static_cast<int*>(e->stats.ptr);
        D.rstep = nullptr;
        if (e->rounding)
            D.rstep = e->rstep_ext[d] ? e->rstep_ext[d] : static_cast<unsigned long long*>(e->rstep_own.ptr) + d;
        D.A.rstep = D.rstep;
        D.tpb_h = (e->tpb_h <= 0 || e->tpb_h > D.ntr) ? D.ntr : e->tpb_h;
        D.tpb_w = (e->tpb_w <= 0 || e->tpb_w > D.ntr) ? D.ntr : e->tpb_w;
        D.tpb_d = (e->tpb_d <= 0 || e->tpb_d > D.ntr) ? D.ntr : e->tpb_d;
        if (t->nz > e->nzmax) e->nzmax = t->nz;
        const int nbh = (D.ntr + D.tpb_h - 1) / D.tpb_h;
        const int nbw = (D.ntr + D.tpb_w - 1) / D.tpb_w;
        const int nbd = (D.ntr + D.tpb_d - 1) / D.tpb_d;
        if (nbh > 32767 || t->nz > 65535) return GETM_ERR_CAPACITY;
        // logical order as k_hpair: tracer batch fastest, then layer group, then tile
        for (int y0 = 0; y0 < t->ny_; y0 += e->tile_y_used)
            for (int x0 = 0; x0 < t->nx_; x0 += e->tile_x_used)
                for (int k0 = 0; k0 < t->nz; k0 += e->hpair_kg)
                    for (int b = 0; b < nbh; b++) wh.push_back(make_int4(d, x0, y0, k0 | (b << 16)));
        for (int b = 0; b < nbw; b++)
            for (int by = 0; by < (t->ny_ + 3) / 4; by++)
                for (int bx = 0; bx < (t->nx_ + 63) / 64; bx++) ww.push_back(make_int4(d, bx, by, b));
        // diffusion work grouped by depth class, so each launch's kernel is unrolled for its columns
        if (t->nz >= 2 && t->nz <= e->pcr_max_nz) {
            const int cls = t->nz <= 8 ? 0 : t->nz <= 16 ? 1 : 2;
            const size_t cols = (size_t)1024 >> (3 + cls);  // columns per PCR block
            std::vector<int4>& wp = wp_class[cls];
            if (t->n2() > 0x7fffffff) return GETM_ERR_CAPACITY;
            for (int b = 0; b < nbd; b++)
                for (size_t q0 = 0; q0 < t->n2(); q0 += cols) wp.push_back(make_int4(d, (int)q0, 0, b));
        } else {
            std::vector<int4>& wd = wd_class[(t->nz - 1) / 8];
            for (int b = 0; b < nbd; b++)
                for (int j = 0; j < t->ny_; j++)
                    for (int bx = 0; bx < (t->nx_ + 63) / 64; bx++) wd.push_back(make_int4(d, bx, j, b));
        }
    }
    // horizontal-stage table in XCD-aware launch order: block b runs logical item
    // (b % 8) * (n / 8) + b / 8 (n padded to a multiple of 8 with domain -1)
    {
        const size_t per = (wh.size() + 7) / 8;
        std::vector<int4> launch(8 * per, make_int4(-1, 0, 0, 0));
        for (size_t b = 0; b < launch.size(); b++) {
            const size_t v = (b % 8) * per + b / 8;
            if (v < wh.size()) launch[b] = wh[v];
        }
        wh.swap(launch);
    }
    std::vector<int4> wd;
    for (int c = 0; c < 8; c++) {
        e->nwork_d_class[c] = (int)wd_class[c].size();
        wd.insert(wd.end(), wd_class[c].begin(), wd_class[c].end());
    }
    for (int c = 0; c < 3; c++) {
        e->nwork_p_class[c] = (int)wp_class[c].size();
        wd.insert(wd.end(), wp_class[c].begin(), wp_class[c].end());
    }
    if (wh.size() > 0x7fffffff || ww.size() > 0x7fffffff || wd.size() > 0x7fffffff) return GETM_ERR_CAPACITY;
    if ((rc = upload_work(e->work_h, wh)) || (rc = upload_work(e->work_w, ww)) || (rc = upload_work(e->work_d, wd)))
        return rc;
    e->nwork_h = (int)wh.size();
    e->nwork_w = (int)ww.size();
    e->nwork_d = (int)wd.size();
    if ((rc = e->doms.ensure((size_t)e->ndom * sizeof(EnsDom<float>)))) return rc;
    GETM_HIP_OK(hipMemcpy(e->doms.ptr, e->host, (size_t)e->ndom * sizeof(EnsDom<float>), hipMemcpyHostToDevice));
    e->committed = 1;
    return GETM_OK;
}

int ensemble_step(void* p, void* stream)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e) return GETM_ERR_ARG;
    if (!e->committed) return GETM_ERR_NOT_PREPARED;
    hipStream_t s = as_stream(stream);
    const EnsDom<float>* doms = static_cast<const EnsDom<float>*>(e->doms.ptr);
    const int d1 = e->ufirst ? 1 : 2;
    int rc;
    if ((rc = launch_hpair_multi(1, d1, e->scheme, e->tile_y_used, e->tile_x_used, e->hpair_kg, doms, static_cast<const int4*>(e->work_h.ptr), e->nwork_h,
                                 s)))
        return rc;
    if ((rc = launch_wcol_multi(e->scheme, doms, static_cast<const int4*>(e->work_w.ptr), e->nwork_w, s))) return rc;
    if ((rc = launch_hpair_multi(2, 3 - d1, e->scheme, e->tile_y_used, e->tile_x_used, e->hpair_kg, doms, static_cast<const int4*>(e->work_h.ptr),
                                 e->nwork_h, s)))
        return rc;
    const int4* wd = static_cast<const int4*>(e->work_d.ptr);
    for (int c = 0; c < 8; c++) {
        if (!e->nwork_d_class[c]) continue;
        if ((rc = launch_vdiff_multi(e->cnpar == 1.0, c, doms, wd, e->nwork_d_class[c], s))) return rc;
        wd += e->nwork_d_class[c];
    }
    for (int c = 0; c < 3; c++) {
        if (!e->nwork_p_class[c]) continue;
        if ((rc = launch_vdiff_pcr_multi(e->cnpar == 1.0, c, doms, wd, e->nwork_p_class[c], s))) return rc;
        wd += e->nwork_p_class[c];
    }
    if (e->rounding) return launch_count_steps_multi(doms, e->ndom, s);
    return GETM_OK;
}

const double* ensemble_thickness(void* p, int d)
{
    Ensemble* e = static_cast<Ensemble*>(p);
    if (!e || d < 0 || d >= e->ndom || !e->committed) return nullptr;
    return e->host[d].hA;
}

}  // extern "C"
