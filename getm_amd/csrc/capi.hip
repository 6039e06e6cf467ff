// C-ABI entry points (include/getm_amd.h): object lifetime, argument checks
// and the stage sequencing of the reference's Python/Cython layer.
#include <stdlib.h>
#include <string.h>

#include <new>

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

extern "C" {

void* create_grid(int imin, int imax, int jmin, int jmax, int kmin, int kmax, int halox, int haloy)
{
    if (imax < imin || jmax < jmin || kmax < kmin || halox < 0 || haloy < 0) return nullptr;
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
    const size_t bytes = t->n2() * (size_t)nk * sizeof(double);
    if ((rc = a->work.ensure(bytes))) return rc;
    hipStream_t s = as_stream(stream);
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
    const int nk = t->nz;
    if ((rc = a->work.ensure(t->n3() * sizeof(double)))) return rc;
    hipStream_t s = as_stream(stream);
    double* tmp = static_cast<double*>(a->work.ptr);
    const double half = 0.5 * timestep;
    // adv1/adv2 as in operators.py:118-121: v first unless ufirst
    const int d1 = ufirst ? 1 : 2, d2 = ufirst ? 2 : 1;
    const Grid* g1 = ufirst ? gu : gv;
    const Grid* g2 = ufirst ? gv : gu;
    const double* vel1 = ufirst ? u : v;
    const double* vel2 = ufirst ? v : u;
    const double* ah1 = ufirst ? Ah_u : Ah_v;
    const double* ah2 = ufirst ? Ah_v : Ah_u;
    // ping-pong var -> tmp -> var (w in place) -> tmp -> var
    if ((rc = uv_sweep(a, d1, nk, t, g1, vel1, ah1, half, h, g1->hn, var, tmp, s))) return rc;
    if ((rc = uv_sweep(a, d2, nk, t, g2, vel2, ah2, half, h, g2->hn, tmp, var, s))) return rc;
    if ((rc = w_sweep(a, t, w, w_var, timestep, h, var, s))) return rc;
    if ((rc = uv_sweep(a, d2, nk, t, g2, vel2, ah2, half, h, g2->hn, var, tmp, s))) return rc;
    if ((rc = uv_sweep(a, d1, nk, t, g1, vel1, ah1, half, h, g1->hn, tmp, var, s))) return rc;
    return GETM_OK;
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

}  // extern "C"
