/*
 * FP64 CPU restatement of pygetm's 3D tracer-transport step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker; never part of the product path.
 *
 * Written from the reference algorithm, keeping its FP64 operation order so
 * that a build without FMA contraction (-ffp-contract=off, x86-64 baseline)
 * reproduces the Fortran bit for bit:
 *   limiters           src/operators/advection_schemes/limiters.h:1-44
 *   x / y sweeps       src/operators/advection_schemes/advection.F90.template:23-221
 *   z sweep            src/operators/advection_schemes/advection.F90.template:223-374
 *   diffusion          src/operators/diffusion.F90:135-314
 *   layer loop         python/src/wrapper.F90:213-273
 *   stage order        python/pygetm/operators.py:95-193
 *   tracer step        python/pygetm/tracer.py:193-276
 * MIN/MAX follow flang's lowering (max(a,b) = a > b ? a : b, folded left to
 * right) and CEILING saturates to the int32 range with NaN -> 0, as measured
 * on amdflang 22, so NaN edge cases (HSIMT at cfl = 1) match too.
 */
#include "getm_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { HSIMT = 1, MUSCL = 2, P2_PDM = 3, SPLMAX13 = 4, SUPERBEE = 5, UPSTREAM = 6 };

static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double fmin3(double a, double b, double c) { return fmin2(fmin2(a, b), c); }

static int fceiling(double x)
{
    if (x != x) return 0;
    double c = ceil(x);
    if (c >= 2147483647.0) return INT_MAX;
    if (c <= -2147483648.0) return INT_MIN;
    return (int)c;
}

/* limiters.h:1-44 */
double oracle_limiter(int scheme, double cfl, double ratio)
{
    switch (scheme) {
    case HSIMT: {
        double kappa = 1.0 - cfl;
        double x = 0.25 * (kappa - (1.0 / 3.0) / kappa);
        double lim = (0.5 + x) + (0.5 - x) * ratio;
        lim = fmax2(0.0, lim);
        return fmin3(2.0 * ratio, lim, 2.0);
    }
    case MUSCL:
        return fmin3(2.0 * ratio, 0.5 * (1.0 + ratio), 2.0);
    case P2_PDM: {
        double x = (1.0 / 6.0) * (1.0 - 2.0 * cfl);
        double lim = (0.5 + x) + (0.5 - x) * ratio;
        return fmin3((2.0 * ratio) / (cfl + 1.e-10), lim, 2.0 / (1.0 - cfl));
    }
    case SPLMAX13:
        return fmin3(2.0 * ratio, (1.0 / 3.0) * fmax2(1.0 + 2.0 * ratio, 2.0 + ratio), 2.0);
    case SUPERBEE:
        return fmax2(fmin2(2.0 * ratio, 1.0), fmin2(ratio, 2.0));
    default: /* UPSTREAM */
        return 0.0;
    }
}

static inline int is_flux_face(int m) { return m == 1 || m == 2; }

/* Limited upwind face value (template:66-88, 166-188): f_up2 = upstream-upstream,
 * f_up = upstream, f_dn = downstream. */
static inline double limited_face_value(int scheme, double cfl, double fu, double fuu, double fd)
{
    double deltaf = fd - fu;
    double deltafu = fu - fuu;
    if (deltaf * deltafu > 0) {
        double ratio = deltafu / deltaf;
        double limiter = oracle_limiter(scheme, cfl, ratio);
        fu = fu + ((0.5 * limiter) * (1.0 - cfl)) * deltaf;
    }
    return fu;
}

static int valid_scheme(int s) { return s >= HSIMT && s <= UPSTREAM; }

/* Per-layer x or y half-sweep.  dir 1: stride between neighbours = 1 (i);
 * dir 2: stride = nx_ (j).  Faces run over [min-1, max] along the sweep
 * direction and [min, max] across it; cells over [min, max] x [min, max]. */
static void sweep_layer(int dir, int scheme, int nx_, int ny_, int halox, int haloy,
                        const int* uvmask, const double* metric1, const double* metric2,
                        const double* huv, const double* uv, const int* tmask, const double* iarea,
                        const double* ah, double dt, double* h, double* f, double* QU, double* flux)
{
    const int imin = halox, imax = nx_ - halox - 1, jmin = haloy, jmax = ny_ - haloy - 1;
    const int s = dir == 1 ? 1 : nx_;
    const int fi0 = dir == 1 ? imin - 1 : imin, fj0 = dir == 1 ? jmin : jmin - 1;
    for (int j = fj0; j <= jmax; j++) {
        for (int i = fi0; i <= imax; i++) {
            const int p = j * nx_ + i;
            if (is_flux_face(uvmask[p])) {
                const double vel = uv[p];
                double q, cfl;
                if (dir == 1) { /* QU = u*hu*dyu, cfl = |u*dt*idxu| */
                    q = (vel * huv[p]) * metric2[p];
                    cfl = fabs((vel * dt) * metric1[p]);
                } else {        /* QV = v*hv*dxv, cfl = |v*dt*idyv| */
                    q = (vel * huv[p]) * metric1[p];
                    cfl = fabs((vel * dt) * metric2[p]);
                }
                double fu, fuu, fd;
                if (vel > 0) {
                    fu = f[p];
                    fuu = fu;
                    if (is_flux_face(uvmask[p - s])) fuu = f[p - s];
                    fd = f[p + s];
                } else {
                    fu = f[p + s];
                    fuu = fu;
                    if (is_flux_face(uvmask[p + s])) fuu = f[p + 2 * s];
                    fd = f[p];
                }
                fu = limited_face_value(scheme, cfl, fu, fuu, fd);
                QU[p] = q;
                flux[p] = q * fu;
            } else {
                QU[p] = 0.0;
                flux[p] = 0.0;
            }
        }
    }
    if (ah) {
        for (int j = fj0; j <= jmax; j++) {
            for (int i = fi0; i <= imax; i++) {
                const int p = j * nx_ + i;
                if (is_flux_face(uvmask[p])) {
                    if (dir == 1)
                        flux[p] = flux[p] - (((ah[p] * huv[p]) * (f[p + s] - f[p])) * metric1[p]) * metric2[p];
                    else
                        flux[p] = flux[p] - (((ah[p] * huv[p]) * (f[p + s] - f[p])) * metric2[p]) * metric1[p];
                }
            }
        }
    }
    for (int j = jmin; j <= jmax; j++) {
        for (int i = imin; i <= imax; i++) {
            const int p = j * nx_ + i;
            if (tmask[p] == 1) {
                double hfo = h[p] * f[p];
                h[p] = h[p] - (dt * (QU[p] - QU[p - s])) * iarea[p];
                double advn = (flux[p] - flux[p - s]) * iarea[p];
                f[p] = (hfo - dt * advn) / h[p];
            }
        }
    }
}

int oracle_advection_uv(int direction, int scheme, int nx, int ny, int nk, int halox, int haloy,
                        const int* uvmask, const double* metric1, const double* metric2,
                        const double* huv, const double* uv, const int* tmask, const double* iarea,
                        const double* ah, double dt, double* h, double* var)
{
    if ((direction != 1 && direction != 2) || !valid_scheme(scheme) || nx < 1 || ny < 1 || nk < 1 ||
        halox < 2 || haloy < 2)
        return 1;
    const int nx_ = nx + 2 * halox, ny_ = ny + 2 * haloy;
    const size_t n2 = (size_t)nx_ * ny_;
    double* QU = calloc(n2, sizeof(double));
    double* flux = calloc(n2, sizeof(double));
    if (!QU || !flux) { free(QU); free(flux); return 2; }
    for (int k = 0; k < nk; k++)
        sweep_layer(direction, scheme, nx_, ny_, halox, haloy, uvmask, metric1, metric2,
                    huv + k * n2, uv + k * n2, tmask, iarea, ah, dt, h + k * n2, var + k * n2, QU, flux);
    free(QU);
    free(flux);
    return 0;
}

/* One column of the z sweep; h and f are strided by n2 (layer stride). */
static int w_column(int scheme, int nz, size_t n2, const double* w, const double* w_var, double dt,
                    double* h, double* f, double* flux, double* cfl0, double* cfls, double* haux,
                    double* faux)
{
    const int itersmax = 200;
    const int km = nz; /* Fortran kmax; Fortran index k -> C k-1 for layers, k for interfaces */
    flux[0] = 0.0;
    flux[km] = 0.0;
    for (int k = 1; k <= km - 1; k++)
        cfl0[k] = (fabs(w_var[k * n2]) * dt) / fmin2(h[(k - 1) * n2], h[k * n2]);

    int iters = 1;
    double itersm1 = 1.0, dtk = dt;
    int it = 1;
    while (it <= iters) {
        if (it == 1) {
            for (int k = 1; k <= km; k++) {
                haux[k] = h[(k - 1) * n2];
                faux[k] = f[(k - 1) * n2];
            }
            for (int k = 1; k <= km - 1; k++) cfls[k] = cfl0[k];
        } else {
            for (int k = 1; k <= km - 1; k++)
                cfls[k] = (fabs(w_var[k * n2]) * dt) / fmin2(haux[k], haux[k + 1]);
        }
        if (iters < itersmax) {
            /* MAXVAL of an empty array is -HUGE */
            double cfl = -1.7976931348623157e308;
            for (int k = 1; k <= km - 1; k++) cfl = fmax2(cfl, cfls[k]);
            int c = fceiling(cfl);
            int iters_new = c > 1 ? c : 1;
            if (iters_new > iters) {
                iters = iters_new > itersmax ? itersmax : iters_new;
                itersm1 = 1.0 / iters;
                dtk = dt * itersm1;
                if (it > 1) {
                    it = 1;
                    continue;
                }
            }
        }
        for (int k = 1; k <= km - 1; k++) {
            double cfl = cfls[k] * itersm1;
            double fu, fuu, fd;
            if (w_var[k * n2] > 0) {
                fu = faux[k];
                fuu = k > 1 ? faux[k - 1] : fu;
                fd = faux[k + 1];
            } else {
                fu = faux[k + 1];
                fuu = k < km - 1 ? faux[k + 2] : fu;
                fd = faux[k];
            }
            fu = limited_face_value(scheme, cfl, fu, fuu, fd);
            flux[k] = w_var[k * n2] * fu;
        }
        const int last = it == iters;
        for (int k = 1; k <= km; k++) {
            double hfo = haux[k] * faux[k];
            double hnew = haux[k] - dtk * (w[k * n2] - w[(k - 1) * n2]);
            double advn = flux[k] - flux[k - 1];
            double fnew = (hfo - dtk * advn) / hnew;
            if (last) {
                h[(k - 1) * n2] = hnew;
                f[(k - 1) * n2] = fnew;
            } else {
                haux[k] = hnew;
                faux[k] = fnew;
            }
        }
        it = it + 1;
    }
    return iters;
}

int oracle_advection_w(int scheme, int nx, int ny, int nz, int halox, int haloy,
                       const double* w, const double* w_var, const int* tmask, double dt,
                       double* h, double* var, int* max_iters)
{
    if (!valid_scheme(scheme) || nx < 1 || ny < 1 || nz < 1 || halox < 0 || haloy < 0) return 1;
    const int nx_ = nx + 2 * halox, ny_ = ny + 2 * haloy;
    const size_t n2 = (size_t)nx_ * ny_;
    double* buf = calloc((size_t)5 * (nz + 2), sizeof(double));
    if (!buf) return 2;
    double *flux = buf, *cfl0 = buf + (nz + 2), *cfls = buf + 2 * (nz + 2), *haux = buf + 3 * (nz + 2),
           *faux = buf + 4 * (nz + 2);
    int mi = 0;
    for (int j = haloy; j < ny_ - haloy; j++) {
        for (int i = halox; i < nx_ - halox; i++) {
            const size_t p = (size_t)j * nx_ + i;
            if (tmask[p] != 1) continue;
            int it = w_column(scheme, nz, n2, w + p, w_var + p, dt, h + p, var + p, flux, cfl0, cfls,
                              haux, faux);
            if (it > mi) mi = it;
        }
    }
    free(buf);
    if (max_iters) *max_iters = mi;
    return 0;
}

int oracle_vdiff_prepare(int nx, int ny, int nz, int halox, int haloy, double molecular,
                         const double* nuh, double dt, double cnpar, const double* dzo,
                         const double* dzn, double* auxo, double* auxn)
{
    if (nx < 1 || ny < 1 || nz < 1 || halox < 0 || haloy < 0) return 1;
    if (nz == 1) return 0;
    const int nx_ = nx + 2 * halox, ny_ = ny + 2 * haloy;
    const size_t n2 = (size_t)nx_ * ny_;
    for (int k = 0; k < nz - 1; k++) {
        for (int j = haloy; j < ny_ - haloy; j++) {
            for (int i = halox; i < nx_ - halox; i++) {
                const size_t p = (size_t)j * nx_ + i;
                /* nuh passed as nuh(:,:,1:nz-1): layer-index k uses interface k+1 */
                double x = (2.0 * dt) * (nuh[(k + 1) * n2 + p] + molecular);
                auxo[k * n2 + p] = ((1.0 - cnpar) * x) / (dzo[(k + 1) * n2 + p] + dzo[k * n2 + p]);
                auxn[k * n2 + p] = (cnpar * x) / (dzn[(k + 1) * n2 + p] + dzn[k * n2 + p]);
            }
        }
    }
    return 0;
}

int oracle_vdiff_apply(int nx, int ny, int nz, int halox, int haloy, const int* mask,
                       const double* dzo, const double* dzn, const double* auxo,
                       const double* auxn, double* var, const double* ea2, const double* ea4)
{
    if (nx < 1 || ny < 1 || nz < 1 || halox < 0 || haloy < 0) return 1;
    const int nx_ = nx + 2 * halox, ny_ = ny + 2 * haloy;
    const size_t n2 = (size_t)nx_ * ny_;
    if (nz == 1) {
        /* diffusion.F90:224-231 operates on the whole array, halos and land included */
        for (size_t p = 0; p < n2; p++) {
            double a4 = dzo[p] * var[p];
            double a2 = dzn[p];
            if (ea4) a4 = a4 + ea4[p];
            if (ea2) a2 = a2 - ea2[p];
            var[p] = a4 / a2;
        }
        return 0;
    }
    double* buf = malloc((size_t)4 * nz * sizeof(double));
    if (!buf) return 2;
    double *a2 = buf, *a4 = buf + nz, *ru = buf + 2 * nz, *qu = buf + 3 * nz;
    for (int j = haloy; j < ny_ - haloy; j++) {
        for (int i = halox; i < nx_ - halox; i++) {
            const size_t p = (size_t)j * nx_ + i;
            if (mask[p] != 1) continue;
            const double* c = var + p;
#define V(k) c[(size_t)(k) * n2]
#define AO(k) auxo[(size_t)(k) * n2 + p]
#define AN(k) auxn[(size_t)(k) * n2 + p]
            const int kt = nz - 1;
            a2[kt] = dzn[kt * n2 + p] + AN(kt - 1);
            a4[kt] = V(kt) * dzo[kt * n2 + p] + (V(kt - 1) - V(kt)) * AO(kt - 1);
            for (int k = 1; k < kt; k++) {
                a2[k] = (dzn[k * n2 + p] + AN(k)) + AN(k - 1);
                a4[k] = (V(k) * dzo[k * n2 + p] + (V(k + 1) - V(k)) * AO(k)) + (V(k - 1) - V(k)) * AO(k - 1);
            }
            a2[0] = dzn[p] + AN(0);
            a4[0] = V(0) * dzo[p] + (V(1) - V(0)) * AO(0);
            if (ea2)
                for (int k = 0; k < nz; k++) a2[k] = a2[k] - ea2[k * n2 + p];
            if (ea4)
                for (int k = 0; k < nz; k++) a4[k] = a4[k] + ea4[k * n2 + p];
            /* a1(k) = -auxn(k-1) (k >= 1), a3(k) = -auxn(k) (k <= nz-2) */
            ru[kt] = (-AN(kt - 1)) / a2[kt];
            qu[kt] = a4[kt] / a2[kt];
            for (int k = kt - 1; k >= 1; k--) {
                double den = a2[k] - (-AN(k)) * ru[k + 1];
                ru[k] = (-AN(k - 1)) / den;
                qu[k] = (a4[k] - (-AN(k)) * qu[k + 1]) / den;
            }
            qu[0] = (a4[0] - (-AN(0)) * qu[1]) / (a2[0] - (-AN(0)) * ru[1]);
            double* vo = var + p;
            vo[0] = qu[0];
            for (int k = 1; k <= kt; k++) vo[(size_t)k * n2] = qu[k] - ru[k] * vo[(size_t)(k - 1) * n2];
#undef V
#undef AO
#undef AN
        }
    }
    free(buf);
    return 0;
}

int oracle_transport_step(int scheme, int nx, int ny, int nz, int halox, int haloy,
                          const int* tmask, const int* umask, const int* vmask,
                          const double* idxu, const double* dyu, const double* dxv,
                          const double* idyv, const double* iarea,
                          const double* ho, const double* hn, const double* hu, const double* hv,
                          const double* u, const double* v, const double* w, const double* nuh,
                          const double* ah_u, const double* ah_v, double dt, double cnpar,
                          double molecular, int ntracers, double* tracers, double* h_out)
{
    if (!valid_scheme(scheme) || ntracers < 0 || nx < 1 || ny < 1 || nz < 1) return 1;
    const int nx_ = nx + 2 * halox, ny_ = ny + 2 * haloy;
    const size_t n3 = (size_t)nx_ * ny_ * nz;
    double* cur = malloc(n3 * sizeof(double));
    double* h = malloc(n3 * sizeof(double));
    double* auxo = calloc(n3, sizeof(double));
    double* auxn = calloc(n3, sizeof(double));
    int rc = 0;
    if (!cur || !h || !auxo || !auxn) { rc = 2; goto out; }
    memcpy(cur, ho, n3 * sizeof(double));
    const double half = 0.5 * dt;
    /* ufirst = False: adv1 = v_3d, adv2 = u_3d (operators.py:159-162) */
    for (int stage = 0; stage < 5; stage++) {
        for (int t = 0; t < ntracers; t++) {
            double* var = tracers + (size_t)t * n3;
            memcpy(h, cur, n3 * sizeof(double));
            if (stage == 0 || stage == 4)
                rc = oracle_advection_uv(2, scheme, nx, ny, nz, halox, haloy, vmask, dxv, idyv, hv, v,
                                         tmask, iarea, ah_v, half, h, var);
            else if (stage == 1 || stage == 3)
                rc = oracle_advection_uv(1, scheme, nx, ny, nz, halox, haloy, umask, idxu, dyu, hu, u,
                                         tmask, iarea, ah_u, half, h, var);
            else
                rc = oracle_advection_w(scheme, nx, ny, nz, halox, haloy, w, w, tmask, dt, h, var, NULL);
            if (rc) goto out;
        }
        if (ntracers > 0) memcpy(cur, h, n3 * sizeof(double));
    }
    if (h_out) memcpy(h_out, cur, n3 * sizeof(double));
    if (ntracers > 0) {
        rc = oracle_vdiff_prepare(nx, ny, nz, halox, haloy, molecular, nuh, dt, cnpar, hn, hn, auxo, auxn);
        if (rc) goto out;
        for (int t = 0; t < ntracers; t++) {
            rc = oracle_vdiff_apply(nx, ny, nz, halox, haloy, tmask, hn, hn, auxo, auxn,
                                    tracers + (size_t)t * n3, NULL, NULL);
            if (rc) goto out;
        }
    }
out:
    free(cur);
    free(h);
    free(auxo);
    free(auxn);
    return rc;
}

/* python/src/momentum.F90:191-216.  idt = 1/dt once (:201); the loop order k, j, i
 * (:202-204) does not matter for the result since w_k depends only on w_{k-1} of
 * the same point; the expression keeps the Fortran's left-to-right order. */
int oracle_w_momentum_3d(int nx, int ny, int nz, int imin, int imax, int jmin, int jmax, const int* mask,
                         const double* dyu, const double* dxv, const double* iarea, const double* ho,
                         const double* hn, const double* pk, const double* qk, double dt, double* w)
{
    if (nx < 1 || ny < 1 || nz < 1 || imin < 1 || jmin < 1 || imax >= nx || jmax >= ny) return 1;
    const size_t n2 = (size_t)nx * ny;
    const double idt = 1.0 / dt;
    for (int k = 1; k <= nz - 1; k++) {
        const size_t L = (size_t)(k - 1) * n2;  /* layer below interface k */
        for (int j = jmin; j <= jmax; j++) {
            for (int i = imin; i <= imax; i++) {
                const size_t p = (size_t)j * nx + i;
                if (mask[p] != 1) continue;
                w[(size_t)k * n2 + p] =
                    w[(size_t)(k - 1) * n2 + p] - (hn[L + p] - ho[L + p]) * idt -
                    (pk[L + p] * dyu[p] - pk[L + p - 1] * dyu[p - 1] + qk[L + p] * dxv[p] - qk[L + p - nx] * dxv[p - nx]) *
                        iarea[p];
            }
        }
    }
    return 0;
}

/* python/src/wrapper.F90:447-486: the activity test over the interior of every
 * layer (:461-467), then the k loop (outermost in the reference; the order of
 * the (i, j) loops does not matter, each column's updates only touch itself). */
int oracle_vertical_advection_to_sources(int nx, int ny, int nz, int halox, int haloy, const int* mask,
                                         const double* c, const double* w, const double* h, double* s)
{
    if (nx < 1 || ny < 1 || nz < 1 || halox < 0 || haloy < 0) return 1;
    const size_t n2 = (size_t)nx * ny;
    int active = 0;
    for (int k = 0; k < nz && !active; k++)
        for (int j = haloy; j < ny - haloy && !active; j++)
            for (int i = halox; i < nx - halox; i++)
                if (w[(size_t)k * n2 + (size_t)j * nx + i] != 0.0) {
                    active = 1;
                    break;
                }
    if (!active) return 0;
    const double upward = -1.0;
    for (int k = 0; k < nz - 1; k++) {
        for (int j = haloy; j < ny - haloy; j++) {
            for (int i = halox; i < nx - halox; i++) {
                const size_t p0 = (size_t)k * n2 + (size_t)j * nx + i, p1 = p0 + n2;
                if (mask[p0] != 1 || mask[p1] != 1) continue;
                const double lw = upward * 0.5 * (w[p0] + w[p1]);
                const double flux = lw > 0.0 ? lw * c[p0] : lw * c[p1];
                s[p0] = s[p0] - flux / h[p0];
                s[p1] = s[p1] + flux / h[p1];
            }
        }
    }
    return 0;
}

/* python/src/wrapper.F90 c_elevation2depth: where (mask /= 0) D = max(H + z, Dmin) */
int oracle_elevation2depth(int n, const double* z, const double* H, double Dmin, const int* mask, double* D)
{
    if (n < 0) return 1;
    for (int p = 0; p < n; p++)
        if (mask[p] != 0) D[p] = fmax2(H[p] + z[p], Dmin);
    return 0;
}

/* python/src/wrapper.F90 c_thickness2vertical_coordinates: the Fortran's whole-array
 * where statements per k, restated per point (each point only reads itself). */
int oracle_thickness2vertical_coordinates(int nx, int ny, int nz, const int* mask, const double* H, const double* h,
                                          double* zc, double* zf)
{
    if (nx < 1 || ny < 1 || nz < 1) return 1;
    const size_t n2 = (size_t)nx * ny;
    for (size_t p = 0; p < n2; p++) {
        if (mask[p] == 0) continue;
        zc[p] = -H[p] + 0.5 * h[p];
        for (int k = 1; k < nz; k++)
            zc[(size_t)k * n2 + p] = zc[(size_t)(k - 1) * n2 + p] + 0.5 * (h[(size_t)(k - 1) * n2 + p] + h[(size_t)k * n2 + p]);
        zf[p] = -H[p];
        for (int k = 1; k <= nz; k++) zf[(size_t)k * n2 + p] = zf[(size_t)(k - 1) * n2 + p] + h[(size_t)(k - 1) * n2 + p];
    }
    return 0;
}

/* python/src/interpolation.F90:9-23 grid_interp_x */
int oracle_interp_x(int nx, int ny, int nz, const double* source, double* target, int ioffset)
{
    if (nx < 1 || ny < 1 || nz < 1 || ioffset < 0 || ioffset > 1) return 1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            const size_t row = ((size_t)k * ny + j) * nx;
            for (int i = 0; i < nx - 1; i++)
                target[row + i + ioffset] = 0.5 * (source[row + i] + source[row + i + 1]);
        }
    return 0;
}

/* python/src/interpolation.F90:25-39 grid_interp_y */
int oracle_interp_y(int nx, int ny, int nz, const double* source, double* target, int joffset)
{
    if (nx < 1 || ny < 1 || nz < 1 || joffset < 0 || joffset > 1) return 1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny - 1; j++)
            for (int i = 0; i < nx; i++) {
                const size_t p = ((size_t)k * ny + j) * nx + i;
                target[p + (size_t)joffset * nx] = 0.5 * (source[p] + source[p + nx]);
            }
    return 0;
}

/* python/src/interpolation.F90:41-55 grid_interp_z */
int oracle_interp_z(int nx, int ny, int nz1, int nz2, const double* source, double* target, int koffset)
{
    if (nx < 1 || ny < 1 || nz1 < 1 || nz2 < 1 || koffset < 0 || nz1 - 1 + koffset > nz2) return 1;
    const size_t n2 = (size_t)nx * ny;
    for (int k = 0; k < nz1 - 1; k++)
        for (size_t p = 0; p < n2; p++)
            target[(size_t)(k + koffset) * n2 + p] = 0.5 * (source[(size_t)k * n2 + p] + source[(size_t)(k + 1) * n2 + p]);
    return 0;
}

/* python/src/interpolation.F90:57-72 grid_interp_xy */
int oracle_interp_xy(int nx1, int ny1, int nx2, int ny2, int nz, const double* source, double* target, int ioffset,
                     int joffset)
{
    if (nx1 < 1 || ny1 < 1 || nx2 < 1 || ny2 < 1 || nz < 1 || ioffset < 0 || joffset < 0 ||
        nx1 - 1 + ioffset > nx2 || ny1 - 1 + joffset > ny2)
        return 1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny1 - 1; j++)
            for (int i = 0; i < nx1 - 1; i++) {
                const size_t p = ((size_t)k * ny1 + j) * nx1 + i;
                target[((size_t)k * ny2 + j + joffset) * nx2 + i + ioffset] =
                    0.25 * (source[p] + source[p + 1] + source[p + nx1] + source[p + nx1 + 1]);
            }
    return 0;
}

/* python/src/momentum.F90:323-340 c_reconstruct_transport_change */
int oracle_reconstruct_transport_change(int nx, int ny, int nz, int jmin, int jmax, double* un, const double* hn,
                                        const double* Uo, double timestep)
{
    if (nx < 1 || ny < 1 || nz < 1 || jmin < 1 || jmax > ny) return 1;
    const double idt = 1.0 / timestep;
    for (int k = 0; k < nz; k++)
        for (int j = jmin - 1; j <= jmax - 1; j++)
            for (int i = 0; i < nx; i++) {
                const size_t p = ((size_t)k * ny + j) * nx + i;
                un[p] = (un[p] * hn[p] - Uo[p]) * idt;
            }
    return 0;
}

/* python/src/wrapper.F90:428-435 c_alpha */
int oracle_alpha(int n, const double* D, double Dmin, double Dcrit, const int* mask, double* alpha)
{
    if (n < 0) return 1;
    for (int p = 0; p < n; p++)
        if (mask[p] != 0) alpha[p] = fmax2(0.0, fmin2(1.0, (D[p] - Dmin) / (Dcrit - Dmin)));
    return 0;
}

/* python/src/wrapper.F90:567-601 c_update_gvc */
int oracle_update_gvc(int nx, int ny, int nz, double dsigma, const double* dbeta, double Dgamma, int kk,
                      const double* D, const int* mask, double* h)
{
    if (nx < 1 || ny < 1 || nz < 1 || kk < 1 || kk > nz) return 1;
    const size_t n2 = (size_t)nx * ny;
    const double bref = dbeta[kk - 1];
    for (size_t p = 0; p < n2; p++) {
        if (mask[p] == 0) continue;
        const double alpha = (fmin2(Dgamma / D[p], 1.0) * dsigma - bref) / (dsigma - bref);
        for (int k = 0; k < nz; k++) h[(size_t)k * n2 + p] = D[p] * (dbeta[k] + alpha * (dsigma - dbeta[k]));
    }
    return 0;
}

/* python/pygetm/vertical_coordinates.py:106-113 Sigma.__call__ */
int oracle_update_sigma(int nx, int ny, int nz, const double* dga, const double* D, const int* mask, double* h)
{
    if (nx < 1 || ny < 1 || nz < 1) return 1;
    const size_t n2 = (size_t)nx * ny;
    for (int k = 0; k < nz; k++)
        for (size_t p = 0; p < n2; p++)
            if (mask[p] != 0) h[(size_t)k * n2 + p] = dga[k] * D[p];
    return 0;
}
