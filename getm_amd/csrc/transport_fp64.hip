// FP64 kernels of the tracer-transport step, bit-identical to the reference.
//
// Compiled with -ffp-contract=off: every multiply and add rounds on its own,
// divisions are IEEE correctly rounded, and the operation order restates the
// Fortran (file:line at each kernel), so results equal the reference's bits.
//
// Layout: C-order (nz, ny_, nx_), i fastest; one lane per (i, j) so a wave
// reads 64 consecutive i and every load/store coalesces.
#include "common.h"
#include "internal.h"

namespace getm {

// ---------------------------------------------------------------------------
// x / y half-sweep, one lane per cell, out of place on the tracer.
// advection.F90.template:23-121 (u2d) and :123-221 (v2d).
//
// The reference first computes every face flux of a layer, then updates the
// cells.  Here each lane recomputes its two faces; a face computed by two
// neighbouring lanes is evaluated with identical operands in identical order,
// so both see the same value as the reference's single evaluation.  Because
// neighbours read the old tracer, the tracer is written to a separate buffer
// (f_out); cells that the sweep does not update are copied through.  The
// thickness h of a cell depends only on its own faces, so it is updated in
// place.
// ---------------------------------------------------------------------------
template <int S, int DIR>
__device__ __forceinline__ void face_flux_f64(const int* __restrict__ uvmask, const double* __restrict__ m1,
                                              const double* __restrict__ m2, const double* __restrict__ huv,
                                              const double* __restrict__ vel, const double* __restrict__ ah,
                                              const double* __restrict__ f, size_t p2, size_t p3, ptrdiff_t pf,
                                              ptrdiff_t s2, ptrdiff_t s3, double dt, double& q, double& flux)
{
    if (!open_face(uvmask[p2])) {
        q = 0.0;
        flux = 0.0;
        return;
    }
    const double u = vel[p3];
    double cfl;
    if (DIR == 1) {  // QU = u*hu*dyu ; cfl = |u*dt*idxu|   (m1 = idxu, m2 = dyu)
        q = (u * huv[p3]) * m2[p2];
        cfl = fabs((u * dt) * m1[p2]);
    } else {         // QV = v*hv*dxv ; cfl = |v*dt*idyv|   (m1 = dxv, m2 = idyv)
        q = (u * huv[p3]) * m1[p2];
        cfl = fabs((u * dt) * m2[p2]);
    }
    // f[pf]: the tracer at the face's lower cell (f is the layer field itself,
    // pf = p3, or one row of it, pf the column), s3 its stride along the sweep
    double fu, fuu, fd;
    if (u > 0.0) {
        fu = f[pf];
        fuu = fu;
        if (open_face(uvmask[p2 - s2])) fuu = f[pf - s3];
        fd = f[pf + s3];
    } else {
        fu = f[pf + s3];
        fuu = fu;
        if (open_face(uvmask[p2 + s2])) fuu = f[pf + 2 * s3];
        fd = f[pf];
    }
    fu = face_value<S, double>(cfl, fu, fuu, fd);
    flux = q * fu;
    if (ah) {  // template:97-105 / 197-205
        const double df = f[pf + s3] - f[pf];
        if (DIR == 1)
            flux = flux - (((ah[p2] * huv[p3]) * df) * m1[p2]) * m2[p2];
        else
            flux = flux - (((ah[p2] * huv[p3]) * df) * m2[p2]) * m1[p2];
    }
}

template <int S, int DIR>
__global__ void __launch_bounds__(256) k_adv_uv_f64(
    int nx_, int ny_, int imin, int imax, int jmin, int jmax, const int* __restrict__ uvmask,
    const double* __restrict__ m1, const double* __restrict__ m2, const double* __restrict__ huv,
    const double* __restrict__ vel, const int* __restrict__ tmask, const double* __restrict__ iarea,
    const double* __restrict__ ah, double dt, double* __restrict__ h, const double* __restrict__ f,
    double* __restrict__ fout)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx_ || j >= ny_) return;
    const size_t n2 = (size_t)nx_ * ny_;
    const size_t p2 = (size_t)j * nx_ + i;
    const size_t p3 = (size_t)k * n2 + p2;
    const bool interior = i >= imin && i <= imax && j >= jmin && j <= jmax;
    if (!interior || tmask[p2] != 1) {
        fout[p3] = f[p3];
        return;
    }
    const ptrdiff_t s2 = DIR == 1 ? 1 : nx_;
    double q0, fl0, q1, fl1;
    face_flux_f64<S, DIR>(uvmask, m1, m2, huv, vel, ah, f, p2 - s2, p3 - s2, p3 - s2, s2, s2, dt, q0, fl0);
    face_flux_f64<S, DIR>(uvmask, m1, m2, huv, vel, ah, f, p2, p3, p3, s2, s2, dt, q1, fl1);
    const double hfo = h[p3] * f[p3];
    const double hnew = h[p3] - (dt * (q1 - q0)) * iarea[p2];
    h[p3] = hnew;
    const double advn = (fl1 - fl0) * iarea[p2];
    fout[p3] = (hfo - dt * advn) / hnew;
}

// x half-sweep in place: one wave per row (j, k), the row's old values first
// copied into LDS (every face of the row reads only that row), then the row's
// updated cells written over them; the arithmetic is k_adv_uv_f64's, operand for
// operand.  Rows of up to ADV_X_ROW cells.
constexpr int ADV_X_ROW = 2048;
template <int S>
__global__ void __launch_bounds__(256) k_adv_x_inplace_f64(
    int nx_, int ny_, int imin, int imax, int jmin, int jmax, const int* __restrict__ uvmask,
    const double* __restrict__ m1, const double* __restrict__ m2, const double* __restrict__ huv,
    const double* __restrict__ vel, const int* __restrict__ tmask, const double* __restrict__ iarea,
    const double* __restrict__ ah, double dt, double* __restrict__ h, double* __restrict__ f)
{
    extern __shared__ double s_rows[];  // [4][nx_]
    const int lane = threadIdx.x, w = threadIdx.y;
    const int j = blockIdx.x * 4 + w, k = blockIdx.y;
    const size_t n2 = (size_t)nx_ * ny_;
    double* row = s_rows + (size_t)w * nx_;
    const size_t r3 = (size_t)k * n2 + (size_t)j * nx_;  // the row's first cell
    if (j < ny_)
        for (int c = lane; c < nx_; c += 64) row[c] = f[r3 + c];
    __syncthreads();
    if (j < jmin || j > jmax) return;
    for (int i = lane; i < nx_; i += 64) {
        const size_t p2 = (size_t)j * nx_ + i, p3 = r3 + i;
        if (i < imin || i > imax || tmask[p2] != 1) continue;
        double q0, fl0, q1, fl1;
        face_flux_f64<S, 1>(uvmask, m1, m2, huv, vel, ah, row, p2 - 1, p3 - 1, i - 1, 1, 1, dt, q0, fl0);
        face_flux_f64<S, 1>(uvmask, m1, m2, huv, vel, ah, row, p2, p3, i, 1, 1, dt, q1, fl1);
        const double hfo = h[p3] * row[i];
        const double hnew = h[p3] - (dt * (q1 - q0)) * iarea[p2];
        h[p3] = hnew;
        const double advn = (fl1 - fl0) * iarea[p2];
        f[p3] = (hfo - dt * advn) / hnew;
    }
}

// In-place x half-sweep of nk layers (GETM_ERR_CAPACITY for rows longer than ADV_X_ROW).
int launch_adv_x_inplace_f64(int scheme, int nk, const Grid& t, const Grid& uv, const double* vel, const double* ah,
                             double dt, double* h, const double* huv, double* f, hipStream_t s)
{
    const int nx_ = t.nx_, ny_ = t.ny_;
    if (nx_ > ADV_X_ROW) return GETM_ERR_CAPACITY;
    if (!uv.mask || !uv.idx || !uv.dy || !t.mask || !t.iarea) return GETM_ERR_ARG;
    dim3 block(64, 4, 1), grid((ny_ + 3) / 4, nk, 1);
    const size_t shm = (size_t)4 * nx_ * sizeof(double);
    ProfileScope ps(GETM_K_ADV_UV_F64, s);
    GETM_SCHEME_SWITCH(scheme, {
        hipLaunchKernelGGL((k_adv_x_inplace_f64<S_>), grid, block, shm, s, nx_, ny_, t.imin, t.imax, t.jmin, t.jmax,
                           uv.mask, uv.idx, uv.dy, huv, vel, t.mask, t.iarea, ah, dt, h, f);
    });
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

// y half-sweep in place: tiles of 64 columns x ADV_Y_ROWS rows of a layer.  A
// tile's faces read two rows beyond it on each side, which the neighbouring
// tiles update, so k_snap_rows first copies the four rows around every tile
// boundary (their old values; 4 / ADV_Y_ROWS of the layer); each tile then
// stages its rows and those halo rows in LDS and writes its updated cells over
// the old ones.  The arithmetic is k_adv_uv_f64's, operand for operand.
constexpr int ADV_Y_ROWS = 32;
__global__ void __launch_bounds__(256) k_snap_rows(int nx_, int ny_, int nb, const double* __restrict__ f,
                                                   double* __restrict__ snap)
{
    // snap[k][b][q][i]: row ADV_Y_ROWS (b + 1) - 2 + q of layer k, q = 0..3
    const int i = blockIdx.x * 64 + threadIdx.x, q = threadIdx.y, b = blockIdx.y, k = blockIdx.z;
    const int r = ADV_Y_ROWS * (b + 1) - 2 + q;
    if (i >= nx_ || r >= ny_) return;
    snap[(((size_t)k * nb + b) * 4 + q) * nx_ + i] = f[(size_t)k * nx_ * ny_ + (size_t)r * nx_ + i];
}
template <int S>
__global__ void __launch_bounds__(256) k_adv_y_inplace_f64(
    int nx_, int ny_, int imin, int imax, int jmin, int jmax, const int* __restrict__ uvmask,
    const double* __restrict__ m1, const double* __restrict__ m2, const double* __restrict__ huv,
    const double* __restrict__ vel, const int* __restrict__ tmask, const double* __restrict__ iarea,
    const double* __restrict__ ah, double dt, double* __restrict__ h, double* __restrict__ f, int nb,
    const double* __restrict__ snap)
{
    constexpr int NR = ADV_Y_ROWS + 4;
    __shared__ double s_t[NR * 64];  // rows j0 - 2 .. j0 + ADV_Y_ROWS + 1, 64 columns each
    const int lane = threadIdx.x, w = threadIdx.y;
    const int i = blockIdx.x * 64 + lane, ty = blockIdx.y, k = blockIdx.z;
    const int j0 = ty * ADV_Y_ROWS;
    const size_t n2 = (size_t)nx_ * ny_;
    if (i < nx_) {
        for (int rr = w; rr < NR; rr += 4) {
            const int j = j0 - 2 + rr;
            if (j < 0 || j >= ny_) continue;
            double x;
            if (j < j0)  // from the tile above: rows 0, 1 of boundary ty
                x = snap[(((size_t)k * nb + (ty - 1)) * 4 + (j - (j0 - 2))) * nx_ + i];
            else if (j >= j0 + ADV_Y_ROWS)  // from the tile below: rows 2, 3 of boundary ty + 1
                x = snap[(((size_t)k * nb + ty) * 4 + (j - (j0 + ADV_Y_ROWS - 2))) * nx_ + i];
            else
                x = f[(size_t)k * n2 + (size_t)j * nx_ + i];
            s_t[rr * 64 + lane] = x;
        }
    }
    __syncthreads();
    if (i >= nx_ || i < imin || i > imax) return;
    for (int rr = 2 + w; rr < 2 + ADV_Y_ROWS; rr += 4) {
        const int j = j0 - 2 + rr;
        if (j < jmin || j > jmax) continue;
        const size_t p2 = (size_t)j * nx_ + i, p3 = (size_t)k * n2 + p2;
        if (tmask[p2] != 1) continue;
        const ptrdiff_t s2 = nx_;
        double q0, fl0, q1, fl1;
        face_flux_f64<S, 2>(uvmask, m1, m2, huv, vel, ah, s_t, p2 - s2, p3 - s2, (rr - 1) * 64 + lane, s2, 64, dt,
                            q0, fl0);
        face_flux_f64<S, 2>(uvmask, m1, m2, huv, vel, ah, s_t, p2, p3, rr * 64 + lane, s2, 64, dt, q1, fl1);
        const double hfo = h[p3] * s_t[rr * 64 + lane];
        const double hnew = h[p3] - (dt * (q1 - q0)) * iarea[p2];
        h[p3] = hnew;
        const double advn = (fl1 - fl0) * iarea[p2];
        f[p3] = (hfo - dt * advn) / hnew;
    }
}

// In-place y half-sweep of nk layers; snap: a work array of at least
// adv_y_snap_doubles(t, nk) doubles.
size_t adv_y_snap_doubles(const Grid& t, int nk)
{
    const int nty = (t.ny_ + ADV_Y_ROWS - 1) / ADV_Y_ROWS;
    return (size_t)nk * (nty > 1 ? nty - 1 : 0) * 4 * t.nx_;
}
int launch_adv_y_inplace_f64(int scheme, int nk, const Grid& t, const Grid& uv, const double* vel, const double* ah,
                             double dt, double* h, const double* huv, double* f, double* snap, hipStream_t s)
{
    const int nx_ = t.nx_, ny_ = t.ny_;
    if (!uv.mask || !uv.dx || !uv.idy || !t.mask || !t.iarea) return GETM_ERR_ARG;
    const int ntx = (nx_ + 63) / 64, nty = (ny_ + ADV_Y_ROWS - 1) / ADV_Y_ROWS, nb = nty - 1;
    if (nty > 65535 || nk > 65535) return GETM_ERR_CAPACITY;
    ProfileScope ps(GETM_K_ADV_UV_F64, s);
    if (nb > 0)
        hipLaunchKernelGGL(k_snap_rows, dim3(ntx, nb, nk), dim3(64, 4, 1), 0, s, nx_, ny_, nb, f, snap);
    GETM_SCHEME_SWITCH(scheme, {
        hipLaunchKernelGGL((k_adv_y_inplace_f64<S_>), dim3(ntx, nty, nk), dim3(64, 4, 1), 0, s, nx_, ny_, t.imin,
                           t.imax, t.jmin, t.jmax, uv.mask, uv.dx, uv.idy, huv, vel, t.mask, t.iarea, ah, dt, h, f, nb,
                           snap);
    });
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

int launch_adv_uv_f64(int direction, int scheme, int nk, const Grid& t, const Grid& uv, const double* vel,
                      const double* ah, double dt, double* h, const double* huv, const double* f_in,
                      double* f_out, hipStream_t s)
{
    const int nx_ = t.nx_, ny_ = t.ny_;
    dim3 block(64, 4, 1);
    dim3 grid((nx_ + 63) / 64, (ny_ + 3) / 4, nk);
    // direction 1 reads (idxu, dyu) from the U grid, direction 2 (dxv, idyv) from the V grid
    const double* m1 = direction == 1 ? uv.idx : uv.dx;
    const double* m2 = direction == 1 ? uv.dy : uv.idy;
    if (!uv.mask || !m1 || !m2 || !t.mask || !t.iarea) return GETM_ERR_ARG;
    ProfileScope ps(GETM_K_ADV_UV_F64, s);
    GETM_SCHEME_SWITCH(scheme, {
        if (direction == 1)
            hipLaunchKernelGGL((k_adv_uv_f64<S_, 1>), grid, block, 0, s, nx_, ny_, t.imin, t.imax, t.jmin,
                               t.jmax, uv.mask, m1, m2, huv, vel, t.mask, t.iarea, ah, dt, h, f_in, f_out);
        else
            hipLaunchKernelGGL((k_adv_uv_f64<S_, 2>), grid, block, 0, s, nx_, ny_, t.imin, t.imax, t.jmin,
                               t.jmax, uv.mask, m1, m2, huv, vel, t.mask, t.iarea, ah, dt, h, f_in, f_out);
    });
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

// ---------------------------------------------------------------------------
// z sweep, one lane per water column, in place.
// advection.F90.template:223-374.
//
// Iteration control is the reference's: iters = max(1, ceil(max cfl)) capped at
// 200, and a restart from the original column whenever a later sub-step needs
// more iterations.  Intermediate sub-step states live in wsh/wsf (same layout
// as h, f); the first sub-step reads h, f and the last writes them, so the
// common single-iteration case touches nothing else.  Within a sub-step the
// column is swept bottom to top with a 4-value register window, which lets the
// update write layer k while the old values of k-1..k+2 are still held.
// ---------------------------------------------------------------------------
template <int S>
__global__ void __launch_bounds__(256) k_adv_w_f64(
    int nx_, int ny_, int nz, int imin, int imax, int jmin, int jmax, const int* __restrict__ tmask,
    const double* __restrict__ w, const double* __restrict__ wv, double dt, double* __restrict__ h,
    double* __restrict__ f, double* __restrict__ wsh, double* __restrict__ wsf, int* __restrict__ stats)
{
    const int i = imin + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = jmin + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > imax || j > jmax) return;
    const size_t n2 = (size_t)nx_ * ny_;
    const size_t p = (size_t)j * nx_ + i;
    if (tmask[p] != 1) return;

    int iters = 1, capped = 0;
    double itersm1 = 1.0, dtk = dt;
    int it = 1;
    while (it <= iters) {
        const double* sh = (it == 1 ? h : wsh) + p;
        const double* sf = (it == 1 ? f : wsf) + p;
        if (iters < kItersMax) {
            double c = -1.7976931348623157e308;  // MAXVAL of an empty array
            double hk = sh[0];
            for (int k = 1; k <= nz - 1; k++) {
                const double hk1 = sh[(size_t)k * n2];
                c = fmax2(c, (fabs(wv[(size_t)k * n2 + p]) * dt) / fmin2(hk, hk1));
                hk = hk1;
            }
            const int ce = fceiling(c);
            const int iters_new = ce > 1 ? ce : 1;
            if (iters_new > iters) {
                if (iters_new > kItersMax) capped = 1;
                iters = iters_new > kItersMax ? kItersMax : iters_new;
                itersm1 = 1.0 / (double)iters;
                dtk = dt * itersm1;
                if (it > 1) {
                    it = 1;
                    continue;
                }
            }
        }
        const bool last = it == iters;
        double* dh = (last ? h : wsh) + p;
        double* df = (last ? f : wsf) + p;
        // window over Fortran layers: fm1 = faux(k-1), f0 = faux(k), f1 = faux(k+1), h0/h1 likewise
        double fm1 = 0.0, f0 = sf[0], h0 = sh[0];
        double f1 = nz > 1 ? sf[n2] : 0.0, h1 = nz > 1 ? sh[n2] : 0.0;
        double flux_prev = 0.0, wprev = w[p];
        for (int k = 1; k <= nz; k++) {
            double f2 = 0.0, h2 = 0.0;
            if (k + 2 <= nz) {
                f2 = sf[(size_t)(k + 1) * n2];
                h2 = sh[(size_t)(k + 1) * n2];
            }
            double fluxk = 0.0;
            if (k <= nz - 1) {
                const double wvk = wv[(size_t)k * n2 + p];
                const double cfl = ((fabs(wvk) * dt) / fmin2(h0, h1)) * itersm1;
                double fu, fuu, fd;
                if (wvk > 0.0) {
                    fu = f0;
                    fuu = k > 1 ? fm1 : fu;
                    fd = f1;
                } else {
                    fu = f1;
                    fuu = k < nz - 1 ? f2 : fu;
                    fd = f0;
                }
                fu = face_value<S, double>(cfl, fu, fuu, fd);
                fluxk = wvk * fu;
            }
            const double wk = w[(size_t)k * n2 + p];
            const double hfo = h0 * f0;
            const double hnew = h0 - dtk * (wk - wprev);
            const double advn = fluxk - flux_prev;
            dh[(size_t)(k - 1) * n2] = hnew;
            df[(size_t)(k - 1) * n2] = (hfo - dtk * advn) / hnew;
            fm1 = f0; f0 = f1; f1 = f2;
            h0 = h1; h1 = h2;
            flux_prev = fluxk;
            wprev = wk;
        }
        it = it + 1;
    }
    if (iters > 1) atomicMax(&stats[0], iters);
    if (capped) atomicAdd(&stats[1], 1);
}

int launch_adv_w_f64(int scheme, const Grid& t, const double* w, const double* wv, double dt, double* h,
                     double* f, double* wsh, double* wsf, int* stats, hipStream_t s)
{
    const int ni = t.imax - t.imin + 1, nj = t.jmax - t.jmin + 1;
    dim3 block(64, 4, 1);
    dim3 grid((ni + 63) / 64, (nj + 3) / 4, 1);
    if (!t.mask) return GETM_ERR_ARG;
    ProfileScope ps(GETM_K_ADV_W_F64, s);
    GETM_SCHEME_SWITCH(scheme, {
        hipLaunchKernelGGL((k_adv_w_f64<S_>), grid, block, 0, s, t.nx_, t.ny_, t.nz, t.imin, t.imax, t.jmin,
                           t.jmax, t.mask, w, wv, dt, h, f, wsh, wsf, stats);
    });
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

// ---------------------------------------------------------------------------
// theta-method vertical diffusion.  diffusion.F90:135-314.
// ---------------------------------------------------------------------------

// auxo/auxn at interior points, levels 0..nz-2 (interface k+1).  diffusion.F90:163-171.
__global__ void __launch_bounds__(256) k_vdiff_prepare_f64(
    int nx_, int ny_, int nz, int imin, int imax, int jmin, int jmax, double molecular,
    const double* __restrict__ nuh, double dt, double cnpar, const double* __restrict__ dzo,
    const double* __restrict__ dzn, double* __restrict__ auxo, double* __restrict__ auxn)
{
    const int i = imin + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = jmin + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > imax || j > jmax) return;
    const size_t n2 = (size_t)nx_ * ny_;
    const size_t p = (size_t)k * n2 + (size_t)j * nx_ + i;
    const double x = (2.0 * dt) * (nuh[p + n2] + molecular);
    auxo[p] = ((1.0 - cnpar) * x) / (dzo[p + n2] + dzo[p]);
    auxn[p] = (cnpar * x) / (dzn[p + n2] + dzn[p]);
}

int launch_vdiff_prepare_f64(const Grid& t, int nz, double molecular, const double* nuh, double dt,
                             double cnpar, const double* dzo, const double* dzn, double* auxo, double* auxn,
                             hipStream_t s)
{
    if (nz <= 1) return GETM_OK;  // diffusion.F90:155
    const int ni = t.imax - t.imin + 1, nj = t.jmax - t.jmin + 1;
    dim3 block(64, 4, 1);
    dim3 grid((ni + 63) / 64, (nj + 3) / 4, nz - 1);
    ProfileScope ps(GETM_K_VDIFF_PREPARE_F64, s);
    hipLaunchKernelGGL(k_vdiff_prepare_f64, grid, block, 0, s, t.nx_, t.ny_, nz, t.imin, t.imax, t.jmin,
                       t.jmax, molecular, nuh, dt, cnpar, dzo, dzn, auxo, auxn);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

// Matrix + Thomas solve per mask==1 column (diffusion.F90:238-307), with
// a1(k) = -auxn(k-1), a3(k) = -auxn(k).  Elimination runs top to bottom as in
// the reference; qu(k) overwrites var(k) once its old value is no longer
// needed and ru(k) goes to scratch, then back-substitution runs bottom to top.
__global__ void __launch_bounds__(256) k_vdiff_apply_f64(
    int nx_, int ny_, int nz, int imin, int imax, int jmin, int jmax, const int* __restrict__ mask,
    const double* __restrict__ dzo, const double* __restrict__ dzn, const double* __restrict__ auxo,
    const double* __restrict__ auxn, double* __restrict__ var, const double* __restrict__ ea2,
    const double* __restrict__ ea4, double* __restrict__ ru)
{
    const int i = imin + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = jmin + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > imax || j > jmax) return;
    const size_t n2 = (size_t)nx_ * ny_;
    const size_t p = (size_t)j * nx_ + i;
    if (mask[p] != 1) return;
    const int kt = nz - 1;
#define AT(a, k) (a)[(size_t)(k) * n2 + p]
    // surface layer
    double vk1 = AT(var, kt);      // var(k+1) in the loop below
    double vk = AT(var, kt - 1);   // var(k)
    double a2 = AT(dzn, kt) + AT(auxn, kt - 1);
    double a4 = vk1 * AT(dzo, kt) + (vk - vk1) * AT(auxo, kt - 1);
    if (ea2) a2 = a2 - AT(ea2, kt);
    if (ea4) a4 = a4 + AT(ea4, kt);
    double rup = (-AT(auxn, kt - 1)) / a2;
    double qup = a4 / a2;
    AT(ru, kt) = rup;
    AT(var, kt) = qup;
    for (int k = kt - 1; k >= 1; k--) {
        const double vkm1 = AT(var, k - 1);
        const double ank = AT(auxn, k), ankm1 = AT(auxn, k - 1);
        a2 = (AT(dzn, k) + ank) + ankm1;
        a4 = (vk * AT(dzo, k) + (vk1 - vk) * AT(auxo, k)) + (vkm1 - vk) * AT(auxo, k - 1);
        if (ea2) a2 = a2 - AT(ea2, k);
        if (ea4) a4 = a4 + AT(ea4, k);
        const double den = a2 - (-ank) * rup;
        const double ruk = (-ankm1) / den;
        const double quk = (a4 - (-ank) * qup) / den;
        AT(ru, k) = ruk;
        AT(var, k) = quk;
        rup = ruk;
        qup = quk;
        vk1 = vk;
        vk = vkm1;
    }
    // bottom layer
    a2 = AT(dzn, 0) + AT(auxn, 0);
    a4 = vk * AT(dzo, 0) + (vk1 - vk) * AT(auxo, 0);
    if (ea2) a2 = a2 - AT(ea2, 0);
    if (ea4) a4 = a4 + AT(ea4, 0);
    double v = (a4 - (-AT(auxn, 0)) * qup) / (a2 - (-AT(auxn, 0)) * rup);
    AT(var, 0) = v;
    for (int k = 1; k <= kt; k++) {
        v = AT(var, k) - AT(ru, k) * v;
        AT(var, k) = v;
    }
#undef AT
}

// nz == 1: var = (dzo*var + ea4)/(dzn - ea2) on the whole padded array (diffusion.F90:224-231).
__global__ void __launch_bounds__(256) k_vdiff_single_f64(size_t n, const double* __restrict__ dzo,
                                                          const double* __restrict__ dzn, double* __restrict__ var,
                                                          const double* __restrict__ ea2,
                                                          const double* __restrict__ ea4)
{
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double a4 = dzo[p] * var[p];
    double a2 = dzn[p];
    if (ea4) a4 = a4 + ea4[p];
    if (ea2) a2 = a2 - ea2[p];
    var[p] = a4 / a2;
}

int launch_vdiff_apply_f64(const Grid& t, int nz, const int* mask, const double* dzo, const double* dzn,
                           const double* auxo, const double* auxn, double* var, const double* ea2,
                           const double* ea4, double* ru, hipStream_t s)
{
    ProfileScope ps(GETM_K_VDIFF_APPLY_F64, s);
    if (nz == 1) {
        const size_t n = t.n2();
        hipLaunchKernelGGL(k_vdiff_single_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, dzo, dzn, var,
                           ea2, ea4);
    } else {
        const int ni = t.imax - t.imin + 1, nj = t.jmax - t.jmin + 1;
        dim3 block(64, 4, 1);
        dim3 grid((ni + 63) / 64, (nj + 3) / 4, 1);
        hipLaunchKernelGGL(k_vdiff_apply_f64, grid, block, 0, s, t.nx_, t.ny_, nz, t.imin, t.imax, t.jmin, t.jmax,
                           mask, dzo, dzn, auxo, auxn, var, ea2, ea4, ru);
    }
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}


// Batched FP64 diffusion of the transport step (dzo = dzn = hn, no ea2/ea4):
// TracerCollection._diffuse (tracer.py:248-276) for every tracer, out of place
// (the stage-5 scratch in, the tracers out), one lane per column and group of
// up to TPL tracers.  auxo/auxn are formed inline with the expressions of
// k_vdiff_prepare_f64 and the solve is k_vdiff_apply_f64's, so results equal
// prepare + per-tracer apply bit for bit.  The matrix does not depend on the
// tracer: its coefficients, the divisions that form them and ru are computed
// once per column for the group (the same operands, so the same bits as per
// tracer), and only the right sides q and the back-substitution are per
// tracer.  The elimination stores ru(k) over the first tracer's scratch value
// var(k), which it no longer needs (it is in a register), so tracers need no
// other scratch.  Columns that are not solved copy their values; nz = 1 applies
// the reference's (dz c)/dz to every point (diffusion.F90:224-231).
template <int TPL>
__global__ void __launch_bounds__(256) k_vdiff_batch_f64(
    int nx_, int ny_, int nz, int ntr, int imin, int imax, int jmin, int jmax, const int* __restrict__ mask,
    const double* __restrict__ hn, const double* __restrict__ nuh, double dt, double cnpar, double molecular,
    double* __restrict__ fin, double* __restrict__ fout)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= nx_ || j >= ny_) return;
    const size_t n2 = (size_t)nx_ * ny_, n3 = n2 * (size_t)nz;
    const size_t p = (size_t)j * nx_ + i;
    const bool solve = nz > 1 && i >= imin && i <= imax && j >= jmin && j <= jmax && mask[p] == 1;
    const int t0 = blockIdx.z * TPL;
    const int nt = ntr - t0 < TPL ? ntr - t0 : TPL;  // tracers of this group (block-uniform)
#define AT(a, k) (a)[(size_t)(k) * n2 + p]
    double* var[TPL];
    double* out[TPL];
#pragma unroll
    for (int u = 0; u < TPL; u++) {
        const int t = t0 + (u < nt ? u : 0);
        var[u] = fin + (size_t)t * n3;
        out[u] = fout + (size_t)t * n3;
    }
    double* ru = var[0];  // ru(k) replaces the first tracer's var(k) once the elimination has passed row k
    if (nz == 1) {
#pragma unroll
        for (int u = 0; u < TPL; u++)
            if (u < nt) out[u][p] = (hn[p] * var[u][p]) / hn[p];
        return;
    }
    if (!solve) {
#pragma unroll
        for (int u = 0; u < TPL; u++)
            if (u < nt)
                for (int k = 0; k < nz; k++) AT(out[u], k) = AT(var[u], k);
        return;
    }
    // auxo(k), auxn(k) of interface k+1 (k_vdiff_prepare_f64)
    auto aux = [&](int k, double& ao, double& an) {
        const double x = (2.0 * dt) * (AT(nuh, k + 1) + molecular);
        ao = ((1.0 - cnpar) * x) / (AT(hn, k + 1) + AT(hn, k));
        an = (cnpar * x) / (AT(hn, k + 1) + AT(hn, k));
    };
    const int kt = nz - 1;
    double ao_lo, an_lo;
    aux(kt - 1, ao_lo, an_lo);
    double vk1[TPL], vk[TPL], qup[TPL];
    double a2 = AT(hn, kt) + an_lo;
#pragma unroll
    for (int u = 0; u < TPL; u++) {
        if (u >= nt) continue;
        vk1[u] = AT(var[u], kt);
        vk[u] = AT(var[u], kt - 1);
        const double a4 = vk1[u] * AT(hn, kt) + (vk[u] - vk1[u]) * ao_lo;
        qup[u] = a4 / a2;
        AT(out[u], kt) = qup[u];
    }
    double rup = (-an_lo) / a2;
    AT(ru, kt) = rup;
    for (int k = kt - 1; k >= 1; k--) {
        const double ank = an_lo, aok = ao_lo;
        aux(k - 1, ao_lo, an_lo);
        a2 = (AT(hn, k) + ank) + an_lo;
        const double den = a2 - (-ank) * rup;
#pragma unroll
        for (int u = 0; u < TPL; u++) {
            if (u >= nt) continue;
            const double vkm1 = AT(var[u], k - 1);
            const double a4 = (vk[u] * AT(hn, k) + (vk1[u] - vk[u]) * aok) + (vkm1 - vk[u]) * ao_lo;
            const double quk = (a4 - (-ank) * qup[u]) / den;
            AT(out[u], k) = quk;
            qup[u] = quk;
            vk1[u] = vk[u];
            vk[u] = vkm1;
        }
        const double ruk = (-an_lo) / den;
        AT(ru, k) = ruk;
        rup = ruk;
    }
    a2 = AT(hn, 0) + an_lo;
    const double den0 = a2 - (-an_lo) * rup;
    double v[TPL];
#pragma unroll
    for (int u = 0; u < TPL; u++) {
        if (u >= nt) continue;
        const double a4 = vk[u] * AT(hn, 0) + (vk1[u] - vk[u]) * ao_lo;
        v[u] = (a4 - (-an_lo) * qup[u]) / den0;
        AT(out[u], 0) = v[u];
    }
    for (int k = 1; k <= kt; k++) {
        const double r = AT(ru, k);
#pragma unroll
        for (int u = 0; u < TPL; u++) {
            if (u >= nt) continue;
            v[u] = AT(out[u], k) - r * v[u];
            AT(out[u], k) = v[u];
        }
    }
#undef AT
}

int launch_vdiff_batch_f64(const Grid& t, int nz, const int* mask, const double* hn, const double* nuh, double dt,
                           double cnpar, double molecular, int ntr, double* fin, double* fout, hipStream_t s)
{
    if (ntr > 65535) return GETM_ERR_CAPACITY;
    ProfileScope ps(GETM_K_VDIFF_APPLY_F64, s);
    // two tracers per lane: one set of coefficients, divisions and ru per column for both
    constexpr int TPL = 2;
    dim3 block(64, 4, 1);
    dim3 grid((t.nx_ + 63) / 64, (t.ny_ + 3) / 4, (ntr + TPL - 1) / TPL);
    hipLaunchKernelGGL(k_vdiff_batch_f64<TPL>, grid, block, 0, s, t.nx_, t.ny_, nz, ntr, t.imin, t.imax, t.jmin,
                       t.jmax, mask, hn, nuh, dt, cnpar, molecular, fin, fout);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

}  // namespace getm
