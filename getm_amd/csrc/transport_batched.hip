// Batched tracer-transport step: T tracers share one velocity field, one
// layer-thickness evolution and one tridiagonal factorisation per column.
//
// Every kernel computes its tracer-independent quantities (face transports
// and CFL numbers, the FP64 layer-thickness update, sub-step counts, the
// diffusion factorisation) once per block and then streams all tracers
// through them, so HBM traffic per tracer is one read and one write per
// kernel.  Kernels are templated on the tracer storage/arithmetic type R:
//   R = double keeps the reference's FP64 operation order and is bit-identical
//          to the per-tracer path (and so to the reference);
//   R = float  stores tracers in FP32 and runs the stencil, limiter and
//          tridiagonal arithmetic in FP32 while the layer-thickness update stays
//          FP64; per-cell update factors are formed in FP64 and rounded once.
//
// Stage order (operators.py:124-133 with ufirst = False):
//   k_hpair<D1=y>: v dt/2, u dt/2   (stages 1-2)
//   k_wcol:        w dt             (stage 3)
//   k_hpair<D1=x>: u dt/2, v dt/2   (stages 4-5)
//   k_vdiff_batch: theta-method diffusion (tracer.py:248-276)
#include <type_traits>

#include "common.h"
#include "internal.h"

namespace getm {

// ---------------------------------------------------------------------------
// Horizontal stage pair on one layer tile.
// ---------------------------------------------------------------------------
constexpr int TX = 64, TY = 16;     // output tile
constexpr int HX = TX + 4, HY = TY + 4;  // input tile incl. halo 2
constexpr int NTHR = 256;

struct FaceC {
    double q;     // transport Q = vel*h*metric (FP64, also drives the h update)
    double cfl;   // |vel*dt*inverse metric|
    int flags;    // bit0 open, bit1 vel > 0, bit2 upstream-upstream face open
};

// Face (gi, gj) of direction DIR (1: U face east of T(gi,gj); 2: V face north of it).
template <int DIR>
__device__ __forceinline__ FaceC face_coef(const StepArgs& A, int gi, int gj, int k, double dt)
{
    FaceC fc;
    const size_t p = (size_t)gj * A.nx_ + gi;
    const int* mask = DIR == 1 ? A.umask : A.vmask;
    if (!open_face(mask[p])) {
        fc.q = 0.0;
        fc.cfl = 0.0;
        fc.flags = 0;
        return fc;
    }
    const size_t p3 = (size_t)k * A.n2 + p;
    const ptrdiff_t s = DIR == 1 ? 1 : A.nx_;
    double vel;
    if (DIR == 1) {
        vel = A.u[p3];
        fc.q = (vel * A.hu[p3]) * A.dyu[p];
        fc.cfl = fabs((vel * dt) * A.idxu[p]);
    } else {
        vel = A.v[p3];
        fc.q = (vel * A.hv[p3]) * A.dxv[p];
        fc.cfl = fabs((vel * dt) * A.idyv[p]);
    }
    const bool pos = vel > 0.0;
    const bool fuu_ok = pos ? open_face(mask[p - s]) : open_face(mask[p + s]);
    fc.flags = 1 | (pos ? 2 : 0) | (fuu_ok ? 4 : 0);
    return fc;
}

// Limited flux through a face from the tracer values of the four cells around
// it along the sweep direction: fm = lower-lower, f0 = lower, f1 = upper,
// fp = upper-upper.
template <int S, typename R>
__device__ __forceinline__ R face_flux(R q, R cfl, int flags, R fm, R f0, R f1, R fp)
{
    if (!(flags & 1)) return R(0);
    R fu, fuu, fd;
    if (flags & 2) {
        fu = f0;
        fuu = (flags & 4) ? fm : fu;
        fd = f1;
    } else {
        fu = f1;
        fuu = (flags & 4) ? fp : fu;
        fd = f0;
    }
    return q * face_value<S, R>(cfl, fu, fuu, fd);
}

// Horizontal diffusion flux correction (template:97-105, 197-205), read from
// global memory: only used when Ah is given.
template <int DIR, typename R>
__device__ __forceinline__ R ah_flux(const StepArgs& A, int gi, int gj, int k, R f0, R f1)
{
    const size_t p = (size_t)gj * A.nx_ + gi;
    const int* mask = DIR == 1 ? A.umask : A.vmask;
    if (!open_face(mask[p])) return R(0);
    const size_t p3 = (size_t)k * A.n2 + p;
    const double* ah = DIR == 1 ? A.ah_u : A.ah_v;
    const double hf = DIR == 1 ? A.hu[p3] : A.hv[p3];
    const double m1 = DIR == 1 ? A.idxu[p] : A.idyv[p];
    const double m2 = DIR == 1 ? A.dyu[p] : A.dxv[p];
    return (((R(ah[p] * hf)) * (f1 - f0)) * R(m1)) * R(m2);
}

// Cell coefficients of one stage: the cell is updated iff interior and wet.
template <typename R>
struct CellC {
    R q0, q1, c0, c1;      // face transports (as R) and CFLs
    int fl0, fl1;
    int upd;
    // FP64: hold, hnew, iA kept for the reference order (hfo - dt*advn)/hnew
    // FP32: a = hold/hnew, b = dt*iA/hnew, f_new = a*f - b*(fl1 - fl0)
    R a, b, iA;
};

template <int DIR, typename R>
__device__ __forceinline__ void cell_coef(const StepArgs& A, int gi, int gj, int k, double dt, double hold,
                                          CellC<R>& c, double& hnew)
{
    const bool inside = gi >= A.imin && gi <= A.imax && gj >= A.jmin && gj <= A.jmax;
    const size_t p = (size_t)gj * A.nx_ + gi;
    c.upd = inside && A.tmask[p] == 1;
    hnew = hold;
    if (!c.upd) return;
    const FaceC f0 = DIR == 1 ? face_coef<1>(A, gi - 1, gj, k, dt) : face_coef<2>(A, gi, gj - 1, k, dt);
    const FaceC f1 = face_coef<DIR>(A, gi, gj, k, dt);
    const double iA = A.iarea[p];
    hnew = hold - (dt * (f1.q - f0.q)) * iA;
    c.q0 = R(f0.q);
    c.q1 = R(f1.q);
    c.c0 = R(f0.cfl);
    c.c1 = R(f1.cfl);
    c.fl0 = f0.flags;
    c.fl1 = f1.flags;
    if constexpr (std::is_same<R, double>::value) {
        c.a = hold;
        c.b = hnew;
        c.iA = iA;
    } else {
        c.a = R(hold / hnew);
        c.b = R((dt * iA) / hnew);
        c.iA = R(iA);
    }
}

template <typename R>
__device__ __forceinline__ R cell_update(const CellC<R>& c, double dt, R f, R fl0, R fl1)
{
    if constexpr (std::is_same<R, double>::value) {
        const double hfo = c.a * f;
        const double advn = (fl1 - fl0) * c.iA;
        return (hfo - dt * advn) / c.b;
    } else {
        return c.a * f - c.b * (fl1 - fl0);
    }
}

// Local coordinates: tile-local (xl, yl), output tile [0,TX) x [0,TY), input
// tile [-2,TX+2) x [-2,TY+2) stored at s_in[yl+2][xl+2].
template <int S, typename R, int D1>
__global__ void __launch_bounds__(NTHR) k_hpair(const StepArgs A, double dt, int ntr, const R* __restrict__ fin,
                                                R* __restrict__ fout, const double* __restrict__ hin,
                                                double* __restrict__ hout)
{
    constexpr int D2 = 3 - D1;
    // region A (first stage): D1 = y -> x in [-2,TX+2), y in [0,TY); D1 = x -> x in [0,TX), y in [-2,TY+2)
    constexpr int AX0 = D1 == 2 ? -2 : 0, AW = D1 == 2 ? HX : TX;
    constexpr int AY0 = D1 == 2 ? 0 : -2, AH = D1 == 2 ? TY : HY;
    constexpr int NA = AW * AH, NPA = (NA + NTHR - 1) / NTHR;
    constexpr int NB = TX * TY, NPB = NB / NTHR;
    __shared__ R s_in[HY][HX];
    __shared__ R s_a[HY][HX];

    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, k = blockIdx.z;
    const size_t lay = (size_t)k * A.n2;

    CellC<R> ca[NPA];
    CellC<R> cb[NPB];
#pragma unroll
    for (int m = 0; m < NPA; m++) {
        const int c = tid + m * NTHR;
        ca[m].upd = 0;
        if (c < NA) {
            const int gi = x0 + AX0 + c % AW, gj = y0 + AY0 + c / AW;
            if (gi >= 0 && gi < A.nx_ && gj >= 0 && gj < A.ny_) {
                double hn1;
                cell_coef<D1, R>(A, gi, gj, k, dt, hin[lay + (size_t)gj * A.nx_ + gi], ca[m], hn1);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < NPB; m++) {
        const int c = tid + m * NTHR;
        const int gi = x0 + c % TX, gj = y0 + c / TX;
        cb[m].upd = 0;
        if (gi < A.nx_ && gj < A.ny_) {
            const size_t p3 = lay + (size_t)gj * A.nx_ + gi;
            CellC<R> tmp;
            double h1, h2;
            cell_coef<D1, R>(A, gi, gj, k, dt, hin[p3], tmp, h1);  // recomputed: identical operands
            cell_coef<D2, R>(A, gi, gj, k, dt, h1, cb[m], h2);
            hout[p3] = h2;
        }
    }

    for (int t = 0; t < ntr; t++) {
        const R* fi = fin + (size_t)t * A.n3 + lay;
        for (int c = tid; c < HX * HY; c += NTHR) {
            const int xl = c % HX, yl = c / HX;
            const int gi = x0 - 2 + xl, gj = y0 - 2 + yl;
            R val = R(0);
            if (gi >= 0 && gi < A.nx_ && gj >= 0 && gj < A.ny_) val = fi[(size_t)gj * A.nx_ + gi];
            s_in[yl][xl] = val;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NPA; m++) {
            const int c = tid + m * NTHR;
            if (c < NA) {
                const int xs = AX0 + c % AW + 2, ys = AY0 + c / AW + 2;  // s_in indices
                R out = s_in[ys][xs];
                if (ca[m].upd) {
                    R v[5];
#pragma unroll
                    for (int o = 0; o < 5; o++) v[o] = D1 == 1 ? s_in[ys][xs - 2 + o] : s_in[ys - 2 + o][xs];
                    R fl0 = face_flux<S, R>(ca[m].q0, ca[m].c0, ca[m].fl0, v[0], v[1], v[2], v[3]);
                    R fl1 = face_flux<S, R>(ca[m].q1, ca[m].c1, ca[m].fl1, v[1], v[2], v[3], v[4]);
                    if (A.ah_u) {
                        const int gi = x0 + xs - 2, gj = y0 + ys - 2;
                        if (D1 == 1) {
                            fl0 = fl0 - ah_flux<1, R>(A, gi - 1, gj, k, v[1], v[2]);
                            fl1 = fl1 - ah_flux<1, R>(A, gi, gj, k, v[2], v[3]);
                        } else {
                            fl0 = fl0 - ah_flux<2, R>(A, gi, gj - 1, k, v[1], v[2]);
                            fl1 = fl1 - ah_flux<2, R>(A, gi, gj, k, v[2], v[3]);
                        }
                    }
                    out = cell_update<R>(ca[m], dt, v[2], fl0, fl1);
                }
                s_a[ys][xs] = out;
            }
        }
        __syncthreads();
        R* fo = fout + (size_t)t * A.n3 + lay;
#pragma unroll
        for (int m = 0; m < NPB; m++) {
            const int c = tid + m * NTHR;
            const int xs = c % TX + 2, ys = c / TX + 2;
            const int gi = x0 + xs - 2, gj = y0 + ys - 2;
            R out = s_a[ys][xs];
            if (cb[m].upd) {
                R v[5];
#pragma unroll
                for (int o = 0; o < 5; o++) v[o] = D2 == 1 ? s_a[ys][xs - 2 + o] : s_a[ys - 2 + o][xs];
                R fl0 = face_flux<S, R>(cb[m].q0, cb[m].c0, cb[m].fl0, v[0], v[1], v[2], v[3]);
                R fl1 = face_flux<S, R>(cb[m].q1, cb[m].c1, cb[m].fl1, v[1], v[2], v[3], v[4]);
                if (A.ah_u) {
                    if (D2 == 1) {
                        fl0 = fl0 - ah_flux<1, R>(A, gi - 1, gj, k, v[1], v[2]);
                        fl1 = fl1 - ah_flux<1, R>(A, gi, gj, k, v[2], v[3]);
                    } else {
                        fl0 = fl0 - ah_flux<2, R>(A, gi, gj - 1, k, v[1], v[2]);
                        fl1 = fl1 - ah_flux<2, R>(A, gi, gj, k, v[2], v[3]);
                    }
                }
                out = cell_update<R>(cb[m], dt, v[2], fl0, fl1);
            }
            if (gi < A.nx_ && gj < A.ny_) fo[(size_t)gj * A.nx_ + gi] = out;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Vertical stage, one lane per column, TB tracers per sweep.
// advection.F90.template:223-374 with the sub-step count decided once from the
// thickness alone: the count and every restart depend only on h and w, so the
// h-only replay of the reference's control flow ends with the same count, and
// the tracer pass then repeats the reference's final (restart-free) pass.
// ---------------------------------------------------------------------------
template <int S, typename R, int TB>
__global__ void __launch_bounds__(256) k_wcol(const StepArgs A, double dt, int ntr, const R* __restrict__ fin,
                                              R* __restrict__ fout, const double* __restrict__ hin,
                                              double* __restrict__ hout, const double* __restrict__ w,
                                              const double* __restrict__ wv, int* __restrict__ stats)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= A.nx_ || j >= A.ny_) return;
    const size_t n2 = A.n2, n3 = A.n3;
    const int nz = A.nz;
    const size_t p = (size_t)j * A.nx_ + i;
    const bool inside = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax;
    if (!inside || A.tmask[p] != 1) {
        for (int k = 0; k < nz; k++) hout[(size_t)k * n2 + p] = hin[(size_t)k * n2 + p];
        for (int t = 0; t < ntr; t++)
            for (int k = 0; k < nz; k++) fout[(size_t)t * n3 + (size_t)k * n2 + p] = fin[(size_t)t * n3 + (size_t)k * n2 + p];
        return;
    }
    const double* hi = hin + p;
    double* ho = hout + p;
    const double* wp = w + p;
    const double* wvp = wv + p;

    // ---- h-only replay of the iteration control
    int iters = 1, capped = 0;
    double itersm1 = 1.0, dtk = dt;
    {
        int it = 1;
        while (it <= iters) {
            const double* sh = it == 1 ? hi : ho;
            if (iters < kItersMax) {
                double c = -1.7976931348623157e308;
                double hk = sh[0];
                for (int k = 1; k <= nz - 1; k++) {
                    const double hk1 = sh[(size_t)k * n2];
                    c = fmax2(c, (fabs(wvp[(size_t)k * n2]) * dt) / fmin2(hk, hk1));
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
            if (iters == 1) break;  // single sub-step: no intermediate thickness needed
            if (it < iters) {
                double wprev = wp[0];
                for (int k = 1; k <= nz; k++) {
                    const double wk = wp[(size_t)k * n2];
                    ho[(size_t)(k - 1) * n2] = sh[(size_t)(k - 1) * n2] - dtk * (wk - wprev);
                    wprev = wk;
                }
            }
            it++;
        }
    }
    if (iters > 1) atomicMax(&stats[0], iters);
    if (capped) atomicAdd(&stats[1], 1);

    // ---- tracer passes, TB tracers at a time, sharing the h sweep
    for (int t0 = 0; t0 < ntr; t0 += TB) {
        const int nb = ntr - t0 < TB ? ntr - t0 : TB;
        for (int it = 1; it <= iters; it++) {
            const double* sh = it == 1 ? hi : ho;
            const R* sf = (it == 1 ? fin : fout) + (size_t)t0 * n3 + p;
            R* df = fout + (size_t)t0 * n3 + p;
            R fm1[TB], f0[TB], f1[TB], fprev[TB];
            double h0 = sh[0], h1 = nz > 1 ? sh[n2] : 0.0;
#pragma unroll
            for (int b = 0; b < TB; b++) {
                fm1[b] = R(0);
                fprev[b] = R(0);
                f0[b] = b < nb ? sf[(size_t)b * n3] : R(0);
                f1[b] = (b < nb && nz > 1) ? sf[(size_t)b * n3 + n2] : R(0);
            }
            double wprev = wp[0];
            for (int k = 1; k <= nz; k++) {
                double h2 = 0.0;
                R f2[TB];
                const bool has2 = k + 2 <= nz;
                if (has2) h2 = sh[(size_t)(k + 1) * n2];
#pragma unroll
                for (int b = 0; b < TB; b++) f2[b] = (has2 && b < nb) ? sf[(size_t)b * n3 + (size_t)(k + 1) * n2] : R(0);
                const double wk = wp[(size_t)k * n2];
                const double hnew = h0 - dtk * (wk - wprev);
                R fl[TB];
                if (k <= nz - 1) {
                    const double wvk = wvp[(size_t)k * n2];
                    const double cfl = ((fabs(wvk) * dt) / fmin2(h0, h1)) * itersm1;
                    const R rcfl = R(cfl), rw = R(wvk);
                    const bool pos = wvk > 0.0;
#pragma unroll
                    for (int b = 0; b < TB; b++) {
                        R fu, fuu, fd;
                        if (pos) {
                            fu = f0[b];
                            fuu = k > 1 ? fm1[b] : fu;
                            fd = f1[b];
                        } else {
                            fu = f1[b];
                            fuu = k < nz - 1 ? f2[b] : fu;
                            fd = f0[b];
                        }
                        fl[b] = rw * face_value<S, R>(rcfl, fu, fuu, fd);
                    }
                } else {
#pragma unroll
                    for (int b = 0; b < TB; b++) fl[b] = R(0);
                }
                // sub-stepped columns replay h in place per batch, so every batch writes it
                if (t0 == 0 || iters > 1) ho[(size_t)(k - 1) * n2] = hnew;
                if constexpr (std::is_same<R, double>::value) {
#pragma unroll
                    for (int b = 0; b < TB; b++) {
                        const double hfo = h0 * f0[b];
                        const double advn = fl[b] - fprev[b];
                        if (b < nb) df[(size_t)b * n3 + (size_t)(k - 1) * n2] = (hfo - dtk * advn) / hnew;
                    }
                } else {
                    const R ca = R(h0 / hnew), cb = R(dtk / hnew);
#pragma unroll
                    for (int b = 0; b < TB; b++)
                        if (b < nb) df[(size_t)b * n3 + (size_t)(k - 1) * n2] = ca * f0[b] - cb * (fl[b] - fprev[b]);
                }
#pragma unroll
                for (int b = 0; b < TB; b++) {
                    fm1[b] = f0[b];
                    f0[b] = f1[b];
                    f1[b] = f2[b];
                    fprev[b] = fl[b];
                }
                h0 = h1;
                h1 = h2;
                wprev = wk;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP32 vertical diffusion of all tracers, one lane per column, the column held
// in registers.  Tracer-independent parts of the theta scheme
// (diffusion.F90:163-198, 238-292) are formed once per column in FP64 and kept
// in LDS as FP32 factors:
//   qu_k = g1_k var_k + g3_k (var_{k+1} - var_k) + g4_k (var_{k-1} - var_k) + g2_k qu_{k+1}
//   var_k = qu_k - ru_k var_{k-1}
// with den_k = a2_k - a3_k ru_{k+1}, g1 = dzo/den, g2 = auxn_k/den,
// g3 = auxo_k/den, g4 = auxo_{k-1}/den (g3 = g4 = 0 when cnpar = 1).
// ---------------------------------------------------------------------------
template <int NZMAX, bool CN1>
__global__ void __launch_bounds__(64) k_vdiff_f32(const StepArgs A, double dt, double cnpar, double molecular,
                                                  const double* __restrict__ nuh, int ntr, const float* __restrict__ fin,
                                                  float* __restrict__ fout)
{
    extern __shared__ float s_coef[];  // [ncoef][nz][64]
    const int lane = threadIdx.x;
    const int i = blockIdx.x * 64 + lane;
    const int j = blockIdx.y;
    const int nz = A.nz;
    const size_t n2 = A.n2, n3 = A.n3;
    constexpr int NC = CN1 ? 3 : 5;
    float* g1 = s_coef;
    float* g2 = s_coef + (size_t)nz * 64;
    float* ru = s_coef + (size_t)2 * nz * 64;
    float* g3 = s_coef + (size_t)3 * nz * 64;
    float* g4 = s_coef + (size_t)4 * nz * 64;
    (void)NC;
    if (i >= A.nx_) return;
    const size_t p = (size_t)j * A.nx_ + i;
    const bool inside = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax;
    const bool solve = inside && A.tmask[p] == 1 && nz > 1;
    if (solve) {
        // factorisation from the surface down, FP64
        const double* hn = A.hn + p;
        double ru_up = 0.0;
        for (int k = nz - 1; k >= 0; k--) {
            const double dz = hn[(size_t)k * n2];
            double an_k = 0.0, ao_k = 0.0, an_km1 = 0.0, ao_km1 = 0.0;
            if (k <= nz - 2) {
                const double x = (2.0 * dt) * (nuh[(size_t)(k + 1) * n2 + p] + molecular);
                const double s = hn[(size_t)(k + 1) * n2] + dz;
                an_k = (cnpar * x) / s;
                ao_k = ((1.0 - cnpar) * x) / s;
            }
            if (k >= 1) {
                const double x = (2.0 * dt) * (nuh[(size_t)k * n2 + p] + molecular);
                const double s = dz + hn[(size_t)(k - 1) * n2];
                an_km1 = (cnpar * x) / s;
                ao_km1 = ((1.0 - cnpar) * x) / s;
            }
            const double a2 = (dz + an_k) + an_km1;
            const double den = k == nz - 1 ? a2 : a2 + an_k * ru_up;  // a2 - a3 ru_{k+1}, a3 = -auxn_k
            const double ruk = -an_km1 / den;
            g1[k * 64 + lane] = float(dz / den);
            g2[k * 64 + lane] = float(an_k / den);
            ru[k * 64 + lane] = float(ruk);
            if (!CN1) {
                g3[k * 64 + lane] = float(ao_k / den);
                g4[k * 64 + lane] = float(ao_km1 / den);
            }
            ru_up = ruk;
        }
    }
    for (int t = 0; t < ntr; t++) {
        const float* fi = fin + (size_t)t * n3 + p;
        float* fo = fout + (size_t)t * n3 + p;
        float vr[NZMAX];
#pragma unroll
        for (int k = 0; k < NZMAX; k++)
            if (k < nz) vr[k] = fi[(size_t)k * n2];
        if (nz == 1) {
            // diffusion.F90:224-231: every point, halos and land included
            const double dz = A.hn[p];
            vr[0] = float((dz * (double)vr[0]) / dz);
        } else if (solve) {
            // forward (top-down) elimination, qu overwrites vr with a window of old values
            float v_up = 0.f, qu_up = 0.f;
#pragma unroll
            for (int k = NZMAX - 1; k >= 0; k--) {
                if (k < nz) {
                    const float vk = vr[k];
                    float q = g1[k * 64 + lane] * vk;
                    if (k < nz - 1) q += g2[k * 64 + lane] * qu_up;
                    if (!CN1) {
                        if (k < nz - 1) q += g3[k * 64 + lane] * (v_up - vk);
                        if (k > 0) q += g4[k * 64 + lane] * (vr[k - 1 >= 0 ? k - 1 : 0] - vk);
                    }
                    vr[k] = q;
                    qu_up = q;
                    v_up = vk;
                }
            }
            float prev = vr[0];
#pragma unroll
            for (int k = 1; k < NZMAX; k++) {
                if (k < nz) {
                    prev = vr[k] - ru[k * 64 + lane] * prev;
                    vr[k] = prev;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NZMAX; k++)
            if (k < nz) fo[(size_t)k * n2] = vr[k];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename R>
int launch_hpair(int d1, int scheme, const StepArgs& A, double dt, int ntr, const R* fin, R* fout,
                 const double* hin, double* hout, hipStream_t s)
{
    dim3 grid((A.nx_ + TX - 1) / TX, (A.ny_ + TY - 1) / TY, A.nz);
    ProfileScope ps(d1 == 2 ? GETM_K_HPAIR_VU : GETM_K_HPAIR_UV, s);
    GETM_SCHEME_SWITCH(scheme, {
        if (d1 == 2)
            hipLaunchKernelGGL((k_hpair<S_, R, 2>), grid, dim3(NTHR), 0, s, A, dt, ntr, fin, fout, hin, hout);
        else
            hipLaunchKernelGGL((k_hpair<S_, R, 1>), grid, dim3(NTHR), 0, s, A, dt, ntr, fin, fout, hin, hout);
    });
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

template <typename R>
int launch_wcol(int scheme, const StepArgs& A, double dt, int ntr, const R* fin, R* fout, const double* hin,
                double* hout, const double* w, const double* wv, int* stats, hipStream_t s)
{
    dim3 block(64, 4, 1);
    dim3 grid((A.nx_ + 63) / 64, (A.ny_ + 3) / 4, 1);
    ProfileScope ps(GETM_K_WCOL, s);
    constexpr int TB = std::is_same<R, double>::value ? 2 : 4;
    GETM_SCHEME_SWITCH(scheme, {
        hipLaunchKernelGGL((k_wcol<S_, R, TB>), grid, block, 0, s, A, dt, ntr, fin, fout, hin, hout, w, wv, stats);
    });
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

int launch_vdiff_f32(const StepArgs& A, double dt, double cnpar, double molecular, const double* nuh, int ntr,
                     const float* fin, float* fout, hipStream_t s)
{
    if (A.nz > 64) return GETM_ERR_CAPACITY;
    const bool cn1 = cnpar == 1.0;
    const size_t shm = (size_t)(cn1 ? 3 : 5) * A.nz * 64 * sizeof(float);
    dim3 grid((A.nx_ + 63) / 64, A.ny_, 1);
    ProfileScope ps(GETM_K_VDIFF_F32, s);
#define GETM_VD(NZM)                                                                                          \
    do {                                                                                                      \
        if (cn1)                                                                                              \
            hipLaunchKernelGGL((k_vdiff_f32<NZM, true>), grid, dim3(64), shm, s, A, dt, cnpar, molecular, nuh, ntr, \
                               fin, fout);                                                                    \
        else                                                                                                  \
            hipLaunchKernelGGL((k_vdiff_f32<NZM, false>), grid, dim3(64), shm, s, A, dt, cnpar, molecular, nuh, ntr, \
                               fin, fout);                                                                    \
    } while (0)
    if (A.nz <= 16) GETM_VD(16);
    else if (A.nz <= 32) GETM_VD(32);
    else if (A.nz <= 48) GETM_VD(48);
    else GETM_VD(64);
#undef GETM_VD
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

template int launch_hpair<float>(int, int, const StepArgs&, double, int, const float*, float*, const double*, double*,
                                 hipStream_t);
template int launch_hpair<double>(int, int, const StepArgs&, double, int, const double*, double*, const double*,
                                  double*, hipStream_t);
template int launch_wcol<float>(int, const StepArgs&, double, int, const float*, float*, const double*, double*,
                                const double*, const double*, int*, hipStream_t);
template int launch_wcol<double>(int, const StepArgs&, double, int, const double*, double*, const double*, double*,
                                 const double*, const double*, int*, hipStream_t);

}  // namespace getm
