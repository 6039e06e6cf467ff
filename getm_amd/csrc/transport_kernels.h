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
#pragma once

#include <type_traits>

#include "common.h"
#include "internal.h"

namespace getm {

// ---------------------------------------------------------------------------
// Horizontal stage pair on one layer tile (v dt/2 then u dt/2, or u then v).
//
// Per block: a 64 x TY output tile of a group of KG (1 or HPAIR_KG) consecutive
// layers and its input tile with halo 2.  The layer-independent terms (the
// faces' mask flags and metrics, the cells' wet flags and inverse areas) are
// loaded once per block into LDS; each layer then loads only its velocities,
// face and cell thicknesses and tracers.
// Per-layer set-up (tracer-independent): every global load of the set-up --
// face fields of both stages, cell thicknesses and the first tracer's
// tile -- is issued before any is used, so the block waits for HBM once.  Face
// transports are formed once and shared through LDS for the FP64 thickness
// update; each face keeps its transport, the cfl-only limiter terms and the
// LDS offsets of its upstream, upstream-upstream and downstream cells, so the
// tracer loop only loads, limits and updates.  Per tracer, each face is
// limited once into an LDS flux row and the next tracer's tile is in flight
// (register prefetch) while the current one is computed.
// Face (x, y) of direction D sits between tile cells (x, y) and (x, y) + e_D.
// Faces outside the reference's face loops (template:61-62, 161-162) are never
// read by an updated cell and are left closed.
// ---------------------------------------------------------------------------
constexpr int TX = 64;      // output tile width (one wave); the height TY is a template parameter

// Row offset j * nx_ (and other products of two array indices): both are below
// 2^24 (create_grid), so the full-rate 24-bit multiply gives the exact product.
__device__ __forceinline__ int mul24(int a, int b) { return (int)__umul24((unsigned)a, (unsigned)b); }

constexpr int HX = TX + 4;  // input tile width incl. halo 2

struct FaceC {
    double q;     // transport Q = vel*h*metric (FP64, also drives the h update)
    double cfl;   // |vel*dt*inverse metric|
    int flags;    // bit0 open, bit1 vel > 0, bit2 upstream-upstream face open
};


// Face of the reference's face loops?  x faces: i in [imin-1, imax], j in [jmin, jmax];
// y faces: i in [imin, imax], j in [jmin-1, jmax].  There the neighbour masks at
// p +- s exist (halo >= 2).
template <int DIR>
__device__ __forceinline__ bool face_in_loop(const StepArgs& A, int gi, int gj)
{
    if (DIR == 1) return gi >= A.imin - 1 && gi <= A.imax && gj >= A.jmin && gj <= A.jmax;
    return gi >= A.imin && gi <= A.imax && gj >= A.jmin - 1 && gj <= A.jmax;
}

// Face transport and CFL number from the face's layer fields (velocity, face
// thickness), its 2-D metrics and its mask flags fl (bit0 in the face loop and
// open, bit1 / bit2 the lower / upper neighbour face along the sweep open):
// QU = u*hu*dyu, cfl = |u*dt*idxu| (template:64-65; QV/cfl likewise :164-165).
__device__ __forceinline__ FaceC face_make2(unsigned fl, double vel, double hf, double mq, double mc, double dt)
{
    FaceC fc{0.0, 0.0, 0};
    if (!(fl & 1u)) return fc;
    fc.q = (vel * hf) * mq;
    fc.cfl = fabs((vel * dt) * mc);
    const bool pos = vel > 0.0;
    const bool fuu_ok = pos ? (fl & 2u) : (fl & 4u);
    fc.flags = 1 | (pos ? 2 : 0) | (fuu_ok ? 4 : 0);
    return fc;
}

// Per-face terms that do not depend on the tracer, plus the LDS offsets of the
// upstream (o_u), upstream-upstream (o_uu) and downstream (o_d) cells, chosen by
// the velocity sign and the neighbour mask (template:66-80).
//   FP64: Q, cfl -> the reference's limiter and face value, bit for bit.
//   FP32: Q, 0.5 (1 - cfl) and the cfl-only parts of the limiter (HSIMT:
//   0.5 +- x; P2_PDM: 0.5 +- x, 2/(cfl + 1e-10), 2/(1 - cfl)), evaluated in FP64 once.
template <int S, typename R>
struct FaceK;

struct FaceOff {
    int open, o_u, o_uu, o_d;
    // base = LDS offset of the lower cell, sd = stride along the sweep.
    // Branch-free so the per-face terms stay in registers.
    __device__ __forceinline__ void set(const FaceC& f, int base, int sd)
    {
        const bool op = f.flags & 1, pos = f.flags & 2, uu = f.flags & 4;
        const int ou = pos ? base : base + sd;
        const int od = pos ? base + sd : base;
        const int ouu = uu ? (pos ? base - sd : base + 2 * sd) : ou;
        open = op;
        o_u = op ? ou : base;
        o_d = op ? od : base;
        o_uu = op ? ouu : base;
    }
};

template <int S>
struct FaceK<S, double> : FaceOff {
    double q, cfl;
    __device__ void set(const FaceC& f, int base, int sd)
    {
        FaceOff::set(f, base, sd);
        q = f.q;
        cfl = f.cfl;
    }
    __device__ __forceinline__ double flux(const double* s) const
    {
        const double fl = q * face_value<S, double>(cfl, s[o_u], s[o_uu], s[o_d]);
        return open ? fl : 0.0;
    }
};

// FP32 limited face value with the cfl-only terms formed once in FP64:
// 0.5 (1 - cfl), and HSIMT: 0.5 +- x; P2_PDM: 0.5 +- x, 2/(cfl + 1e-10), 2/(1 - cfl)
// (limiters.h:1-44).  Branch-free: the limited value is formed always and selected.
template <int S>
struct LimF {
    static constexpr int NL = S == HSIMT ? 2 : (S == P2_PDM ? 4 : 0);
    float ch;
    float l[NL > 0 ? NL : 1];
    __device__ __forceinline__ void set(double cfl)
    {
        ch = float(0.5 * (1.0 - cfl));
        if constexpr (S == HSIMT) {
            const double kappa = 1.0 - cfl;
            const double x = 0.25 * (kappa - (1.0 / 3.0) / kappa);
            l[0] = float(0.5 + x);
            l[1] = float(0.5 - x);
        } else if constexpr (S == P2_PDM) {
            const double x = (1.0 / 6.0) * (1.0 - 2.0 * cfl);
            l[0] = float(0.5 + x);
            l[1] = float(0.5 - x);
            l[2] = float(2.0 / (cfl + 1.e-10));
            l[3] = float(2.0 / (1.0 - cfl));
        }
    }
    // limiter of the ratio r (limiters.h:1-44) with the cfl-only terms of set()
    __device__ __forceinline__ float limit(float r) const
    {
        if constexpr (S == SUPERBEE) return fmax2(fmin2(2.f * r, 1.f), fmin2(r, 2.f));
        else if constexpr (S == MUSCL) return fmin3(2.f * r, 0.5f * (1.f + r), 2.f);
        else if constexpr (S == SPLMAX13)
            return fmin3(2.f * r, (1.f / 3.f) * fmax2(__builtin_fmaf(2.f, r, 1.f), 2.f + r), 2.f);
        else if constexpr (S == HSIMT) return fmin3(2.f * r, fmax2(0.f, __builtin_fmaf(l[1], r, l[0])), 2.f);
        else if constexpr (S == P2_PDM) return fmin3(r * l[2], __builtin_fmaf(l[1], r, l[0]), l[3]);
        else return 0.f;
    }
    __device__ __forceinline__ float value(float fu, float fuu, float fd) const
    {
        if constexpr (S == UPSTREAM) {
            return fu;
        } else {
            const float deltaf = fd - fu;
            const float deltafu = fu - fuu;
            // hardware reciprocal (1 ulp): deltaf == 0 gives inf/NaN here, but
            // then deltaf * deltafu > 0 fails and fu is selected below
            const float r = deltafu * __builtin_amdgcn_rcpf(deltaf);
            const float lim = limit(r);
            // FP32 only: the contractions are written out (the build keeps FP64 uncontracted)
            const float lim_fu = __builtin_fmaf(lim * ch, deltaf, fu);
            return deltaf * deltafu > 0.f ? lim_fu : fu;
        }
    }
    // Limited correction of the upstream value, (0.5 lim (1 - cfl)) deltaf where
    // deltaf deltafu > 0 and 0 elsewhere (template:82-88), from the downstream
    // and upstream differences deltaf = fd - fu, deltafu = fu - fuu.
    __device__ __forceinline__ float corr(float deltaf, float deltafu) const
    {
        if constexpr (S == UPSTREAM) {
            return 0.f;
        } else {
            const float r = deltafu * __builtin_amdgcn_rcpf(deltaf);
            const float c = (limit(r) * ch) * deltaf;
            return deltaf * deltafu > 0.f ? c : 0.f;
        }
    }
};

template <int S>
struct FaceK<S, float> : FaceOff {
    float q;
    LimF<S> lim;
    __device__ void set(const FaceC& f, int base, int sd)
    {
        FaceOff::set(f, base, sd);
        q = float(f.q);
        lim.set(f.cfl);
    }
    __device__ __forceinline__ float flux(const float* s) const
    {
        const float val = lim.value(s[o_u], s[o_uu], s[o_d]);
        return open ? q * val : 0.f;
    }
};

// Compensated FP32 (GETM_F32C): tile values and transports in FP64, the limiter
// in FP32 on the FP64 differences (LimF: the cfl-only terms formed once), and
// the limited face value fu + corr and the flux Q (fu + corr) in FP64.  The
// FP32 limiter only shapes the flux; every face's FP64 flux enters both of its
// cells, so the update conserves to FP64 rounding.
template <int S>
struct FaceKC : FaceOff {
    double q;
    LimF<S> lim;
    __device__ void set(const FaceC& f, int base, int sd)
    {
        FaceOff::set(f, base, sd);
        q = f.q;
        lim.set(f.cfl);
    }
    __device__ __forceinline__ double flux(const double* s) const
    {
        const double fu = s[o_u];
        const double val = fu + (double)lim.corr(float(s[o_d] - fu), float(fu - s[o_uu]));
        return open ? q * val : 0.0;
    }
};

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


// Per-cell terms of one stage; the cell is updated iff interior and wet.
//   FP64: hold, hnew, iA for the reference order (h f - dt (dF iA)) / hnew
//   FP32: a = (hold - hnew)/hnew, b = dt iA/hnew:  f_new = f + (a f - b dF).
// The FP32 increment form matters: the factors are rounded once per step and,
// in a steady flow, identically every step, so in the direct form
// f_new = (hold/hnew) f - b dF the rounding of hold/hnew (up to 2^-25 of f)
// would accumulate linearly with the step count; in the increment form it is
// scaled by |hold/hnew - 1| (a Courant number) and the only O(ulp(f)) error is
// the final, data-dependent rounding.
// own = LDS offset of the cell, fl1 = offset of its upper face flux, rb = the
// cell's random bits of this stage (stochastic rounding; sr_cell_bits of its
// index in the domain's 3-D array).
template <typename R>
struct CellK {
    R a, b, iA;
    int upd, own, fl1;
    unsigned rb;
};

// C (compensated FP32, R = double): the FP32 form's increment factors in FP64,
// f_new = f + (a f - b dF) evaluated in FP64.
template <typename R, bool C = false>
__device__ __forceinline__ void cell_set(CellK<R>& c, bool upd, double hold, double hnew, double iA, double dt)
{
    c.upd = upd;
    // cells that are not updated still run the (discarded) update: defined factors
    c.a = R(0);
    c.b = R(1);
    c.iA = R(0);
    if (!upd) return;
    if constexpr (C) {
        const double rh = rcp_nr(hnew);
        c.a = (hold - hnew) * rh;
        c.b = (dt * iA) * rh;
        c.iA = 0.0;
    } else if constexpr (std::is_same<R, double>::value) {
        c.a = hold;
        c.b = hnew;
        c.iA = iA;
    } else {
        const double rh = rcp_nr(hnew);
        c.a = float((hold - hnew) * rh);
        c.b = float((dt * iA) * rh);
        c.iA = 0.f;
    }
}

template <typename R, bool C = false>
__device__ __forceinline__ R cell_update(const CellK<R>& c, double dt, R f, R fl0, R fl1, const SRound& sr)
{
    if constexpr (C) {
        return f + fma(c.a, f, -(c.b * (fl1 - fl0)));
    } else if constexpr (std::is_same<R, double>::value) {
        const double hfo = c.a * f;
        const double advn = (fl1 - fl0) * c.iA;
        return (hfo - dt * advn) / c.b;
    } else {
        return sr.add(f, __builtin_fmaf(c.a, f, -(c.b * (fl1 - fl0))), c.rb);
    }
}

// Stage seed of a rounding site in this step (0: round to nearest; FP64 never rounds here).
template <typename R>
__device__ __forceinline__ unsigned stage_seed(const StepArgs& A, unsigned stage)
{
    if constexpr (std::is_same<R, double>::value) return 0u;
    else return A.rstep ? sr_stage_seed(*A.rstep, stage) : 0u;
}
__device__ __forceinline__ SRound tracer_round(unsigned stage_seed, int tracer)
{
    return SRound{stage_seed ? sr_tracer_key(tracer) : 0u};
}

// Tile-local coordinates: output [0,TX) x [0,TY); input [-2,TX+2) x [-2,TY+2)
// stored at LDS offset (y+2)*HX + (x+2).  Region A (first stage) cells:
// D1 = y -> x in [-2,TX+2), y in [0,TY); D1 = x -> x in [0,TX), y in [-2,TY+2).
// sr_stage: stage id of the first stage's rounding (the second takes sr_stage + 1);
// tg0: index in the domain of the first tracer of fin/fout.
// GROUP: the block walks several layers and keeps the layer-independent terms
// in LDS; otherwise (one layer) they are loaded straight into registers.
// HDIFF: horizontal diffusion fluxes (Ah given).  A separate instantiation: the
// branch alone, even never taken, holds ~45 VGPRs of its operands.
// C: compensated FP32 storage (R = double): fin/fout hold the FP32 values and
// lin/lout their int16 residuals; tiles are decoded to FP64 in LDS, faces are
// FaceKC, cells the FP64 increment form, and updated cells are encoded again.
template <int S, typename R, int D1, int TY, bool GROUP, bool HDIFF, int TXW = TX, bool C = false, int LPC = 1>
__device__ __forceinline__ void hpair_body(const StepArgs& A, double dt, int x0, int y0, int k0, int k1, int ntr,
                                           const std::conditional_t<C, float, R>* __restrict__ fin,
                                           std::conditional_t<C, float, R>* __restrict__ fout,
                                           const double* __restrict__ hin, double* __restrict__ hout,
                                           unsigned sr_stage, int tg0, const double* hin_dry, bool write_dry,
                                           const short* __restrict__ lin = nullptr, short* __restrict__ lout = nullptr)
{
    static_assert(!C || std::is_same<R, double>::value, "compensated storage works on FP64 values");
    using T = std::conditional_t<C, float, R>;  // stored tracer type
    using FK = std::conditional_t<C, FaceKC<S>, FaceK<S, R>>;
    // tile width of this instantiation (the file-wide 64 unless given)
    constexpr int TX = TXW, HX = TXW + 4;
    constexpr int HY = TY + 4;       // input tile height incl. halo 2
    constexpr int NT = HX * HY;      // cells of the input tile
    constexpr int NTHR = TX * TY / LPC;  // LPC output cells per lane
    static_assert(NTHR * LPC == TX * TY, "cells per lane must divide the tile");
    constexpr int D2 = 3 - D1;
    constexpr int SD1 = D1 == 1 ? 1 : HX, SD2 = D2 == 1 ? 1 : HX;  // LDS strides along the sweeps
    // region A cells
    constexpr int AX0 = D1 == 2 ? -2 : 0, AW = D1 == 2 ? HX : TX;
    constexpr int AY0 = D1 == 2 ? 0 : -2, AH = D1 == 2 ? TY : HY;
    constexpr int NA = AW * AH, NPA = (NA + NTHR - 1) / NTHR;
    // region A faces: D1 = y -> x in [-2,TX+2), y in [-1,TY); D1 = x -> x in [-1,TX), y in [-2,TY+2)
    constexpr int FAX0 = D1 == 2 ? -2 : -1, FAW = D1 == 2 ? HX : TX + 1;
    constexpr int FAY0 = D1 == 2 ? -1 : -2, FAH = D1 == 2 ? TY + 1 : HY;
    constexpr int NFA = FAW * FAH, NPFA = (NFA + NTHR - 1) / NTHR;
    // region B faces: D2 = x -> x in [-1,TX), y in [0,TY); D2 = y -> x in [0,TX), y in [-1,TY)
    constexpr int FBX0 = D2 == 1 ? -1 : 0, FBW = D2 == 1 ? TX + 1 : TX;
    constexpr int FBY0 = D2 == 1 ? 0 : -1, FBH = D2 == 1 ? TY : TY + 1;
    constexpr int NFB = FBW * FBH, NPFB = (NFB + NTHR - 1) / NTHR;
    constexpr int NB = TX * TY, NPB = NB / NTHR;
    constexpr int NPIN = (NT + NTHR - 1) / NTHR;

    // LDS: the first tile buffer, and one pool that holds the other two tile
    // buffers, the stage-A tile and the flux rows while tracers are processed,
    // and the FP64 face transports (s_q) and stage-A thickness (s_h) during each
    // layer's set-up (a smaller footprint leaves room for a third block per CU)
    constexpr int NQ = NFA > NFB ? NFA : NFB;
    constexpr int POOL_T = (int)sizeof(R) * (3 * NT + NFA + NFB);
    constexpr int POOL_S = (int)sizeof(double) * (NQ + NT);
    static_assert(POOL_S <= POOL_T, "set-up arrays must fit in the tracer-loop pool");
    __shared__ R s_in0[NT];
    __shared__ __attribute__((aligned(16))) unsigned char s_pool[POOL_T];
    R* const s_in1 = reinterpret_cast<R*>(s_pool);
    R* const s_in2 = s_in1 + NT;
    R* const s_a = s_in2 + NT;
    R* const s_fa = s_a + NT;
    R* const s_fb = s_fa + NFA;
    double* const s_q = reinterpret_cast<double*>(s_pool);  // FP64 face transports of the stage being set up
    double* const s_h = s_q + NQ;                           // FP64 thickness after stage A
    auto s_in = [&](int b) -> R* { return b == 0 ? s_in0 : (b == 1 ? s_in1 : s_in2); };
    // layer-independent terms, formed once per block: face metrics (transport and
    // inverse CFL metric) and mask flags, cell inverse areas and wet flags
    constexpr int GFA = GROUP ? NFA : 1, GFB = GROUP ? NFB : 1, GT = GROUP ? NT : 1;
    __shared__ double s_mqa[GFA], s_mca[GFA], s_mqb[GFB], s_mcb[GFB];
    __shared__ double s_ia[GT];
    __shared__ unsigned char s_ffa[GFA], s_ffb[GFB], s_cf[GT];

    const int tid = threadIdx.x;
    auto inarr = [&](int gi, int gj) { return gi >= 0 && gi < A.nx_ && gj >= 0 && gj < A.ny_; };
    auto wet = [&](int gi, int gj, int tm) {
        return gi >= A.imin && gi <= A.imax && gj >= A.jmin && gj <= A.jmax && tm == 1;
    };

    // ---- layer-independent set-up (into LDS for a layer group).
    // faces: 2-D loads (masks of the face and of its two neighbours along the sweep,
    // transport and CFL metrics) -> flags bit0 in the face loop and open, bit1 the
    // lower neighbour face open, bit2 the upper one (the upstream-upstream faces)
    auto face2d = [&](auto dirc, int gi, int gj, unsigned char& fl, double& mq, double& mc) -> int {
        constexpr int DIR = decltype(dirc)::value;
        fl = 0;
        mq = 0.0;
        mc = 0.0;
        if (!face_in_loop<DIR>(A, gi, gj)) return -1;
        const unsigned p = (unsigned)(mul24(gj, A.nx_) + gi);
        const unsigned s = DIR == 1 ? 1u : (unsigned)A.nx_;  // p - s >= 0 inside the face loops (halo >= 2)
        const int* mask = DIR == 1 ? A.umask : A.vmask;
        // every load issued before any is used (no load waits for the mask)
        const int m0 = at32(mask, 4 * p), mlo = at32(mask, 4 * (p - s)), mhi = at32(mask, 4 * (p + s));
        mq = at32(DIR == 1 ? A.dyu : A.dxv, 8 * p);   // metric of the transport
        mc = at32(DIR == 1 ? A.idxu : A.idyv, 8 * p);  // inverse metric of the CFL number
        fl = open_face(m0) ? (1 | (open_face(mlo) ? 2 : 0) | (open_face(mhi) ? 4 : 0)) : 0;
        return (int)p;  // in the face loop (open or closed)
    };
    // cells: bit0 in the array, bit1 updated (interior and wet); returns the in-layer offset or -1
    auto cell2d = [&](int gi, int gj, unsigned char& cf, double& ia) -> int {
        cf = 0;
        ia = 0.0;
        if (!inarr(gi, gj)) return -1;
        const unsigned p = (unsigned)(mul24(gj, A.nx_) + gi);
        cf = 1 | (wet(gi, gj, at32(A.tmask, 4 * p)) ? 2 : 0);
        ia = at32(A.iarea, 8 * p);
        return (int)p;
    };
    if constexpr (GROUP) {
#pragma unroll
        for (int m = 0; m < NPFA; m++) {
            const int c = tid + m * NTHR;
            if (c < NFA) {
                unsigned char fl;
                double mq, mc;
                face2d(std::integral_constant<int, D1>(), x0 + FAX0 + c % FAW, y0 + FAY0 + c / FAW, fl, mq, mc);
                s_ffa[c] = fl;
                s_mqa[c] = mq;
                s_mca[c] = mc;
            }
        }
#pragma unroll
        for (int m = 0; m < NPFB; m++) {
            const int c = tid + m * NTHR;
            if (c < NFB) {
                unsigned char fl;
                double mq, mc;
                face2d(std::integral_constant<int, D2>(), x0 + FBX0 + c % FBW, y0 + FBY0 + c / FBW, fl, mq, mc);
                s_ffb[c] = fl;
                s_mqb[c] = mq;
                s_mcb[c] = mc;
            }
        }
        // region A cells (region B's are a subset)
#pragma unroll
        for (int m = 0; m < NPA; m++) {
            const int c = tid + m * NTHR;
            if (c < NA) {
                const int xl = AX0 + c % AW, yl = AY0 + c / AW;
                const int own = (yl + 2) * HX + (xl + 2);
                unsigned char cf;
                double ia;
                cell2d(x0 + xl, y0 + yl, cf, ia);
                s_cf[own] = cf;
                s_ia[own] = ia;
            }
        }
        __syncthreads();
    }

    const unsigned ss_a = stage_seed<R>(A, sr_stage), ss_b = stage_seed<R>(A, sr_stage + 1);
    for (int k = k0; k < k1; k++) {
        // the lane index through an empty asm: the layer-invariant index arithmetic
        // below is recomputed per layer instead of being hoisted out of the loop
        // and held in registers across it
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        // in-layer offsets of this lane's input-tile cells (-1 outside the array),
        // output cells (-1 outside), open faces and region A cells (-1: nothing to load)
        int toff[NPIN], ooff[NPB], foa[NPFA], fob[NPFB], coa[NPA];
#pragma unroll
        for (int m = 0; m < NPIN; m++) {
            const int c = tid + m * NTHR;
            toff[m] = -1;
            if (c < NT) {
                const int gi = x0 - 2 + c % HX, gj = y0 - 2 + c / HX;
                if (inarr(gi, gj)) toff[m] = mul24(gj, A.nx_) + gi;
            }
        }
#pragma unroll
        for (int m = 0; m < NPB; m++) {
            const int c = tid + m * NTHR;
            const int gi = x0 + c % TX, gj = y0 + c / TX;
            ooff[m] = (gi < A.nx_ && gj < A.ny_) ? mul24(gj, A.nx_) + gi : -1;
        }
        // one layer: the layer-independent terms of this lane's faces and cells in registers
        unsigned char rffa[NPFA], rffb[NPFB], rcfa[NPA], rcfb[NPB];
        double rmqa[NPFA], rmca[NPFA], rmqb[NPFB], rmcb[NPFB], ria_a[NPA], ria_b[NPB];
#pragma unroll
        for (int m = 0; m < NPFA; m++) {
            const int c = tid + m * NTHR;
            if constexpr (GROUP) {
                foa[m] = (c < NFA && (s_ffa[c] & 1)) ? mul24(y0 + FAY0 + c / FAW, A.nx_) + x0 + FAX0 + c % FAW : -1;
            } else {
                rffa[m] = 0;
                rmqa[m] = rmca[m] = 0.0;
                foa[m] = c < NFA ? face2d(std::integral_constant<int, D1>(), x0 + FAX0 + c % FAW, y0 + FAY0 + c / FAW,
                                          rffa[m], rmqa[m], rmca[m])
                                 : -1;
            }
        }
#pragma unroll
        for (int m = 0; m < NPFB; m++) {
            const int c = tid + m * NTHR;
            if constexpr (GROUP) {
                fob[m] = (c < NFB && (s_ffb[c] & 1)) ? mul24(y0 + FBY0 + c / FBW, A.nx_) + x0 + FBX0 + c % FBW : -1;
            } else {
                rffb[m] = 0;
                rmqb[m] = rmcb[m] = 0.0;
                fob[m] = c < NFB ? face2d(std::integral_constant<int, D2>(), x0 + FBX0 + c % FBW, y0 + FBY0 + c / FBW,
                                          rffb[m], rmqb[m], rmcb[m])
                                 : -1;
            }
        }
        // bit m: region A cell m is updated (its thickness comes from hin; the
        // others' from hin_dry, which holds the same values where nothing updates them)
        unsigned wet_a = 0;
#pragma unroll
        for (int m = 0; m < NPA; m++) {
            const int c = tid + m * NTHR;
            coa[m] = -1;
            rcfa[m] = 0;
            ria_a[m] = 0.0;
            if (c < NA) {
                const int xl = AX0 + c % AW, yl = AY0 + c / AW;
                if constexpr (GROUP) {
                    const unsigned char cf = s_cf[(yl + 2) * HX + (xl + 2)];
                    if (cf & 1) coa[m] = mul24(y0 + yl, A.nx_) + x0 + xl;
                    if (cf & 2) wet_a |= 1u << m;
                } else {
                    coa[m] = cell2d(x0 + xl, y0 + yl, rcfa[m], ria_a[m]);
                    if (rcfa[m] & 2) wet_a |= 1u << m;
                }
            }
        }
        if constexpr (!GROUP) {
#pragma unroll
            for (int m = 0; m < NPB; m++) {
                const int c = tid + m * NTHR;
                cell2d(x0 + c % TX, y0 + c / TX, rcfb[m], ria_b[m]);
            }
        }
        const size_t lay = (size_t)k * A.n2;
        // tracer tiles and output cells go through buffer resources of one layer of
        // one tracer (base and extent in SGPRs), lanes with nothing to access taking
        // the offset RSRC_SKIP: every lane issues the same loads and stores in every
        // iteration of the tracer loop, without exec-mask branches, so the compiler
        // counts the memory operations in flight and waits for a tile's loads
        // without waiting for the previous tracer's stores.  Tracers t >= ntr get an
        // empty resource (nothing is read).
        const unsigned lay_bytes = (unsigned)A.n2 * (unsigned)sizeof(T);
        T pf[NPIN];
        short pl[C ? NPIN : 1];
        auto load_tile = [&](int t) {
            const Rsrc rf = make_rsrc(fin + (size_t)t * A.n3 + lay, t < ntr ? lay_bytes : 0u);
#pragma unroll
            for (int m = 0; m < NPIN; m++)
                pf[m] = ld_val<T>(rf, toff[m] >= 0 ? (unsigned)toff[m] * (unsigned)sizeof(T) : RSRC_SKIP);
            if constexpr (C) {
                const Rsrc rl = make_rsrc(lin + (size_t)t * A.n3 + lay, t < ntr ? (unsigned)A.n2 * 2u : 0u);
#pragma unroll
                for (int m = 0; m < NPIN; m++)
                    pl[m] = ld_i16(rl, toff[m] >= 0 ? (unsigned)toff[m] * 2u : RSRC_SKIP, 0);
            }
        };
        auto store_tile = [&](int buf) {
            R* const dst = s_in(buf);
#pragma unroll
            for (int m = 0; m < NPIN; m++) {
                const int c = tid + m * NTHR;
                if constexpr (C) {
                    if (c < NT) dst[c] = cdecode(pf[m], pl[m]);
                } else {
                    if (c < NT) dst[c] = pf[m];
                }
            }
        };
        // ---- every per-layer load of the set-up in flight at once
        double va[NPFA], ha[NPFA], vb[NPFB], hb[NPFB], h0a[NPA];
#pragma unroll
        for (int m = 0; m < NPFA; m++) {
            va[m] = 0.0;
            ha[m] = 0.0;
            if (foa[m] >= 0) {
                va[m] = at32(D1 == 1 ? A.u + lay : A.v + lay, 8u * (unsigned)foa[m]);
                ha[m] = at32(D1 == 1 ? A.hu + lay : A.hv + lay, 8u * (unsigned)foa[m]);
            }
        }
#pragma unroll
        for (int m = 0; m < NPFB; m++) {
            vb[m] = 0.0;
            hb[m] = 0.0;
            if (fob[m] >= 0) {
                vb[m] = at32(D2 == 1 ? A.u + lay : A.v + lay, 8u * (unsigned)fob[m]);
                hb[m] = at32(D2 == 1 ? A.hu + lay : A.hv + lay, 8u * (unsigned)fob[m]);
            }
        }
#pragma unroll
        for (int m = 0; m < NPA; m++)
            h0a[m] = coa[m] >= 0 ? at32(((wet_a >> m) & 1u ? hin : hin_dry) + lay, 8u * (unsigned)coa[m]) : 0.0;
        if (ntr > 0) load_tile(0);

        // ---- stage A faces: transports to LDS for the thickness update
        FK fa[NPFA];
#pragma unroll
        for (int m = 0; m < NPFA; m++) {
            const int c = tid + m * NTHR;
            FaceC fc{0.0, 0.0, 0};
            if (c < NFA) {
                if constexpr (GROUP) fc = face_make2(s_ffa[c], va[m], ha[m], s_mqa[c], s_mca[c], dt);
                else fc = face_make2(rffa[m], va[m], ha[m], rmqa[m], rmca[m], dt);
            }
            if (c < NFA) s_q[c] = fc.q;
            const int base = (FAY0 + c / FAW + 2) * HX + (FAX0 + c % FAW + 2);
            fa[m].set(fc, c < NFA ? base : 0, SD1);
        }
        if (ntr > 0) store_tile(0);
        __syncthreads();
        // ---- stage A cells: h1 = h0 - (dt (Q1 - Q0)) iA
        CellK<R> ca[NPA];
#pragma unroll
        for (int m = 0; m < NPA; m++) {
            const int c = tid + m * NTHR;
            ca[m].upd = 0;
            ca[m].a = R(0);
            ca[m].b = R(1);
            ca[m].iA = R(0);
            if (c < NA) {
                const int xl = AX0 + c % AW, yl = AY0 + c / AW;
                const int own = (yl + 2) * HX + (xl + 2);
                const int f1 = (yl - FAY0) * FAW + (xl - FAX0);
                ca[m].own = own;
                ca[m].fl1 = f1;
                ca[m].rb = sr_cell_bits(ss_a, (unsigned)lay + (unsigned)(mul24(y0 + yl, A.nx_) + x0 + xl));
                double h1 = 0.0;
                if (coa[m] >= 0) {
                    const bool upd = GROUP ? (s_cf[own] & 2) : (rcfa[m] & 2);
                    const double ia = GROUP ? s_ia[own] : ria_a[m];
                    h1 = h0a[m];
                    if (upd) h1 = h0a[m] - (dt * (s_q[f1] - s_q[f1 - (D1 == 1 ? 1 : FAW)])) * ia;
                    cell_set<R, C>(ca[m], upd, h0a[m], h1, ia, dt);
                }
                s_h[own] = h1;
            }
        }
        __syncthreads();
        // ---- stage B faces
        FK fb[NPFB];
#pragma unroll
        for (int m = 0; m < NPFB; m++) {
            const int c = tid + m * NTHR;
            FaceC fc{0.0, 0.0, 0};
            if (c < NFB) {
                if constexpr (GROUP) fc = face_make2(s_ffb[c], vb[m], hb[m], s_mqb[c], s_mcb[c], dt);
                else fc = face_make2(rffb[m], vb[m], hb[m], rmqb[m], rmcb[m], dt);
            }
            if (c < NFB) s_q[c] = fc.q;
            const int base = (FBY0 + c / FBW + 2) * HX + (FBX0 + c % FBW + 2);
            fb[m].set(fc, c < NFB ? base : 0, SD2);
        }
        __syncthreads();
        // ---- stage B cells: h2 = h1 - (dt (Q1 - Q0)) iA, written as the pair's thickness
        CellK<R> cb[NPB];
#pragma unroll
        for (int m = 0; m < NPB; m++) {
            const int c = tid + m * NTHR;
            const int xl = c % TX, yl = c / TX;
            const int own = (yl + 2) * HX + (xl + 2);
            const int f1 = (yl - FBY0) * FBW + (xl - FBX0);
            cb[m].upd = 0;
            cb[m].a = R(0);
            cb[m].b = R(1);
            cb[m].iA = R(0);
            cb[m].own = own;
            cb[m].fl1 = f1;
            cb[m].rb = sr_cell_bits(ss_b, (unsigned)lay + (unsigned)ooff[m]);
            if (ooff[m] >= 0) {
                const double h1 = s_h[own];
                const bool upd = GROUP ? (s_cf[own] & 2) : (rcfb[m] & 2);
                const double ia = GROUP ? s_ia[own] : ria_b[m];
                double h2 = h1;
                if (upd) h2 = h1 - (dt * (s_q[f1] - s_q[f1 - (D2 == 1 ? 1 : FBW)])) * ia;
                cell_set<R, C>(cb[m], upd, h1, h2, ia, dt);
                if (upd || write_dry) at32(hout + lay, 8u * (unsigned)ooff[m]) = h2;
            }
        }

        // ---- tracer loop: tile t+1 was loaded during iteration t-1 and is stored to
        // LDS now, tile t+2 is loaded during this iteration (three LDS tile buffers)
        // the prefetches past the last tracer read nothing (empty resource) and fill
        // a tile buffer that no tracer reads; they keep the loop's memory operations
        // the same in every iteration
        load_tile(1);
        // as many stores (to nowhere) as a tracer's stage-B cells issue: entering the
        // loop, tile 1's loads then have the same stores after them as every later
        // tile's loads have (the previous tracer's), so the wait for them is the same
        {
            const Rsrc none = make_rsrc(fout, 0u);
#pragma unroll
            for (int m = 0; m < NPB; m++) {
                st_val<T>(none, RSRC_SKIP, T(0));
                if constexpr (C) st_i16(none, RSRC_SKIP, 0, 0);
            }
        }
        __syncthreads();  // s_q and s_h (read above) share the pool with the tile buffers
        for (int t = 0; t < ntr; t++) {
            const R* sin = s_in(t % 3);
            const SRound sr_a = tracer_round(ss_a, tg0 + t), sr_b = tracer_round(ss_b, tg0 + t);
            store_tile((t + 1) % 3);
            load_tile(t + 2);
            // stage A faces
#pragma unroll
            for (int m = 0; m < NPFA; m++) {
                const int c = tid + m * NTHR;
                if (c < NFA) {
                    R fl = fa[m].flux(sin);
                    if constexpr (HDIFF) {
                        if (fa[m].open) {
                            const int xs = FAX0 + c % FAW, ys = FAY0 + c / FAW;
                            const int base = (ys + 2) * HX + xs + 2;
                            fl = fl - ah_flux<D1, R>(A, x0 + xs, y0 + ys, k, sin[base], sin[base + SD1]);
                        }
                    }
                    s_fa[c] = fl;
                }
            }
            __syncthreads();
            // stage A cells
#pragma unroll
            for (int m = 0; m < NPA; m++) {
                const int c = tid + m * NTHR;
                if (c < NA) {
                    // formed for every cell and selected (no branch); cells not updated keep
                    // their value
                    const R f = sin[ca[m].own];
                    const R upd = cell_update<R, C>(ca[m], dt, f, s_fa[ca[m].fl1 - (D1 == 1 ? 1 : FAW)],
                                                    s_fa[ca[m].fl1], sr_a);
                    s_a[ca[m].own] = ca[m].upd ? upd : f;
                }
            }
            __syncthreads();
            // stage B faces
#pragma unroll
            for (int m = 0; m < NPFB; m++) {
                const int c = tid + m * NTHR;
                if (c < NFB) {
                    R fl = fb[m].flux(s_a);
                    if constexpr (HDIFF) {
                        if (fb[m].open) {
                            const int xs = FBX0 + c % FBW, ys = FBY0 + c / FBW;
                            const int base = (ys + 2) * HX + xs + 2;
                            fl = fl - ah_flux<D2, R>(A, x0 + xs, y0 + ys, k, s_a[base], s_a[base + SD2]);
                        }
                    }
                    s_fb[c] = fl;
                }
            }
            __syncthreads();
            // stage B cells -> HBM
            const Rsrc ro = make_rsrc(fout + (size_t)t * A.n3 + lay, lay_bytes);
#pragma unroll
            for (int m = 0; m < NPB; m++) {
                const R f = s_a[cb[m].own];
                const R upd = cell_update<R, C>(cb[m], dt, f, s_fb[cb[m].fl1 - (D2 == 1 ? 1 : FBW)], s_fb[cb[m].fl1],
                                                sr_b);
                const R out = cb[m].upd ? upd : f;
                // FP32: only updated cells are stored (the scratch buffer's other cells are
                // never read, and the tracers' own keep their values); FP64 stores every
                // cell, as its step ends with a copy of the whole scratch buffer
                const bool st = (C || std::is_same<R, float>::value) ? (bool)cb[m].upd : ooff[m] >= 0;
                if constexpr (C) {
                    float hi;
                    short lo;
                    cencode(out, hi, lo);
                    const Rsrc rlo = make_rsrc(lout + (size_t)t * A.n3 + lay, (unsigned)A.n2 * 2u);
                    st_f32(ro, st ? (unsigned)ooff[m] * 4u : RSRC_SKIP, 0, hi);
                    st_i16(rlo, st ? (unsigned)ooff[m] * 2u : RSRC_SKIP, 0, lo);
                } else {
                    st_val<T>(ro, st ? (unsigned)ooff[m] * (unsigned)sizeof(T) : RSRC_SKIP, out);
                }
            }
        }
        // no barrier between tracers is needed after the stage-B cells: they read s_a
        // and s_fb, which the next tracer writes only after its first barrier, and the
        // tile buffer the next tracer fills was last read two tracers back.  The last
        // tracer's stage-B cells still read the pool, where the next layer's set-up
        // writes s_q and s_h.
        __syncthreads();
    }
}

// Single domain, 1-D grid in XCD-aware order.  Work-groups go to the 8 XCDs
// round-robin (block b to XCD b % 8), and each XCD has its own L2, so block b
// takes logical item v = (b % 8) * (gridDim.x / 8) + b / 8: every XCD walks a
// contiguous range of the logical order (tracer batch fastest, then layer
// group, then tile; the compensated form: tile, then layer).  With layer
// groups first the blocks resident on one XCD are the same tile at consecutive
// layer groups, and the tile's 2-D fields (3 masks, 5 metrics) come from that
// XCD's L2 instead of being fetched again from HBM for every group.
// Every tracer-batch block writes the same thickness.
// FP32: at most 128 VGPRs (four waves per SIMD), two 512-lane blocks per CU at TY = 8.
// KG: layers per block (1 or HPAIR_KG); HDIFF: horizontal diffusion (A.ah_u set).
// Occupancy floor of the pair kernels: FP32 four waves per SIMD (128 VGPRs), FP64
// two; the compensated form's 768-lane blocks (12-row tiles) six and its 512-lane
// blocks (16-row tiles, two cells per lane) four, so that two blocks share a CU and
// cover each other's barrier phases.
// LPC: output cells per lane (1, or 2 for the compensated form's 16-row tiles).
template <typename R, bool C, bool HDIFF>
constexpr int hpair_waves_per_eu(int lanes)
{
    return sizeof(R) == 4 ? 4 : (C && !HDIFF && lanes == 768 ? 6 : (C && !HDIFF && lanes == 512 ? 4 : 2));
}

template <int S, typename R, int D1, int TY, int KG, bool HDIFF, int TXW, bool C = false, int LPC = 1>
__global__ void __launch_bounds__(TXW * TY / LPC) __attribute__((amdgpu_waves_per_eu(hpair_waves_per_eu<R, C, HDIFF>(TXW * TY / LPC)))) k_hpair(const StepArgs A, double dt, int ntr_all, int tpb, int ntx,
                                                   int nty, int nbatch, int ox, int oy,
                                                   const std::conditional_t<C, float, R>* __restrict__ fin_all,
                                                   std::conditional_t<C, float, R>* __restrict__ fout_all,
                                                   const double* __restrict__ hin,
                                                   double* __restrict__ hout, unsigned sr_stage,
                                                   const double* hin_dry, int write_dry,
                                                   const short* __restrict__ lin_all, short* __restrict__ lout_all)
{
    const unsigned b = blockIdx.x, per = gridDim.x / 8;
    unsigned v = (b % 8) * per + b / 8;
    const unsigned ngroups = (unsigned)(A.nz + KG - 1) / KG;
    const unsigned nitems = (unsigned)nbatch * ngroups * ntx * nty;
    if (v >= nitems) return;
    const int tbatch = v % nbatch;
    v /= nbatch;
    int k0, tx, ty;
    if constexpr (C) {
        // compensated form (many tracers per block): tiles before layers, so that
        // neighbouring tiles of a layer share their halo lines in L2 (5% faster at
        // 256 tracers)
        tx = v % ntx;
        v /= ntx;
        ty = v % nty;
        k0 = (int)(v / nty) * KG;
    } else {
        // layer groups before tiles: a tile's 2-D masks and metrics stay in L2 for
        // all its layers (and the index arithmetic keeps the FP64 and FP32 kernels
        // at 68-80 VGPRs, three 512-lane blocks per CU: 35% faster in FP64)
        k0 = (int)(v % ngroups) * KG;
        v /= ngroups;
        tx = v % ntx;
        ty = v / ntx;
    }
    const int k1 = k0 + KG < A.nz ? k0 + KG : A.nz;
    const int tb0 = tbatch * tpb;
    const int ntr = ntr_all - tb0 < tpb ? ntr_all - tb0 : tpb;
    const size_t off = (size_t)tb0 * A.n3;
    hpair_body<S, R, D1, TY, (KG > 1), HDIFF, TXW, C, LPC>(A, dt, ox + tx * TXW, oy + ty * TY, k0, k1, ntr, fin_all + off,
                                                      fout_all + off, hin, hout, sr_stage, tb0, hin_dry,
                                                      write_dry != 0, C ? lin_all + off : nullptr,
                                                      C ? lout_all + off : nullptr);
}

// Ensemble: one work item per block {domain, x0, y0, first layer of the group | batch << 16}; the
// table is built in the XCD-aware order of k_hpair (domain -1 = padding).
// STAGE 1: stages 1-2 (tracers -> scratch, ho -> hA); STAGE 2: stages 4-5
// (tracers -> scratch, hB -> hA).
template <int S, typename R, int D1, int TY, int STAGE, int KG, int TXW = TX, int LPC = 1>
__global__ void __launch_bounds__(TXW * TY / LPC) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 2))) k_hpair_multi(const EnsDom<R>* __restrict__ doms,
                                                         const int4* __restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    if (wk.x < 0) return;  // padding of the XCD-ordered table
    const EnsDom<R>& D = doms[wk.x];
    const int k0 = wk.w & 0xffff, tb0 = (wk.w >> 16) * D.tpb_h;
    const int k1 = k0 + KG < D.A.nz ? k0 + KG : D.A.nz;
    const int ntr = D.ntr - tb0 < D.tpb_h ? D.ntr - tb0 : D.tpb_h;
    const size_t off = (size_t)tb0 * D.A.n3;
    hpair_body<S, R, D1, TY, (KG > 1), false, TXW, false, LPC>(D.A, 0.5 * D.dt, wk.y, wk.z, k0, k1, ntr, D.tracers + off,
                                              D.scratch + off,
                             STAGE == 1 ? D.ho : D.hB, D.hA, STAGE == 1 ? SR_PAIR1 : SR_PAIR2, tb0,
                             // stages 4-5 end in hA, where stages 1-2 left the cells
                             // that nothing updates (the w stage does not copy them)
                             STAGE == 1 ? D.ho : D.hA, STAGE == 1);
}

// ---------------------------------------------------------------------------
// Vertical stage, one lane per column, TB tracers per sweep.
// advection.F90.template:223-374.  The sub-step count and every restart depend
// only on h and w, so the count is decided once from a thickness-only replay
// of the reference's control flow, and the tracer pass then repeats the
// reference's final (restart-free) pass.  The thickness after s sub-steps is
// recomputed per layer by repeating the reference's own update
// h <- h - dtk (w_k - w_{k-1}) s times, so nothing is kept in memory and
// blocks that handle different tracer batches of a column never share
// scratch; batch 0 writes the final thickness.  Most columns need one
// sub-step: the tracer pass runs first with one, taking the Courant maximum
// the replay would take, and only a column that needs more is replayed and
// swept again (from fin, so fin and fout must be different buffers).
// ---------------------------------------------------------------------------
// C: compensated FP32 storage (R = double): fin/fout hold FP32 values and
// lin/lout their int16 residuals; the column is decoded to FP64, limited in
// FP32 (LimF on the FP64 differences) and updated in FP64 by the increment form.
template <int S, typename R, int TB, bool C = false>
__device__ __forceinline__ void wcol_body(const StepArgs& A, double dt, int i, int j, int tb0, int tb1, bool first,
                                          bool copy_dry,
                                          const std::conditional_t<C, float, R>* __restrict__ fin,
                                          std::conditional_t<C, float, R>* __restrict__ fout,
                                          const double* __restrict__ hin, double* __restrict__ hout,
                                          const double* __restrict__ w, const double* __restrict__ wv,
                                          int* __restrict__ stats, const short* __restrict__ lin = nullptr,
                                          short* __restrict__ lout = nullptr)
{
    static_assert(!C || std::is_same<R, double>::value, "compensated storage works on FP64 values");
    using T = std::conditional_t<C, float, R>;
    if (i >= A.nx_ || j >= A.ny_) return;
    const size_t n2 = A.n2, n3 = A.n3;
    const int nz = A.nz;
    const size_t p = (size_t)j * A.nx_ + i;
    const bool inside = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax;
    if (!inside || A.tmask[p] != 1) {
        // copy_dry = false: the next stage reads these columns' thickness from hin
        if (first && copy_dry)
            for (int k = 0; k < nz; k++) hout[(size_t)k * n2 + p] = hin[(size_t)k * n2 + p];
        // FP32: the tracers of columns that are not updated keep their values in fout
        // (no stage writes them); FP64 copies them (its step ends with a full copy)
        if constexpr (std::is_same<R, double>::value && !C)
            for (int t = tb0; t < tb1; t++)
                for (int k = 0; k < nz; k++)
                    fout[(size_t)t * n3 + (size_t)k * n2 + p] = fin[(size_t)t * n3 + (size_t)k * n2 + p];
        return;
    }
    const double* hi = hin + p;
    const double* wp = w + p;
    const double* wvp = wv + p;
    double dtk = dt;
    // thickness of 0-based layer kk after s sub-steps
    auto h_at = [&](int kk, int steps) {
        double h = hi[(size_t)kk * n2];
        if (steps > 0) {
            const double dw = wp[(size_t)(kk + 1) * n2] - wp[(size_t)kk * n2];
            for (int st = 0; st < steps; st++) h = h - dtk * dw;
        }
        return h;
    };

    // ---- tracer pass for TB tracers at a time, sharing the thickness sweep.  With
    // cmax != nullptr (only for iters == 1) it also returns the largest Courant
    // number at the initial thickness, in the order of the reference's MAXVAL.
    auto sweep = [&](int t0, int iters, double itersm1, double* cmax) {
        const int nb = tb1 - t0 < TB ? tb1 - t0 : TB;
        double c = -1.7976931348623157e308;  // MAXVAL of an empty array
        for (int it = 1; it <= iters; it++) {
            const T* sf = (it == 1 ? fin : fout) + (size_t)t0 * n3 + p;
            T* df = fout + (size_t)t0 * n3 + p;
            const short* sl = C ? (it == 1 ? lin : lout) + (size_t)t0 * n3 + p : nullptr;
            short* dl = C ? lout + (size_t)t0 * n3 + p : nullptr;
            // tracer b of the sweep at in-tracer offset o
            auto ld = [&](int b, size_t o) -> R {
                if constexpr (C) return cdecode(sf[(size_t)b * n3 + o], sl[(size_t)b * n3 + o]);
                else return sf[(size_t)b * n3 + o];
            };
            SRound sr[TB];
            const unsigned ss = stage_seed<R>(A, SR_WCOL + it);
#pragma unroll
            for (int b = 0; b < TB; b++) sr[b] = tracer_round(ss, t0 + b);
            const bool write_h = first && t0 == tb0 && it == iters;
            R fm1[TB], f0[TB], f1[TB], fprev[TB];
            double h0 = h_at(0, it - 1), h1 = nz > 1 ? h_at(1, it - 1) : 0.0;
#pragma unroll
            for (int b = 0; b < TB; b++) {
                fm1[b] = R(0);
                fprev[b] = R(0);
                f0[b] = b < nb ? ld(b, 0) : R(0);
                f1[b] = (b < nb && nz > 1) ? ld(b, n2) : R(0);
            }
            double wprev = wp[0];
            for (int k = 1; k <= nz; k++) {
                double h2 = 0.0;
                R f2[TB];
                const bool has2 = k + 2 <= nz;
                if (has2) h2 = h_at(k + 1, it - 1);
#pragma unroll
                for (int b = 0; b < TB; b++) f2[b] = (has2 && b < nb) ? ld(b, (size_t)(k + 1) * n2) : R(0);
                const double wk = wp[(size_t)k * n2];
                const double hnew = h0 - dtk * (wk - wprev);
                R fl[TB];
                if (k <= nz - 1) {
                    const double wvk = wvp[(size_t)k * n2];
                    const double c0 = (fabs(wvk) * dt) / fmin2(h0, h1);
                    if (cmax) c = fmax2(c, c0);
                    const double cfl = c0 * itersm1;
                    const R rw = R(wvk);
                    const bool pos = wvk > 0.0;
                    LimF<S> lf;
                    if constexpr (!std::is_same<R, double>::value || C) lf.set(cfl);
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
                        if constexpr (C)
                            fl[b] = rw * (fu + (double)lf.corr(float(fd - fu), float(fu - fuu)));
                        else if constexpr (std::is_same<R, double>::value)
                            fl[b] = rw * face_value<S, double>(cfl, fu, fuu, fd);
                        else
                            fl[b] = rw * lf.value(fu, fuu, fd);
                    }
                } else {
#pragma unroll
                    for (int b = 0; b < TB; b++) fl[b] = R(0);
                }
                if (write_h) hout[(size_t)(k - 1) * n2 + p] = hnew;
                if constexpr (C) {
                    // increment form in FP64: f + ((h0 - hnew)/hnew f - dt/hnew dF)
                    const double rh = rcp_nr(hnew);
                    const double ca = (h0 - hnew) * rh, cb = dtk * rh;
#pragma unroll
                    for (int b = 0; b < TB; b++)
                        if (b < nb) {
                            float hi;
                            short lo;
                            cencode(f0[b] + fma(ca, f0[b], -(cb * (fl[b] - fprev[b]))), hi, lo);
                            df[(size_t)b * n3 + (size_t)(k - 1) * n2] = hi;
                            dl[(size_t)b * n3 + (size_t)(k - 1) * n2] = lo;
                        }
                } else if constexpr (std::is_same<R, double>::value) {
#pragma unroll
                    for (int b = 0; b < TB; b++) {
                        const double hfo = h0 * f0[b];
                        const double advn = fl[b] - fprev[b];
                        if (b < nb) df[(size_t)b * n3 + (size_t)(k - 1) * n2] = (hfo - dtk * advn) / hnew;
                    }
                } else {
                    // increment form, as in cell_update: f + ((h0 - hnew)/hnew f - dt/hnew dF)
                    const double rh = rcp_nr(hnew);
                    const R ca = R((h0 - hnew) * rh), cb = R(dtk * rh);
                    const unsigned rb = sr_cell_bits(ss, (unsigned)((size_t)(k - 1) * n2 + p));
#pragma unroll
                    for (int b = 0; b < TB; b++)
                        if (b < nb)
                            df[(size_t)b * n3 + (size_t)(k - 1) * n2] =
                                sr[b].add(f0[b], __builtin_fmaf(ca, f0[b], -(cb * (fl[b] - fprev[b]))), rb);
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
        if (cmax) *cmax = c;
    };

    // ---- first tracer group with one sub-step, taking the Courant maximum on the
    // way (the same expressions, in the same order, as the replay's first check):
    // when it needs no more than one sub-step, that pass is already the result.
    int iters = 1, capped = 0;
    double itersm1 = 1.0;
    {
        double c1;
        sweep(tb0, 1, 1.0, &c1);
        const int ce = fceiling(c1);
        if (ce <= 1) {
            for (int t0 = tb0 + TB; t0 < tb1; t0 += TB) sweep(t0, 1, 1.0, nullptr);
            return;
        }
    }

    // ---- thickness-only replay of the iteration control (template:277-322)
    {
        int it = 1;
        while (it <= iters) {
            if (iters < kItersMax) {
                double c = -1.7976931348623157e308;  // MAXVAL of an empty array
                double hk = h_at(0, it - 1);
                for (int k = 1; k <= nz - 1; k++) {
                    const double hk1 = h_at(k, it - 1);
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
            if (iters == 1) break;
            it++;
        }
    }
    if (first) {
        if (iters > 1) atomicMax(&stats[0], iters);
        if (capped) atomicAdd(&stats[1], 1);
    }
    for (int t0 = tb0; t0 < tb1; t0 += TB) sweep(t0, iters, itersm1, nullptr);
}
template <int S, typename R, int TB, bool C = false>
__global__ void __launch_bounds__(256) k_wcol(const StepArgs A, double dt, int ntr, int tpb,
                                              const std::conditional_t<C, float, R>* __restrict__ fin,
                                              std::conditional_t<C, float, R>* __restrict__ fout,
                                              const double* __restrict__ hin,
                                              double* __restrict__ hout, const double* __restrict__ w,
                                              const double* __restrict__ wv, int* __restrict__ stats, int copy_dry,
                                              const short* __restrict__ lin, short* __restrict__ lout)
{
    unsigned bx, by, bz;
    if (!xcd_block((A.nx_ + 63) / 64, (A.ny_ + 3) / 4, (ntr + tpb - 1) / tpb, bx, by, bz)) return;
    const int tb0 = bz * tpb;
    const int tb1 = tb0 + tpb < ntr ? tb0 + tpb : ntr;
    wcol_body<S, R, TB, C>(A, dt, bx * 64 + threadIdx.x, by * 4 + threadIdx.y, tb0, tb1, bz == 0, copy_dry != 0, fin,
                           fout, hin, hout, w, wv, stats, lin, lout);
}

// Compensated w stage for one row of 64 columns and a batch of tracers, NW waves
// per block.  The tracer-independent terms of every level are formed once per
// block, in parallel over levels, into LDS: per cell the increment
// coefficients a = (h0 - hnew)/hnew, b = dt/hnew of cell_update (FP64), per
// face the transport w (FP64) and the limiter's cfl terms (LimF), with the
// reference's expressions (advection.F90.template:239-322 for one sub-step).
// Columns whose Courant maximum needs a single sub-step are then swept from
// registers: each wave takes a contiguous share of the tracers and, per tracer,
// loads the whole column (FP32 values and int16 residuals) at once, forms the
// faces from a window of decoded values and stores every level once.  The few
// columns that need sub-steps take wcol_body (thickness replay, FP64 windows).
// Levels above the surface (nz .. NZMAX - 1) have zero coefficients and read
// the surface cell, so the unrolled sweep has no level tests but the stores';
// the reference's zero flux through the surface is the zero transport there.
template <int S, int NZMAX, int NW>
struct WcolLds {
    double a[NZMAX][64], b[NZMAX][64], q[NZMAX][64], c0[NZMAX][64];
    float ch[NZMAX][64];
    float l[LimF<S>::NL > 0 ? LimF<S>::NL : 1][NZMAX][64];
    int single[64];
};

template <int S, int NZMAX, int NW>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4)))
k_wcol_c(const StepArgs A, double dt, int ntr_all, int tpb, const float* __restrict__ fin_all,
         float* __restrict__ fout_all, const double* __restrict__ hin, double* __restrict__ hout,
         const double* __restrict__ w, const double* __restrict__ wv, int* __restrict__ stats,
         const short* __restrict__ lin_all, short* __restrict__ lout_all)
{
    static_assert(NZMAX % 8 == 0, "depth classes are multiples of 8 layers");
    extern __shared__ __attribute__((aligned(16))) unsigned char s_raw[];
    WcolLds<S, NZMAX, NW>& L = *reinterpret_cast<WcolLds<S, NZMAX, NW>*>(s_raw);
    constexpr int NL = LimF<S>::NL;
    const int nz = A.nz;
    const size_t n2 = A.n2, n3 = A.n3;
    unsigned bx, by, bz;
    if (!xcd_block((A.nx_ + 63) / 64, A.ny_, (ntr_all + tpb - 1) / tpb, bx, by, bz)) return;
    const int lane = threadIdx.x, tid = threadIdx.y * 64 + threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int j = by;
    auto column = [&](int l, size_t& p) {  // wet interior column of lane l of the row?
        const int i = bx * 64 + l;
        p = (size_t)j * A.nx_ + (i < A.nx_ ? i : 0);
        return i < A.nx_ && i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax && A.tmask[p] == 1;
    };
    // ---- per-level terms, items (level, lane)
    for (int it = tid; it < NZMAX * 64; it += 64 * NW) {
        const int k = it / 64, l = it % 64;
        size_t p;
        const bool wet = column(l, p);
        double a = 0.0, b = 0.0, q = 0.0, c0 = 0.0;
        LimF<S> lf;
        lf.set(0.0);
        if (wet && k < nz) {
            // cell k: one sub-step of the thickness (template:277-290)
            const double h0 = hin[(size_t)k * n2 + p];
            const double hnew = h0 - dt * (w[(size_t)(k + 1) * n2 + p] - w[(size_t)k * n2 + p]);
            const double rh = rcp_nr(hnew);
            a = (h0 - hnew) * rh;
            b = dt * rh;
            if (k >= 1) {
                // face k between cells k - 1 and k
                const double wvk = wv[(size_t)k * n2 + p];
                c0 = (fabs(wvk) * dt) / fmin2(hin[(size_t)(k - 1) * n2 + p], h0);
                q = wvk;
                lf.set(c0 * 1.0);
            }
        }
        L.a[k][l] = a;
        L.b[k][l] = b;
        L.q[k][l] = q;
        L.c0[k][l] = c0;
        L.ch[k][l] = lf.ch;
#pragma unroll
        for (int m = 0; m < NL; m++) L.l[m][k][l] = lf.l[m];
    }
    __syncthreads();
    if (wave == 0) {
        // Courant maximum in the order of the reference's MAXVAL: one sub-step?
        double c = -1.7976931348623157e308;
        for (int k = 1; k <= nz - 1; k++) c = fmax2(c, L.c0[k][lane]);
        L.single[lane] = fceiling(c) <= 1;
    }
    __syncthreads();
    const int tb0 = bz * tpb;
    const int ntr = ntr_all - tb0 < tpb ? ntr_all - tb0 : tpb;
    if (bz == 0) {
        // thickness of single-step columns (the others' is written by wcol_body)
        for (int it = tid; it < NZMAX * 64; it += 64 * NW) {
            const int k = it / 64, l = it % 64;
            size_t p;
            if (k < nz && column(l, p) && L.single[l]) {
                const double h0 = hin[(size_t)k * n2 + p];
                hout[(size_t)k * n2 + p] = h0 - dt * (w[(size_t)(k + 1) * n2 + p] - w[(size_t)k * n2 + p]);
            }
        }
    }
    size_t p;
    const bool wet = column(lane, p);
    if (!wet) return;
    // this wave's tracers
    const int per = (ntr + NW - 1) / NW;
    const int t_lo = tb0 + wave * per, t_hi = t_lo + per < tb0 + ntr ? t_lo + per : tb0 + ntr;
    if (t_lo >= t_hi) return;
    if (!L.single[lane]) {
        wcol_body<S, double, 2, true>(A, dt, bx * 64 + lane, j, t_lo, t_hi, bz == 0 && t_lo == tb0, false, fin_all,
                                      fout_all, hin, hout, w, wv, stats, lin_all, lout_all);
        return;
    }
    const unsigned v4 = (unsigned)p * 4u, v2 = (unsigned)p * 2u;
    for (int t = t_lo; t < t_hi; t++) {
        // the layer size and the lane through an empty asm: per-level offsets and LDS
        // loads are formed where used, not hoisted out of the tracer loop
        unsigned m2 = (unsigned)n2;
        asm volatile("" : "+s"(m2));
        int ln = lane;
        asm volatile("" : "+v"(ln));
        auto lev = [&](int k) -> unsigned { return (unsigned)(k < nz ? k : nz - 1) * m2; };
        const size_t o = (size_t)t * n3;
        const Rsrc rfi = make_rsrc(fin_all + o, (unsigned)(n3 * 4)), rli = make_rsrc(lin_all + o, (unsigned)(n3 * 2));
        const Rsrc rfo = make_rsrc(fout_all + o, (unsigned)(n3 * 4)), rlo = make_rsrc(lout_all + o, (unsigned)(n3 * 2));
        float hi[NZMAX];
        unsigned lp[NZMAX / 2];
#pragma unroll
        for (int k = 0; k < NZMAX; k++) {
            hi[k] = ld_f32(rfi, v4, lev(k) * 4u);
            const unsigned l = (unsigned short)ld_i16(rli, v2, lev(k) * 2u);
            lp[k / 2] = k % 2 ? lp[k / 2] | (l << 16) : l;
        }
        auto value = [&](int k) -> double {
            return cdecode(hi[k], (short)(k % 2 ? lp[k / 2] >> 16 : lp[k / 2] & 0xffffu));
        };
        // window: fm2 = f[k-2], fm1 = f[k-1], f0 = f[k], f1 = f[k+1] at face k
        double fm2 = 0.0, fm1 = value(0), f0 = value(NZMAX > 1 ? 1 : 0), fprev = 0.0;
#pragma unroll
        for (int k = 1; k <= NZMAX; k++) {
            double fl = 0.0;
            const double f1 = value(k + 1 < NZMAX ? k + 1 : NZMAX - 1);
            if (k < NZMAX) {
                const double q = L.q[k][ln];
                LimF<S> lf;
                lf.ch = L.ch[k][ln];
#pragma unroll
                for (int m = 0; m < NL; m++) lf.l[m] = L.l[m][k][ln];
                const bool pos = q > 0.0;
                const double fu = pos ? fm1 : f0;
                const double fuu = pos ? (k > 1 ? fm2 : fm1) : f1;
                const double fd = pos ? f0 : fm1;
                fl = q * (fu + (double)lf.corr(float(fd - fu), float(fu - fuu)));
            }
            // cell k - 1, increment form in FP64: f + ((h0 - hnew)/hnew f - dt/hnew dF)
            const double a = L.a[k - 1][ln], b = L.b[k - 1][ln];
            const double x = fm1 + fma(a, fm1, -(b * (fl - fprev)));
            if (k - 1 < NZMAX - 8 || k - 1 < nz) {
                float h;
                short l;
                cencode(x, h, l);
                st_f32(rfo, v4, (unsigned)(k - 1) * m2 * 4u, h);
                st_i16(rlo, v2, (unsigned)(k - 1) * m2 * 2u, l);
            }
            fprev = fl;
            fm2 = fm1;
            fm1 = f0;
            f0 = f1;
        }
    }
}

// Ensemble: work item {domain, bx, by, batch}; scratch -> tracers, hA -> hB (updated columns only).
template <int S, typename R, int TB>
__global__ void __launch_bounds__(256) k_wcol_multi(const EnsDom<R>* __restrict__ doms, const int4* __restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const EnsDom<R>& D = doms[wk.x];
    const int tb0 = wk.w * D.tpb_w;
    const int tb1 = tb0 + D.tpb_w < D.ntr ? tb0 + D.tpb_w : D.ntr;
    wcol_body<S, R, TB>(D.A, D.dt, wk.y * 64 + threadIdx.x, wk.z * 4 + threadIdx.y, tb0, tb1, wk.w == 0, false, D.scratch,
                        D.tracers, D.hA, D.hB, D.w, D.w, D.stats);
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
// FP32 theta-method column solve in increment form.  With c = c_old + d the
// reference system (diffusion.F90:233-313, dzo = dzn = h)
//   -an_{k-1} c_{k-1} + (h_k + an_k + an_{k-1}) c_k - an_k c_{k+1} = h_k c_old_k + explicit part
// becomes the same matrix acting on d with right-hand side
//   an_k (c_{k+1} - c_k) + an_{k-1} (c_{k-1} - c_k) + ao_k (c_{k+1} - c_k) + ao_{k-1} (c_{k-1} - c_k)
// (old values), eliminated from the surface down as the reference does:
//   q_k = e_k (c_{k+1} - c_k + q_{k+1}) - ru_k (c_{k-1} - c_k) [+ g3_k (c_{k+1} - c_k) + g4_k (c_{k-1} - c_k)]
//   d_0 = q_0,  d_k = q_k - ru_k d_{k-1},  c_k = c_old_k + d_k
// with e_k = an_k/den_k, ru_k = -an_{k-1}/den_k, g3_k = ao_k/den_k, g4_k = ao_{k-1}/den_k
// formed in FP64 and rounded once.  The factors are the same every step in a
// steady state, so their rounding must not multiply c itself (a direct solve
// c_k = (h_k/den_k) c_old_k + ... would drift by up to 2^-25 of c per step); here it
// multiplies the differences only, and a well-mixed column stays exactly as it is.
// The old column and the eliminated q_k are both kept in registers; columns of
// nz <= 16 interleave two tracers, deeper ones take one tracer at a time.
// tg0: index in the domain of the first tracer of fin/fout.
// SRC: tracer sources ea4 (ntr x nz levels, FP64, the reference's ea4 = source *
// dt * scale with the surface flux in the top layer; cnpar = 1 only).  The
// reference solves (dz + K) c_new = dz c + ea4 (diffusion.F90:233-313); with
// s = ea4 / dz that is the solve from c + s, so the differences are taken of
// c + s and the stored value is c + (s + d), rounded once.
// Factorisation of one column from the surface down, FP64, into the FP32 LDS
// factors e2, ru (and g3, g4 unless cnpar = 1) of lane `lane`.  The
// coefficients of interface k (an, ao) serve level k as its lower and level
// k - 1 as its upper neighbour, so each is formed once; the reciprocals are
// refined hardware estimates (the factors are rounded to FP32 afterwards).
template <bool CN1>
__device__ __forceinline__ void vdiff_factor(const StepArgs& A, double dt, double cnpar, double molecular,
                                             const double* __restrict__ nuh, size_t p, int lane, float* e2, float* ru,
                                             float* g3, float* g4)
{
    const size_t n2 = A.n2;
    const int nz = A.nz;
    const double* hn = A.hn + p;
    double ru_up = 0.0, an_k = 0.0, ao_k = 0.0;
    double dz = hn[(size_t)(nz - 1) * n2];
    for (int k = nz - 1; k >= 0; k--) {
        double an_km1 = 0.0, ao_km1 = 0.0, dz_lo = 0.0;
        if (k >= 1) {
            dz_lo = hn[(size_t)(k - 1) * n2];
            const double x = (2.0 * dt) * (nuh[(size_t)k * n2 + p] + molecular);
            const double rs = rcp_nr(dz + dz_lo);
            an_km1 = (cnpar * x) * rs;
            ao_km1 = ((1.0 - cnpar) * x) * rs;
        }
        const double a2 = (dz + an_k) + an_km1;
        const double den = k == nz - 1 ? a2 : a2 + an_k * ru_up;  // a2 - a3 ru_{k+1}, a3 = -an_k
        const double rden = rcp_nr(den);
        const double ruk = -an_km1 * rden;
        e2[k * 64 + lane] = float(an_k * rden);
        ru[k * 64 + lane] = float(ruk);
        if (!CN1) {
            g3[k * 64 + lane] = float(ao_k * rden);
            g4[k * 64 + lane] = float(ao_km1 * rden);
        }
        ru_up = ruk;
        an_k = an_km1;
        ao_k = ao_km1;
        dz = dz_lo;
    }
}

template <int NZMAX, bool CN1, bool SRC = false>
__device__ __forceinline__ void vdiff_body(const StepArgs& A, double dt, double cnpar, double molecular,
                                           const double* __restrict__ nuh, int i, int j, int lane, int ntr,
                                           const float* __restrict__ fin, float* __restrict__ fout, int tg0,
                                           const double* __restrict__ ea4 = nullptr)
{
    extern __shared__ float s_coef[];  // [ncoef][nz][64]
    constexpr int NT = NZMAX <= 16 ? 2 : 1;
    static_assert(NZMAX % 8 == 0, "depth classes are multiples of 8 layers");
    const int nz = A.nz;
    // the caller guarantees NZMAX - 8 < nz <= NZMAX, so only the top eight
    // unrolled levels need a run-time test
    auto has = [&](int k) { return k < NZMAX - 8 || k < nz; };
    auto has_up = [&](int k) { return k < NZMAX - 9 || k < nz - 1; };
    const size_t n2 = A.n2, n3 = A.n3;
    float* e2 = s_coef;
    float* ru = s_coef + (size_t)nz * 64;
    float* g3 = s_coef + (size_t)2 * nz * 64;
    float* g4 = s_coef + (size_t)3 * nz * 64;
    if (i >= A.nx_) return;
    const size_t p = (size_t)j * A.nx_ + i;
    const bool inside = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax;
    const bool wetcol = inside && A.tmask[p] == 1;
    const bool solve = wetcol && nz > 1;
    // columns that are not solved keep their values in fout (no stage writes them),
    // except that with nz = 1 every point takes the reference's (dz c)/dz
    if (!solve && nz > 1) return;
    if (solve) vdiff_factor<CN1>(A, dt, cnpar, molecular, nuh, p, lane, e2, ru, g3, g4);
    const unsigned ss = stage_seed<float>(A, SR_VDIFF);
    for (int t = 0; t < ntr; t += NT) {
        const int nt = ntr - t < NT ? ntr - t : NT;
        float c[NT][NZMAX], q[NT][NZMAX];
        float s[SRC ? NT : 1][SRC ? NZMAX : 1];
#pragma unroll
        for (int b = 0; b < NT; b++) {
            // nz = 1: the column's value is in fin if the stages updated it, else in fout
            const float* fi = (wetcol ? fin : fout) + (size_t)(t + (b < nt ? b : 0)) * n3 + p;
#pragma unroll
            for (int k = 0; k < NZMAX; k++)
                if (has(k)) c[b][k] = fi[(size_t)k * n2];
            if constexpr (SRC) {
                const double* e = ea4 + (size_t)(t + (b < nt ? b : 0)) * n3 + p;
#pragma unroll
                for (int k = 0; k < NZMAX; k++)
                    if (has(k)) s[b][k] = float(e[(size_t)k * n2] / A.hn[(size_t)k * n2 + p]);
            }
        }
        // value whose differences drive the solve: c, or c + s with sources
        auto cv = [&](int b, int k) -> float {
            if constexpr (SRC) return c[b][k] + s[b][k];
            else return c[b][k];
        };
        if (nz == 1) {
            // diffusion.F90:224-231: every point, halos and land included
            const double dz = A.hn[p];
#pragma unroll
            for (int b = 0; b < NT; b++) {
                if constexpr (SRC)  // diffusion.F90:224-231 with ea4
                    c[b][0] = float((dz * (double)c[b][0] + ea4[(size_t)(t + (b < nt ? b : 0)) * n3 + p]) / dz);
                else
                    c[b][0] = float((dz * (double)c[b][0]) / dz);
            }
        } else if (solve) {
            float q_up[NT];
#pragma unroll
            for (int b = 0; b < NT; b++) q_up[b] = 0.f;
#pragma unroll
            for (int k = NZMAX - 1; k >= 0; k--) {
                if (has(k)) {
                    const float ek = has_up(k) ? e2[k * 64 + lane] : 0.f;
                    const float rk = ru[k * 64 + lane];  // 0 at k = 0
                    float e3 = 0.f, e4 = 0.f;
                    if (!CN1) {
                        e3 = has_up(k) ? g3[k * 64 + lane] : 0.f;
                        e4 = g4[k * 64 + lane];
                    }
#pragma unroll
                    for (int b = 0; b < NT; b++) {
                        const float ck = cv(b, k);
                        const float dup = has_up(k) ? cv(b, k + 1 < NZMAX ? k + 1 : k) - ck : 0.f;
                        const float ddn = k > 0 ? cv(b, k > 0 ? k - 1 : 0) - ck : 0.f;
                        float qk = __builtin_fmaf(ek, dup + q_up[b], -(rk * ddn));
                        if constexpr (!CN1 && SRC) {
                            // the explicit part acts on the old values, without ea4/dz
                            // (diffusion.F90:238-262: a4 = c dzo + ea4 + auxo differences of c)
                            const float cu = has_up(k) ? c[b][k + 1 < NZMAX ? k + 1 : k] - c[b][k] : 0.f;
                            const float cd = k > 0 ? c[b][k > 0 ? k - 1 : 0] - c[b][k] : 0.f;
                            qk += e3 * cu + e4 * cd;
                        } else if (!CN1) {
                            qk += e3 * dup + e4 * ddn;
                        }
                        q[b][k] = qk;
                        q_up[b] = qk;
                    }
                }
            }
            // compiler-only barrier: ru is read again from LDS below instead of being
            // kept in registers from the elimination (nz more live values per lane)
            asm volatile("" ::: "memory");
            float d[NT];
            SRound sr[NT];
            const unsigned rb0 = sr_cell_bits(ss, (unsigned)p);
#pragma unroll
            for (int b = 0; b < NT; b++) {
                sr[b] = tracer_round(ss, tg0 + t + b);
                d[b] = q[b][0];
                if constexpr (SRC) c[b][0] = sr[b].add(c[b][0], s[b][0] + d[b], rb0);
                else c[b][0] = sr[b].add(c[b][0], d[b], rb0);
            }
#pragma unroll
            for (int k = 1; k < NZMAX; k++) {
                if (has(k)) {
                    const float r = ru[k * 64 + lane];
                    const unsigned rb = sr_cell_bits(ss, (unsigned)((size_t)k * n2 + p));
#pragma unroll
                    for (int b = 0; b < NT; b++) {
                        d[b] = __builtin_fmaf(-r, d[b], q[b][k]);
                        if constexpr (SRC) c[b][k] = sr[b].add(c[b][k], s[b][k] + d[b], rb);
                        else c[b][k] = sr[b].add(c[b][k], d[b], rb);
                    }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < NT; b++)
            if (b < nt) {
                float* fo = fout + (size_t)(t + b) * n3 + p;
#pragma unroll
                for (int k = 0; k < NZMAX; k++)
                    if (has(k)) fo[(size_t)k * n2] = c[b][k];
            }
    }
}
template <int NZMAX, bool CN1, bool SRC>
__global__ void __launch_bounds__(64) k_vdiff_f32(const StepArgs A, double dt, double cnpar, double molecular,
                                                  const double* __restrict__ nuh, int ntr_all, int tpb,
                                                  const float* __restrict__ fin_all, float* __restrict__ fout_all,
                                                  const double* __restrict__ ea4_all)
{
    const int tb0 = blockIdx.z * tpb;
    const int ntr = ntr_all - tb0 < tpb ? ntr_all - tb0 : tpb;
    vdiff_body<NZMAX, CN1, SRC>(A, dt, cnpar, molecular, nuh, blockIdx.x * 64 + threadIdx.x, blockIdx.y, threadIdx.x,
                                ntr, fin_all + (size_t)tb0 * A.n3, fout_all + (size_t)tb0 * A.n3, tb0,
                                SRC ? ea4_all + (size_t)tb0 * A.n3 : nullptr);
}

// Factorisation of vdiff_factor plus what the compensated solve needs per
// column: the content weights w / H (FP64, in LDS), H = sum_k h_k.
// Back-substitution d_k = q_k - ru_k d_{k-1} (FP32 ru as stored, FP64
// arithmetic) makes d a linear map of the eliminated right side q, so the
// column's content change sum_k h_k d_k equals sum_k w_k q_k with
//   w_{nz-1} = h_{nz-1},  w_k = h_k - ru_{k+1} w_{k+1},
// formed here once for every tracer of the column; the solve's uniform shift
// (sum_k h_k d_k) / H is then sum_k (w_k / H) q_k.
template <bool CN1>
__device__ __forceinline__ void vdiff_factor_w(const StepArgs& A, double dt, double cnpar, double molecular,
                                               const double* __restrict__ nuh, size_t p, int lane, float* e2,
                                               float* ru, float* g3, float* g4, double* wl, int nzmax)
{
    const size_t n2 = A.n2;
    const int nz = A.nz;
    const double* hn = A.hn + p;
    double ru_up = 0.0, an_k = 0.0, ao_k = 0.0, w_up = 0.0, H = 0.0;
    float ru_up_f = 0.f;
    double dz = hn[(size_t)(nz - 1) * n2];
    for (int k = nz - 1; k >= 0; k--) {
        double an_km1 = 0.0, ao_km1 = 0.0, dz_lo = 0.0;
        if (k >= 1) {
            dz_lo = hn[(size_t)(k - 1) * n2];
            const double x = (2.0 * dt) * (nuh[(size_t)k * n2 + p] + molecular);
            const double rs = rcp_nr(dz + dz_lo);
            an_km1 = (cnpar * x) * rs;
            ao_km1 = ((1.0 - cnpar) * x) * rs;
        }
        const double a2 = (dz + an_k) + an_km1;
        const double den = k == nz - 1 ? a2 : a2 + an_k * ru_up;  // a2 - a3 ru_{k+1}, a3 = -an_k
        const double rden = rcp_nr(den);
        const double ruk = -an_km1 * rden;
        const float ruk_f = float(ruk);
        e2[k * 64 + lane] = float(an_k * rden);
        ru[k * 64 + lane] = ruk_f;
        if (!CN1) {
            g3[k * 64 + lane] = float(ao_k * rden);
            g4[k * 64 + lane] = float(ao_km1 * rden);
        }
        const double wk = k == nz - 1 ? dz : fma(-(double)ru_up_f, w_up, dz);
        wl[k * 64 + lane] = wk;
        H += dz;
        w_up = wk;
        ru_up = ruk;
        ru_up_f = ruk_f;
        an_k = an_km1;
        ao_k = ao_km1;
        dz = dz_lo;
    }
    const double rH = 1.0 / H;
    for (int k = 0; k < nz; k++) wl[k * 64 + lane] *= rH;
    // levels nz .. nzmax - 1 of the depth class solve nothing: zero factors and weights
    for (int k = nz; k < nzmax; k++) {
        e2[k * 64 + lane] = 0.f;
        ru[k * 64 + lane] = 0.f;
        if (!CN1) {
            g3[k * 64 + lane] = 0.f;
            g4[k * 64 + lane] = 0.f;
        }
        wl[k * 64 + lane] = 0.0;
    }
}

// Compensated FP32 column solve (GETM_F32C): the factorisation and elimination
// of vdiff_body (FP32 factors in LDS, FP32 right side q), with the differences
// that drive the solve taken of the FP64-decoded old values.  NW waves share a
// block's 64 columns and their factorisation (made by wave 0), each wave
// solving its own tracers, one lane per column.  The exact increments conserve
// the column's content, sum_k h_k d_k = 0 (the right side is a sum of
// differences, and the matrix minus diag(h) has zero column sums); the FP32
// elimination leaves that sum off by ~1e-7 of |d|, so the defect
// D = sum_k h_k d_k is removed as the uniform shift
// D / sum_k h_k = sum_k (w_k / H) q_k (FP64, weights from vdiff_factor_w), and
// the column conserves to FP64 rounding.  Back-substitution runs in FP64 from
// the bottom and stores each level as it goes: c (+ s) + (d - D / H), encoded
// once.  The old column is held in registers as its FP32 values and its int16
// residuals (two to a register), q as FP32, so every tracer value is read and
// written once.  The solve always runs NZMAX levels: the levels above the
// surface (nz .. NZMAX - 1) have zero factors and weights (so their q and d are
// 0) and read the surface cell; only real levels are stored.  Memory is reached
// through buffer resources: a lane offset per element size and a wave-uniform
// layer offset (the launcher checks that a tracer's FP64 extent stays below
// 4 GiB).
template <int NZMAX, bool CN1, bool SRC, int NW>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(3)))
k_vdiff_c(const StepArgs A, double dt, double cnpar, double molecular, const double* __restrict__ nuh, int ntr_all,
          int tpb, const float* __restrict__ fin_all, const short* __restrict__ lin_all, float* __restrict__ fout_all,
          short* __restrict__ lout_all, const double* __restrict__ ea4_all)
{
    extern __shared__ float s_coef[];  // e2, ru[, g3, g4]: [NZMAX][64] FP32; then w / H: [NZMAX][64] FP64
    static_assert(NZMAX % 8 == 0, "depth classes are multiples of 8 layers");
    const int nz = A.nz;
    const size_t n3 = A.n3;
    float* e2 = s_coef;
    float* ru = s_coef + NZMAX * 64;
    float* g3 = s_coef + 2 * NZMAX * 64;
    float* g4 = s_coef + 3 * NZMAX * 64;
    double* wl = reinterpret_cast<double*>(s_coef + (CN1 ? 2 : 4) * NZMAX * 64);
    unsigned bx, by, bz;
    if (!xcd_block((A.nx_ + 63) / 64, A.ny_, (ntr_all + tpb - 1) / tpb, bx, by, bz)) return;
    const int lane = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int i = bx * 64 + lane, j = by;
    const bool inx = i < A.nx_;
    const unsigned p = (unsigned)j * (unsigned)A.nx_ + (unsigned)(inx ? i : 0);
    const bool inside = inx && i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax;
    const bool wetcol = inside && A.tmask[p] == 1;
    const bool solve = wetcol && nz > 1;
    if (wave == 0 && solve) vdiff_factor_w<CN1>(A, dt, cnpar, molecular, nuh, p, lane, e2, ru, g3, g4, wl, NZMAX);
    __syncthreads();
    // columns that are not solved keep their values in fout (no stage writes them),
    // except that with nz = 1 every point takes the reference's (dz c)/dz
    if (!inx || (!solve && nz > 1)) return;
    const int tb0 = bz * tpb;
    const int ntr = ntr_all - tb0 < tpb ? ntr_all - tb0 : tpb;
    // lane byte offsets of the column's cell in an FP32, int16 and FP64 layer;
    // the layer's own offset is wave-uniform (k n2 times the element size)
    const unsigned v4 = p * 4u, v2 = p * 2u, v8 = p * 8u;
    const Rsrc rh = make_rsrc(A.hn, (unsigned)(n3 * 8));
    for (int t = wave; t < ntr; t += NW) {
        // the layer size and the lane through an empty asm: the per-level offsets
        // and LDS factor loads below are formed where used instead of being
        // hoisted out of the tracer loop and held in registers across it
        unsigned n2 = (unsigned)A.n2;
        asm volatile("" : "+s"(n2));
        int ln = lane;
        asm volatile("" : "+v"(ln));
        // layer of level k (the levels above the surface read the surface cell)
        auto lev = [&](int k) -> unsigned { return (unsigned)(k < nz ? k : nz - 1) * n2; };
        const size_t o = (size_t)(tb0 + t) * n3;  // wave-uniform
        const Rsrc rfi = make_rsrc(fin_all + o, (unsigned)(n3 * 4)), rli = make_rsrc(lin_all + o, (unsigned)(n3 * 2));
        const Rsrc rfo = make_rsrc(fout_all + o, (unsigned)(n3 * 4)), rlo = make_rsrc(lout_all + o, (unsigned)(n3 * 2));
        const Rsrc rei = make_rsrc(SRC ? ea4_all + o : A.hn, (unsigned)(n3 * 8));
        auto store = [&](unsigned layer, double x) {
            float h;
            short l;
            cencode(x, h, l);
            st_f32(rfo, v4, layer * 4u, h);
            st_i16(rlo, v2, layer * 2u, l);
        };
        auto src = [&](unsigned layer) -> double {  // ea4 / h of the cell (sources)
            return ld_f64(rei, v8, layer * 8u) / ld_f64(rh, v8, layer * 8u);
        };
        if (nz == 1) {
            // diffusion.F90:224-231: every point, halos and land included; the
            // value is in fin if the stages updated the column, else in fout
            const double dz = ld_f64(rh, v8, 0);
            const double c = wetcol ? cdecode(ld_f32(rfi, v4, 0), ld_i16(rli, v2, 0))
                                    : cdecode(ld_f32(rfo, v4, 0), ld_i16(rlo, v2, 0));
            if constexpr (SRC) store(0, (dz * c + ld_f64(rei, v8, 0)) / dz);
            else store(0, (dz * c) / dz);
            continue;
        }
        float hi[NZMAX];
        unsigned lp[NZMAX / 2];  // residuals of levels 2m (low half) and 2m + 1 (high half)
#pragma unroll
        for (int k = 0; k < NZMAX; k++) {
            hi[k] = ld_f32(rfi, v4, lev(k) * 4u);
            const unsigned l = (unsigned short)ld_i16(rli, v2, lev(k) * 2u);
            lp[k / 2] = k % 2 ? lp[k / 2] | (l << 16) : l;
        }
        // old value (plus ea4/h with sources) of level k
        auto value = [&](int k) -> double {
            const double c = cdecode(hi[k], (short)(k % 2 ? lp[k / 2] >> 16 : lp[k / 2] & 0xffffu));
            if constexpr (SRC) return c + src(lev(k));
            else return c;
        };
        // old value without the sources, for the explicit part of cnpar < 1
        auto raw = [&](int k) -> double {
            return cdecode(hi[k], (short)(k % 2 ? lp[k / 2] >> 16 : lp[k / 2] & 0xffffu));
        };
        constexpr bool RAW = !CN1 && SRC;
        float q[NZMAX];
        double shift = 0.0;
        {
            // elimination from the top down; window (c_{k+1}, c_k, c_{k-1}), and the
            // same window of the old values without ea4/h when the explicit part
            // needs them (diffusion.F90:238-262: a4 = c dzo + ea4 + auxo differences of c)
            double c_up = 0.0, c_k = value(NZMAX - 1);
            double r_up = 0.0, r_k = RAW ? raw(NZMAX - 1) : 0.0;
            float q_up = 0.f;
#pragma unroll
            for (int k = NZMAX - 1; k >= 0; k--) {
                const float ek = e2[k * 64 + ln];  // 0 at and above the surface
                const float rk = ru[k * 64 + ln];  // 0 at the bottom and above the surface
                const double c_dn = k > 0 ? value(k > 0 ? k - 1 : 0) : 0.0;
                const float dup = k < NZMAX - 1 ? float(c_up - c_k) : 0.f;
                const float ddn = k > 0 ? float(c_dn - c_k) : 0.f;
                float qk = __builtin_fmaf(ek, dup + q_up, -(rk * ddn));
                if constexpr (RAW) {
                    const double r_dn = k > 0 ? raw(k > 0 ? k - 1 : 0) : 0.0;
                    const float rup = k < NZMAX - 1 ? float(r_up - r_k) : 0.f;
                    const float rdn = k > 0 ? float(r_dn - r_k) : 0.f;
                    qk += g3[k * 64 + ln] * rup + g4[k * 64 + ln] * rdn;
                    r_up = r_k;
                    r_k = r_dn;
                } else if (!CN1) {
                    qk += g3[k * 64 + ln] * dup + g4[k * 64 + ln] * ddn;
                }
                q[k] = qk;
                shift = fma((double)qk, wl[k * 64 + ln], shift);
                q_up = qk;
                c_up = c_k;
                c_k = c_dn;
            }
        }
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < NZMAX; k++) {
            d = k == 0 ? (double)q[0] : fma(-(double)ru[k * 64 + ln], d, (double)q[k]);
            if (k < NZMAX - 8 || k < nz) {
                // the registers pass through an empty asm, so the compiler decodes them
                // again here instead of keeping the elimination's FP64 values live
                float h = hi[k];
                unsigned l = lp[k / 2];
                asm volatile("" : "+v"(h), "+v"(l));
                double c = cdecode(h, (short)(k % 2 ? l >> 16 : l & 0xffffu));
                if constexpr (SRC) c += src(lev(k));
                store(lev(k), c + (d - shift));
            }
        }
    }
}

// Ensemble: work item {domain, bx, row, batch}; scratch -> tracers.
template <int NZMAX, bool CN1>
__global__ void __launch_bounds__(64) k_vdiff_multi(const EnsDom<float>* __restrict__ doms,
                                                    const int4* __restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const EnsDom<float>& D = doms[wk.x];
    const int tb0 = wk.w * D.tpb_d;
    const int ntr = D.ntr - tb0 < D.tpb_d ? D.ntr - tb0 : D.tpb_d;
    const size_t off = (size_t)tb0 * D.A.n3;
    vdiff_body<NZMAX, CN1>(D.A, D.dt, D.cnpar, D.molecular, D.nuh, wk.y * 64 + threadIdx.x, wk.z, threadIdx.x, ntr,
                           D.scratch + off, D.tracers + off, tb0);
}

// ---------------------------------------------------------------------------
// FP32 column solve by parallel cyclic reduction, for shallow columns (nz <= 32):
// a block of 1024 lanes owns COLS = 1024 / NP consecutive columns (memory order),
// one lane per (column, level) with NP = the power of two >= nz (levels nz ..
// NP-1 are identity rows), so a wave holds 64 / NP whole columns.  The system is
// the increment form of k_vdiff_f32 (the reference's matrix,
// diffusion.F90:233-313, acting on d = c_new - c_old):
//   row k:  l_k d_{k-1} + m_k d_k + u_k d_{k+1} = r_k,
//   l_k = -an_{k-1},  m_k = h_k + an_k + an_{k-1},  u_k = -an_k,
//   r_k = an_k (c_{k+1} - c_k) + an_{k-1} (c_{k-1} - c_k) [+ ao_k (..) + ao_{k-1} (..)].
// Step s (stride 2^s) eliminates rows k +- 2^s from row k.  The factors a_s, g_s
// of the log2 NP steps depend only on the matrix: each lane forms them once for
// its row and keeps them in registers; every tracer then only reduces its
// right-hand side, r <- r + a_s r_{k-2^s} + g_s r_{k+2^s}, d = r / m.  Tracer
// tiles move between HBM and LDS by rows (coalesced) and are read by columns.
// More arithmetic per cell than the per-lane Thomas elimination, but a
// dependency chain of log2 NP instead of 2 nz and 16-32x more lanes per column:
// it pays when there are too few columns to fill the GPU.
// ---------------------------------------------------------------------------
template <int NP, bool CN1>
__device__ __forceinline__ void pcr_body(const StepArgs& A, double dt, double cnpar, double molecular,
                                         const double* __restrict__ nuh, size_t p0, int ntr,
                                         const float* __restrict__ fin, float* __restrict__ fout, int tg0)
{
    constexpr int COLS = 1024 / NP, LDW = COLS + 1;  // padded rows: column reads are conflict-free
    constexpr int LOG = NP == 8 ? 3 : NP == 16 ? 4 : 5;
    static_assert(NP == 8 || NP == 16 || NP == 32, "wave PCR covers columns of up to 32 levels");
    __shared__ float s_t[NP * LDW];
    __shared__ unsigned char s_solve[COLS];  // columns that are solved (the others keep their values in fout)
    const int tid = threadIdx.x, lane = tid % 64;
    const int k = tid % NP, cl = tid / NP;  // this lane's level and column in the tile
    const int nz = A.nz;
    const size_t n2 = A.n2, n3 = A.n3;
    const size_t pc = p0 + cl;
    const bool incol = pc < n2, lev = k < nz;
    bool solve = false;
    if (incol) {
        const int i = (int)(pc % A.nx_), j = (int)(pc / A.nx_);
        solve = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax && A.tmask[pc] == 1;
    }
    if (k == 0) s_solve[cl] = solve;  // visible after the first barrier of the tracer loop
    // row k in FP64, rounded once
    float l = 0.f, m = 1.f, u = 0.f, eu = 0.f, el = 0.f, ou = 0.f, ol = 0.f;
    if (solve && lev) {
        const double hk = A.hn[(size_t)k * n2 + pc];
        double an_up = 0.0, an_lo = 0.0, ao_up = 0.0, ao_lo = 0.0;
        if (k <= nz - 2) {
            const double x = (2.0 * dt) * (nuh[(size_t)(k + 1) * n2 + pc] + molecular);
            const double rs = rcp_nr(A.hn[(size_t)(k + 1) * n2 + pc] + hk);
            an_up = (cnpar * x) * rs;
            ao_up = ((1.0 - cnpar) * x) * rs;
        }
        if (k >= 1) {
            const double x = (2.0 * dt) * (nuh[(size_t)k * n2 + pc] + molecular);
            const double rs = rcp_nr(hk + A.hn[(size_t)(k - 1) * n2 + pc]);
            an_lo = (cnpar * x) * rs;
            ao_lo = ((1.0 - cnpar) * x) * rs;
        }
        l = float(-an_lo);
        u = float(-an_up);
        m = float((hk + an_up) + an_lo);
        eu = float(an_up);
        el = float(an_lo);
        ou = float(ao_up);
        ol = float(ao_lo);
    }
    // reduction factors of the matrix, kept for every tracer
    float fa[LOG], fg[LOG];
#pragma unroll
    for (int st = 0; st < LOG; st++) {
        const int sg = 1 << st;
        const bool hl = k >= sg, hu = k + sg < NP;  // neighbours stay in this column's lanes
        const int src_l = lane - (hl ? sg : 0), src_u = lane + (hu ? sg : 0);
        const float ml = __shfl(m, src_l), ll = __shfl(l, src_l), ul = __shfl(u, src_l);
        const float mu = __shfl(m, src_u), lu = __shfl(l, src_u), uu = __shfl(u, src_u);
        const float a = hl ? -l * __builtin_amdgcn_rcpf(ml) : 0.f;
        const float g = hu ? -u * __builtin_amdgcn_rcpf(mu) : 0.f;
        m = m + (hl ? a * ul : 0.f) + (hu ? g * lu : 0.f);
        l = hl ? a * ll : 0.f;
        u = hu ? g * uu : 0.f;
        fa[st] = a;
        fg[st] = g;
    }
    const float minv = 1.f / m;
    const unsigned ss = stage_seed<float>(A, SR_VDIFF);
    const unsigned rb = sr_cell_bits(ss, (unsigned)((size_t)k * n2 + pc));
    const int ncols = (size_t)COLS < n2 - p0 ? COLS : (int)(n2 - p0);
    for (int t = 0; t < ntr; t++) {
        const float* fi = fin + (size_t)t * n3 + p0;
        float* fo = fout + (size_t)t * n3 + p0;
        // rows in, coalesced
        for (int e = tid; e < nz * COLS; e += 1024) {
            const int kk = e / COLS, cc = e % COLS;
            if (cc < ncols) s_t[kk * LDW + cc] = fi[(size_t)kk * n2 + cc];
        }
        __syncthreads();
        const float c = (incol && lev) ? s_t[k * LDW + cl] : 0.f;
        const float cdn = __shfl(c, lane - (k >= 1 ? 1 : 0));
        const float cup = __shfl(c, lane + (k + 1 < NP ? 1 : 0));
        float r = 0.f;
        if (solve && lev) {
            const float ddn = k >= 1 ? cdn - c : 0.f;
            const float dup = k + 1 < nz ? cup - c : 0.f;
            r = el * ddn + eu * dup;
            if (!CN1) r += ol * ddn + ou * dup;
        }
#pragma unroll
        for (int st = 0; st < LOG; st++) {
            const int sg = 1 << st;
            const bool hl = k >= sg, hu = k + sg < NP;
            const float rl = __shfl(r, lane - (hl ? sg : 0));
            const float rh = __shfl(r, lane + (hu ? sg : 0));
            r = r + (hl ? fa[st] * rl : 0.f) + (hu ? fg[st] * rh : 0.f);
        }
        __syncthreads();  // every lane has read its value before the tile is overwritten
        if (incol && lev) s_t[k * LDW + cl] = tracer_round(ss, tg0 + t).add(c, r * minv, rb);
        __syncthreads();
        // rows out, coalesced (solved columns only)
        for (int e = tid; e < nz * COLS; e += 1024) {
            const int kk = e / COLS, cc = e % COLS;
            if (cc < ncols && s_solve[cc]) fo[(size_t)kk * n2 + cc] = s_t[kk * LDW + cc];
        }
        __syncthreads();
    }
}

// Single domain: blockIdx.x = tile of 1024 / NP columns, blockIdx.y = tracer batch (2 <= nz <= 32).
template <int NP, bool CN1>
__global__ void __launch_bounds__(1024) k_vdiff_pcr(const StepArgs A, double dt, double cnpar, double molecular,
                                                    const double* __restrict__ nuh, int ntr_all, int tpb,
                                                    const float* __restrict__ fin_all, float* __restrict__ fout_all)
{
    const int tb0 = blockIdx.y * tpb;
    const int ntr = ntr_all - tb0 < tpb ? ntr_all - tb0 : tpb;
    pcr_body<NP, CN1>(A, dt, cnpar, molecular, nuh, (size_t)blockIdx.x * (1024 / NP), ntr,
                      fin_all + (size_t)tb0 * A.n3, fout_all + (size_t)tb0 * A.n3, tb0);
}

// Ensemble: one work item per block {domain, first column, 0, batch}.
template <int NP, bool CN1>
__global__ void __launch_bounds__(1024) k_vdiff_pcr_multi(const EnsDom<float>* __restrict__ doms,
                                                          const int4* __restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const EnsDom<float>& D = doms[wk.x];
    const int tb0 = wk.w * D.tpb_d;
    const int ntr = D.ntr - tb0 < D.tpb_d ? D.ntr - tb0 : D.tpb_d;
    const size_t off = (size_t)tb0 * D.A.n3;
    pcr_body<NP, CN1>(D.A, D.dt, D.cnpar, D.molecular, D.nuh, (size_t)wk.y, ntr, D.scratch + off, D.tracers + off,
                      tb0);
}

static inline int nbatch(int ntr, int tpb) { return tpb <= 0 ? 1 : (ntr + tpb - 1) / tpb; }

template <typename R, bool C>
inline int launch_hpair_impl(int d1, int scheme, int ty, const StepArgs& A, double dt, int ntr, int tpb,
                             const std::conditional_t<C, float, R>* fin, std::conditional_t<C, float, R>* fout,
                             const short* lin, short* lout, const double* hin, double* hout, unsigned sr_stage,
                             hipStream_t s, const double* hin_dry, bool write_dry)
{
    if (!hin_dry) hin_dry = hin;
    if (tpb <= 0 || tpb > ntr) tpb = ntr;
    const int nb = nbatch(ntr, tpb);
    if (ty != 4 && ty != 8 && ty != 16 && !(C && ty == 12)) return GETM_ERR_ARG;
    // Tile shape.  Stages 4-5 sweep x first, so their first stage covers the
    // tile's rows plus two on each side: 32-wide tiles of twice the height
    // (same lane count) cut that overhead from (ty + 4)/ty to (2 ty + 4)/(2 ty).
    // That pays where the per-layer set-up dominates, i.e. for few tracers per
    // block (coastal, 2 tracers: 0.768 -> 0.732 ms); with many, the narrower
    // rows cost more in memory access than they save (256 tracers: 35.5 -> 40.3
    // ms).  Stages 1-2 sweep y first and keep 64-wide rows; so does the FP64 form
    // (measured for FP32 only).
    const bool narrow = std::is_same<R, float>::value && d1 == 1 && ty <= 8 && tpb <= HPAIR_KG;
    const int txw = narrow ? 32 : TX;
    const int tye = narrow ? 2 * ty : ty;
    // FP32 and compensated forms: tiles over the interior [imin, imax] x [jmin,
    // jmax] only (they store updated cells only, and the thickness of the halo
    // ring, which no stage changes, is copied by launch_h_ring); 64-wide tiles
    // over the 1028 columns of a 1024-column domain with its halos would add a
    // 17th tile for 2 interior columns.  FP64 covers every cell: its step copies
    // the whole scratch buffer back.
    const bool inner = C || std::is_same<R, float>::value;
    const int ox = inner ? A.imin : 0, oy = inner ? A.jmin : 0;
    const int wx = inner ? A.imax - A.imin + 1 : A.nx_, wy = inner ? A.jmax - A.jmin + 1 : A.ny_;
    if (wx <= 0 || wy <= 0) return GETM_OK;  // no interior: no cell to update
    const int ntx = (wx + txw - 1) / txw, nty = (wy + tye - 1) / tye;
    const bool ah = A.ah_u != nullptr;
    const int kg = ah ? 1 : hpair_group(tpb, std::is_same<R, float>::value);
    const long nitems = (long)nb * ((A.nz + kg - 1) / kg) * ntx * nty;
    if (nitems > 0x7fffff00L) return GETM_ERR_CAPACITY;
    dim3 grid((unsigned)(8 * ((nitems + 7) / 8)));
    ProfileScope ps(d1 == 2 ? GETM_K_HPAIR_VU : GETM_K_HPAIR_UV, s);
#define GETM_HPK(D1V, TYV, TXV, KGV, AHV, LPCV)                                                                     \
    GETM_SCHEME_SWITCH(scheme, {                                                                                      \
        hipLaunchKernelGGL((k_hpair<S_, R, D1V, TYV, KGV, AHV, TXV, C, LPCV>), grid, dim3(TXV * TYV / LPCV), 0, s, A,  \
                           dt, ntr, tpb, ntx, nty, nb, ox, oy, fin, fout, hin, hout, sr_stage, hin_dry,               \
                           (int)write_dry, lin, lout);                                                                \
    })
#define GETM_HP(D1V, TYV, TXV)                            \
    if (ah) {                                             \
        GETM_HPK(D1V, TYV, TXV, 1, true, 1);              \
    } else if constexpr (std::is_same<R, float>::value) {  \
        if (kg == HPAIR_KG) {                             \
            GETM_HPK(D1V, TYV, TXV, HPAIR_KG, false, 1);  \
        } else {                                          \
            GETM_HPK(D1V, TYV, TXV, 1, false, 1);         \
        }                                                 \
    } else {                                              \
        GETM_HPK(D1V, TYV, TXV, 1, false, 1);             \
    }
    // FP32 and compensated forms, 16-row tiles of one layer: two cells per lane
    // (512-lane blocks, two per CU, each covering the other's barrier phases).  At
    // 256 tracers: compensated x-first pair 43.3 -> 40.9 ms, y-first 42.1 -> 38.1
    // ms; FP32 pairs 38.7 / 39.3 -> 28.4 / 29.1 ms (34.2 / 34.4 ms on 8-row tiles)
    {
        if (!ah && ty == 16 && kg == 1) {
            if (d1 == 2) {
                GETM_HPK(2, 16, TX, 1, false, 2);
            } else {
                GETM_HPK(1, 16, TX, 1, false, 2);
            }
            GETM_HIP_OK(hipGetLastError());
            return GETM_OK;
        }
    }
    if (d1 == 2) {
        if (ty == 16) {
            GETM_HP(2, 16, TX);
        } else if (C && ty == 12) {
            GETM_HP(2, 12, TX);
        } else if (ty == 8) {
            GETM_HP(2, 8, TX);
        } else {
            GETM_HP(2, 4, TX);
        }
    } else if (narrow) {
        if (ty == 8) {
            GETM_HP(1, 16, 32);
        } else {
            GETM_HP(1, 8, 32);
        }
    } else {
        if (ty == 16) {
            GETM_HP(1, 16, TX);
        } else if (C && ty == 12) {
            GETM_HP(1, 12, TX);
        } else if (ty == 8) {
            GETM_HP(1, 8, TX);
        } else {
            GETM_HP(1, 4, TX);
        }
    }
#undef GETM_HP
#undef GETM_HPK
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

}  // namespace getm
