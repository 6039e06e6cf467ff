// Tracer inventories and the FP32 inventory fixer.
//
// Inventory of tracer t: sum over wet cells of area * h * c
// (the reference's domain integral, python/pygetm/simulation.py:1408-1466),
// accumulated in FP64 with Neumaier compensation in a fixed order (lane
// column sums, then a butterfly over each wave and the waves of a block in
// order, then a fixed-order sum over blocks), so it is deterministic.
//
// Fixer (FP32 tracers): FP32 storage rounds every cell once per stage, which
// makes the inventory drift by ~1e-11 relative per step.  The fixer restores
// a target inventory M exactly up to one cell's rounding: the defect
// D = M - I is shared in proportion to each cell's absolute mass, which in
// units of the cell's ulp is q_i = r |c_i| / ulp(c_i) with r = D / I_abs and
// I_abs = sum of area * h * |c| (I_abs = I for a tracer of one sign; with mixed
// signs no cell moves by more than |D| / I_abs of its value, however small I);
// with S_i the prefix sum of |c|/ulp in a fixed cell order, cell i changes by
//   n_i = floor(r S_i + u) - floor(r S_{i-1} + u)  ulps   (n_i in {-1, 0, 1} in practice),
// which is exact in FP32 and touches ~|r| * 2^23 cells per tracer.  The phase
// u in [0, 1) is a hash of the bits of r, so a defect that repeats from step
// to step does not flip the same cells every time (expected n_i = q_i).
// Flipped cells carry different masses, so a residual of a few tens of
// cell-ulps remains (of the order of the largest cell's ulp mass when a tracer
// sits in a few cells, e.g. a fresh point release).  The fix kernel
// accumulates the fixed inventory and each block's largest-mass cell on the
// way; a greedy pass (one wave per tracer, cells decided in a fixed order)
// then changes cells near the tracer's largest-mass cell: first by at most one
// ulp each, over the 64-column chunk of that cell in an order rotated by a hash
// of the residual (so no cell takes the residual step after step), then, if a
// residual remains, by rint(R / ulp mass) ulps (at most 4096, i.e. 5e-4 of the
// cell's own value) along a wider walk.  Either way the residual R never grows
// and ends below half the ulp mass of the smallest cell passed.
#include <type_traits>

#include "common.h"
#include "internal.h"

namespace getm {

constexpr int IBX = 64, IBY = 4, INV_THREADS = IBX * IBY;
constexpr int INV_TCH = 8;  // tracers per block of k_inventory_tr
constexpr int INV_SLOTS = 4;  // partial sums per (tracer, block): S, C, sum |c|/ulp, absolute mass

__device__ __forceinline__ void neumaier(double& s, double& c, double x)
{
    const double t = s + x;
    if (fabs(s) >= fabs(x))
        c += (s - t) + x;
    else
        c += (x - t) + s;
    s = t;
}

// Sum of two compensated pairs (s1, c1) + (s2, c2).  Symmetric in its two
// arguments (the error term comes from the larger of s1, s2 either way), so a
// butterfly over a wave leaves the same pair in every lane.
__device__ __forceinline__ void neumaier_merge(double& s, double& c, double s2, double c2)
{
    const double t = s + s2;
    const double e = fabs(s) >= fabs(s2) ? (s - t) + s2 : (s2 - t) + s;
    c = (c + c2) + e;
    s = t;
}

// |c| / ulp(c) for an FP32 value (0 for zero and subnormals): |c| 2^(23 - e)
// with e = eb - 127 from the biased exponent eb, the power of two built as an
// FP64 with exponent field 1023 + 150 - eb (exact, no exp2 per value).
// Infinities and NaN give NaN.
__device__ __forceinline__ double q_of(float c)
{
    const unsigned eb = (__float_as_uint(c) >> 23) & 0xffu;
    if (eb == 0u) return 0.0;
    if (eb == 255u) return (double)c * 0.0;
    return (double)fabsf(c) * __hiloint2double((int)((1173u - eb) << 20), 0);
}

// Partial sums per (tracer, block), INV_SLOTS of them: mass (hi, lo), the sum
// of |c|/ulp(c) (FP32 only, 0 otherwise) and the absolute mass (area h |c|).
// COMP: compensated FP32 tracers (f the FP32 values, lo their int16 residuals).
template <typename R, bool COMP = false>
__global__ void __launch_bounds__(INV_THREADS) k_inventory(const StepArgs A, const double* __restrict__ h,
                                                           const double* __restrict__ area, int ntr,
                                                           const R* __restrict__ f, double* __restrict__ part,
                                                           const short* __restrict__ lo = nullptr)
{
    __shared__ double s_m[IBY], s_c[IBY], s_q[IBY], s_a[IBY];
    const int tid = threadIdx.y * IBX + threadIdx.x;
    const int i = blockIdx.x * IBX + threadIdx.x, j = blockIdx.y * IBY + threadIdx.y;
    const int t = blockIdx.z;
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    double m = 0.0, mc = 0.0, q = 0.0, ab = 0.0;
    if (i < A.nx_ && j < A.ny_) {
        const size_t p = (size_t)j * A.nx_ + i;
        const bool wet = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax && A.tmask[p] == 1;
        if (wet) {
            const R* fc = f + (size_t)t * A.n3 + p;
            const double a = area[p];
            for (int k = 0; k < A.nz; k++) {
                const R c = fc[(size_t)k * A.n2];
                const double ah = a * h[(size_t)k * A.n2 + p];
                if constexpr (COMP) {
                    const double x = cdecode(c, lo[(size_t)t * A.n3 + p + (size_t)k * A.n2]);
                    neumaier(m, mc, ah * x);
                    ab = fma(ah, fabs(x), ab);
                } else {
                    neumaier(m, mc, ah * (double)c);
                    ab = fma(ah, fabs((double)c), ab);
                    if constexpr (std::is_same<R, float>::value) q += q_of(c);
                }
            }
        }
    }
    // a butterfly over each wave (compensated for the mass), then the waves in
    // order: a fixed order, so deterministic
    for (int d = 1; d < 64; d <<= 1) {
        neumaier_merge(m, mc, __shfl_xor(m, d), __shfl_xor(mc, d));
        q += __shfl_xor(q, d);
        ab += __shfl_xor(ab, d);
    }
    if (threadIdx.x == 0) {
        s_m[threadIdx.y] = m;
        s_c[threadIdx.y] = mc;
        s_q[threadIdx.y] = q;
        s_a[threadIdx.y] = ab;
    }
    __syncthreads();
    if (tid == 0) {
        double S = s_m[0], C = s_c[0], Q = s_q[0], AB = s_a[0];
        for (int w = 1; w < IBY; w++) {
            neumaier_merge(S, C, s_m[w], s_c[w]);
            Q += s_q[w];
            AB += s_a[w];
        }
        double* pt = part + (size_t)t * INV_SLOTS * nblk;
        pt[blk] = S;
        pt[nblk + blk] = C;
        pt[2 * nblk + blk] = Q;
        pt[3 * nblk + blk] = AB;
    }
}

// Partial sums per (tracer, block) of the FP64 and compensated forms, TCH
// tracers per block: each lane sums its column over k in order for every
// tracer of the chunk, with the column's area * h loaded once for all of them;
// the block combines its lanes by a butterfly over each wave and the four
// waves in order (a fixed order: deterministic), and writes the same
// partial sums per block as k_inventory (the c/ulp sum 0; the absolute mass
// for the compensated form only).  1-D XCD-aware grid over
// (column block x, column block y, tracer chunk), so that neighbouring column
// blocks share their L2.
template <typename R, bool COMP, int TCH>
__global__ void __launch_bounds__(INV_THREADS) k_inventory_tr(const StepArgs A, const double* __restrict__ h,
                                                              const double* __restrict__ area, int ntr,
                                                              const R* __restrict__ f,
                                                              const short* __restrict__ lo,
                                                              double* __restrict__ part)
{
    __shared__ double s_s[IBY][TCH], s_c[IBY][TCH], s_a[IBY][TCH];
    const unsigned nbx = (A.nx_ + IBX - 1) / IBX, nby = (A.ny_ + IBY - 1) / IBY;
    unsigned bx, by, bz;
    if (!xcd_block(nbx, nby, (ntr + TCH - 1) / TCH, bx, by, bz)) return;
    const int i = bx * IBX + threadIdx.x, j = by * IBY + threadIdx.y;
    const int t0 = bz * TCH;
    const int nt = ntr - t0 < TCH ? ntr - t0 : TCH;
    const int nblk = nbx * nby;
    const int blk = by * nbx + bx;
    // per lane: the column's sum for each tracer of the chunk (40-odd terms: plain
    // FP64 sums, exact to ~1e-15 of the column's sum of magnitudes); compensation
    // starts with the merge of the lanes' sums
    double m[TCH], mc[TCH], ma[COMP ? TCH : 1];
#pragma unroll
    for (int tt = 0; tt < TCH; tt++) m[tt] = mc[tt] = 0.0;
#pragma unroll
    for (int tt = 0; tt < (COMP ? TCH : 1); tt++) ma[tt] = 0.0;
    if (i < A.nx_ && j < A.ny_) {
        const unsigned p = (unsigned)j * A.nx_ + i;
        const bool wet = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax && A.tmask[p] == 1;
        if (wet) {
            const double a = area[p];
            const unsigned n2 = (unsigned)A.n2;
            const unsigned vs = p * (unsigned)sizeof(R), v2 = p * 2u, v8 = p * 8u;
            const Rsrc rh = make_rsrc(h, (unsigned)(A.n3 * 8));
            Rsrc rf[TCH], rl[TCH];
#pragma unroll
            for (int tt = 0; tt < TCH; tt++) {
                const size_t o = (size_t)(t0 + (tt < nt ? tt : 0)) * A.n3;
                rf[tt] = make_rsrc(f + o, (unsigned)(A.n3 * sizeof(R)));
                rl[tt] = make_rsrc(COMP ? (const void*)(lo + o) : (const void*)(f + o), (unsigned)(A.n3 * 2));
            }
#pragma unroll 4
            for (int k = 0; k < A.nz; k++) {
                const unsigned L = (unsigned)k * n2;
                const double ah = a * ld_f64(rh, v8, L * 8u);
                // (tracers past the chunk's end read tracer t0 again; their sums are not written)
#pragma unroll
                for (int tt = 0; tt < TCH; tt++) {
                    double x;
                    if constexpr (COMP)
                        x = cdecode(ld_f32(rf[tt], vs, L * 4u), ld_i16(rl[tt], v2, L * 2u));
                    else if constexpr (sizeof(R) == 8)
                        x = ld_f64(rf[tt], vs, L * 8u);
                    else
                        x = (double)ld_f32(rf[tt], vs, L * 4u);
                    m[tt] = fma(ah, x, m[tt]);
                    if constexpr (COMP) ma[tt] = fma(ah, fabs(x), ma[tt]);
                }
            }
        }
    }
#pragma unroll
    for (int tt = 0; tt < TCH; tt++) {
        for (int d = 1; d < 64; d <<= 1) neumaier_merge(m[tt], mc[tt], __shfl_xor(m[tt], d), __shfl_xor(mc[tt], d));
        double a = 0.0;
        if constexpr (COMP) {
            a = ma[tt];
            for (int d = 1; d < 64; d <<= 1) a += __shfl_xor(a, d);
        }
        if (threadIdx.x == 0) {
            s_s[threadIdx.y][tt] = m[tt];
            s_c[threadIdx.y][tt] = mc[tt];
            s_a[threadIdx.y][tt] = a;
        }
    }
    __syncthreads();
    if (threadIdx.y == 0 && threadIdx.x < nt) {
        const int tt = threadIdx.x;
        double S = s_s[0][tt], C = s_c[0][tt], AB = s_a[0][tt];
        for (int w = 1; w < IBY; w++) {
            neumaier_merge(S, C, s_s[w][tt], s_c[w][tt]);
            AB += s_a[w][tt];
        }
        double* pt = part + (size_t)(t0 + tt) * INV_SLOTS * nblk;
        pt[blk] = S;
        pt[nblk + blk] = C;
        pt[2 * nblk + blk] = 0.0;
        pt[3 * nblk + blk] = AB;
    }
}

// Fixed-order combination of per-lane compensated sums (s, c) over a wave:
// lane 0's pair, then lane 1's, ... (every lane ends with the same total).
__device__ __forceinline__ double wave_neumaier_total(double s, double c)
{
    double S = 0.0, C = 0.0;
    for (int l = 0; l < 64; l++) {
        neumaier(S, C, __shfl(s, l));
        C += __shfl(c, l);
    }
    return S + C;
}

// Exclusive prefix over the wave's lanes, in lane order (same value order every call).
__device__ __forceinline__ double wave_exclusive_prefix(double x, int lane)
{
    double run = 0.0, mine = 0.0;
    for (int l = 0; l < 64; l++) {
        if (l == lane) mine = run;
        run += __shfl(x, l);
    }
    return mine;
}

// One wave per tracer: inventory = compensated sum over blocks; for the fixer
// also the exclusive prefix of the per-block |c|/ulp sums and the ratio
// D / I_abs of the defect to the absolute mass.
// Lane l owns the contiguous block range [l seg, (l + 1) seg), so the prefix is
// a lane-local running sum started at the lane's exclusive offset.
__global__ void __launch_bounds__(64) k_inventory_reduce(int ntr, int nblk, double* __restrict__ part,
                                                         double* __restrict__ inv, const double* __restrict__ target,
                                                         double* __restrict__ ratio)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    if (t >= ntr) return;
    double* pt = part + (size_t)t * INV_SLOTS * nblk;
    const int seg = (nblk + 63) / 64;
    const int b0 = lane * seg, b1 = b0 + seg < nblk ? b0 + seg : nblk;
    double S = 0.0, C = 0.0, Q = 0.0, AB = 0.0;
    for (int b = b0; b < b1; b++) {
        neumaier(S, C, pt[b]);
        C += pt[nblk + b];
        Q += pt[2 * nblk + b];
        AB += pt[3 * nblk + b];
    }
    double abt = 0.0;  // absolute mass, lanes in order
    for (int l = 0; l < 64; l++) abt += __shfl(AB, l);
    double run = wave_exclusive_prefix(Q, lane);
    for (int b = b0; b < b1; b++) {
        const double qb = pt[2 * nblk + b];
        pt[2 * nblk + b] = run;  // exclusive prefix, in place
        run += qb;
    }
    const double I = wave_neumaier_total(S, C);
    if (lane == 0) {
        if (inv) inv[t] = I;
        // the fixers move each cell in proportion to its absolute mass: ratio
        // D / I_abs (the FP64 form has no fixer and no absolute mass: D / I)
        const double den = abt > 0.0 ? abt : I;
        if (target && ratio) ratio[t] = (den != 0.0) ? (target[t] - I) / den : 0.0;
    }
}

// Same block decomposition and summation order as k_inventory, so the block
// prefixes from k_inventory_reduce continue exactly into each block.
__global__ void __launch_bounds__(INV_THREADS) k_fix_f32(const StepArgs A, int ntr, float* __restrict__ f,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ ratio,
                                                         const double* __restrict__ area, double* __restrict__ part2)
{
    __shared__ double s_w[IBY], s_m[IBY], s_c[IBY], s_x[IBY];
    __shared__ int s_l[IBY];
    const int tid = threadIdx.y * IBX + threadIdx.x;
    const int lane = threadIdx.x, wv = threadIdx.y;
    const int i = blockIdx.x * IBX + threadIdx.x, j = blockIdx.y * IBY + threadIdx.y;
    const int t = blockIdx.z;
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    const double r = ratio[t];
    const unsigned long long rb = (unsigned long long)__double_as_longlong(r);
    const double u = (double)(sr_hash((unsigned)rb) ^ sr_hash((unsigned)(rb >> 32) + (unsigned)t)) * 0x1p-32;
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
    // exclusive prefix of the lanes' sums in lane order, from the block's offset:
    // a scan within each wave, then the totals of the waves before
    double incl = q;
    for (int d = 1; d < 64; d <<= 1) {
        const double o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    double excl = __shfl_up(incl, 1);
    if (lane == 0) excl = 0.0;
    if (lane == 63) s_w[wv] = incl;
    __syncthreads();
    double base = part[(size_t)t * INV_SLOTS * nblk + 2 * nblk + blk];
    for (int w2 = 0; w2 < wv; w2++) base += s_w[w2];
    double m = 0.0, mc = 0.0, mx = 0.0;
    if (wet) {
        const double a = area[p];
        double S0 = base + excl;
        for (int k = 0; k < A.nz; k++) {
            float c = fc[(size_t)k * A.n2];
            const double S1 = S0 + q_of(c);
            const double n = floor(r * S1 + u) - floor(r * S0 + u);
            if (n != 0.0) {
                const int e = ilogbf(c);
                c = c + (float)(n * exp2((double)(e - 23)));
                fc[(size_t)k * A.n2] = c;
            }
            const double cm = (a * A.hn[(size_t)k * A.n2 + p]) * (double)c;
            neumaier(m, mc, cm);
            mx = fmax(mx, fabs(cm));
            S0 = S1;
        }
    }
    // the block's fixed inventory (compensated butterfly over each wave, then the
    // waves in order) and its largest cell mass with the lowest lane holding it
    for (int d = 1; d < 64; d <<= 1) neumaier_merge(m, mc, __shfl_xor(m, d), __shfl_xor(mc, d));
    int ml = mx > 0.0 ? tid : INV_THREADS;
    for (int d = 1; d < 64; d <<= 1) {
        const double ox = __shfl_xor(mx, d);
        const int ol = __shfl_xor(ml, d);
        if (ox > mx || (ox == mx && ol < ml)) {
            mx = ox;
            ml = ol;
        }
    }
    if (lane == 0) {
        s_m[wv] = m;
        s_c[wv] = mc;
        s_x[wv] = mx;
        s_l[wv] = ml;
    }
    __syncthreads();
    if (tid == 0) {
        double S = s_m[0], C = s_c[0], X = s_x[0];
        int L = s_l[0];
        for (int w2 = 1; w2 < IBY; w2++) {
            neumaier_merge(S, C, s_m[w2], s_c[w2]);
            if (s_x[w2] > X || (s_x[w2] == X && s_l[w2] < L)) {
                X = s_x[w2];
                L = s_l[w2];
            }
        }
        const int lx = X > 0.0 ? L : -1;
        double* pt = part2 + (size_t)t * 4 * nblk;
        pt[blk] = S;
        pt[nblk + blk] = C;
        pt[2 * nblk + blk] = X;
        // column of the block's largest-mass cell (-1: no mass in this block)
        pt[3 * nblk + blk] = lx < 0 ? -1.0
                                    : (double)((size_t)(blockIdx.y * IBY + lx / IBX) * A.nx_ +
                                               (blockIdx.x * IBX + lx % IBX));
    }
}

// One wave per tracer: remaining defect after the dithering pass (resid[t])
// and the column of the largest-mass cell (resid[ntr + t], -1 if none; ties go
// to the lowest block).
__global__ void __launch_bounds__(64) k_residual(int ntr, int nblk, const double* __restrict__ part2,
                                                 const double* __restrict__ target, double* __restrict__ resid)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    if (t >= ntr) return;
    const double* pt = part2 + (size_t)t * 4 * nblk;
    const int seg = (nblk + 63) / 64;
    const int b0 = lane * seg, b1 = b0 + seg < nblk ? b0 + seg : nblk;
    double S = 0.0, C = 0.0, X = 0.0, P = -1.0;
    for (int b = b0; b < b1; b++) {
        neumaier(S, C, pt[b]);
        C += pt[nblk + b];
        if (pt[2 * nblk + b] > X) {
            X = pt[2 * nblk + b];
            P = pt[3 * nblk + b];
        }
    }
    const double tot = wave_neumaier_total(S, C);
    double XB = 0.0, PB = -1.0;
    for (int l = 0; l < 64; l++) {
        const double xl = __shfl(X, l), pl = __shfl(P, l);
        if (xl > XB) {
            XB = xl;
            PB = pl;
        }
    }
    if (lane == 0) {
        resid[ntr + t] = PB;
        // target - tot, compensated
        double R = 0.0, RC = 0.0;
        neumaier(R, RC, target[t]);
        neumaier(R, RC, -tot);
        resid[t] = R + RC;
    }
}

// One wave per tracer.  Spread pass: the chunk of the largest-mass cell (row j0,
// 64 columns), one ulp per cell at most, layers and lanes in an order rotated by
// a hash of the residual.  Then greedy changes of up to GREEDY_MAX_ULPS ulps per
// cell if a residual remains.  The walk visits chunks of 64 consecutive columns and, in each, the layers
// bottom to top; the 64 cells of one layer are loaded together and then decided
// one lane after the other with the running residual R, which every lane holds.
// Chunks: first the 64 columns centred on the largest-mass cell in the rows
// j0, j0-1, j0+1, ..., j0 +- GREEDY_ROWS/2 (a tracer that sits in a few cells,
// such as a fresh point release, has all its cells of intermediate mass there,
// upstream and downstream), then the whole grid in memory order from that
// column on, wrapping round.  A cell whose ulp mass exceeds 2|R| would change by
// 0 ulps and is skipped.  The walk ends when |R| <= tol or after max_cells
// cells (wet or not).
constexpr double GREEDY_MAX_ULPS = 4096.0;
constexpr int GREEDY_ROWS = 63;

__global__ void __launch_bounds__(64) k_greedy_f32(const StepArgs A, int ntr, float* __restrict__ f,
                                                   const double* __restrict__ area, const double* __restrict__ resid,
                                                   const double* __restrict__ target, int max_cells)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    if (t >= ntr) return;
    double R = resid[t];
    const double tol = 1e-14 * fabs(target[t]);  // two orders below the conservation target
    if (fabs(R) <= tol) return;
    float* ft = f + (size_t)t * A.n3;
    const double P = resid[ntr + t];
    const size_t n2 = (size_t)A.n2;
    const size_t p0 = (P >= 0.0 && P < (double)n2) ? (size_t)P : 0;
    const int i0 = (int)(p0 % A.nx_), j0 = (int)(p0 / A.nx_);
    // ---- spread: one ulp per cell over the chunk of the largest-mass cell (row j0,
    // columns i0-32 .. i0+31), layers from k_start on and lanes from l_rot on
    {
        const unsigned long long rb = (unsigned long long)__double_as_longlong(R);
        const unsigned hs = sr_hash((unsigned)rb ^ sr_hash((unsigned)(rb >> 32) + (unsigned)t));
        const int k_start = (int)(hs % (unsigned)A.nz), l_rot = (int)((hs >> 24) & 63u);
        const int i = i0 - 32 + lane;
        const bool inr = i >= 0 && i < A.nx_;
        const size_t p = inr ? (size_t)j0 * A.nx_ + i : 0;
        const bool col = inr && i >= A.imin && i <= A.imax && j0 >= A.jmin && j0 <= A.jmax && A.tmask[p] == 1;
        const double a = col ? area[p] : 0.0;
        for (int kk = 0; kk < A.nz; kk++) {
            const int k = k_start + kk < A.nz ? k_start + kk : k_start + kk - A.nz;
            const size_t pk = (size_t)k * n2 + p;
            float c = col ? ft[pk] : 0.f;
            const double m = col ? a * A.hn[pk] : 0.0;
            const int e = ilogbf(c);
            const bool normal = col && c != 0.f && e >= -126;
            const double ulp = normal ? exp2((double)(e - 23)) : 0.0;
            const double wm = m * ulp;
            unsigned long long cand = __ballot(normal && wm > 0.0 && wm <= 2.0 * fabs(R));
            // lanes in the order l_rot, l_rot + 1, ..., 63, 0, ..., l_rot - 1
            cand = l_rot ? (cand >> l_rot) | (cand << (64 - l_rot)) : cand;
            while (cand) {
                const int l = (__ffsll((long long)cand) - 1 + l_rot) & 63;
                cand &= cand - 1;
                const double wl = __shfl(wm, l);
                const double n = fmin(fmax(rint(R / wl), -1.0), 1.0);
                if (n == 0.0) continue;
                double dR = 0.0;
                if (lane == l) {
                    const float cn = c + (float)(n * ulp);
                    dR = m * ((double)cn - (double)c);
                    c = cn;
                    ft[pk] = cn;
                }
                R -= __shfl(dR, l);
                if (fabs(R) <= tol) return;
            }
        }
    }
    // ---- walk with up to GREEDY_MAX_ULPS per cell
    const size_t nchunks = GREEDY_ROWS + (n2 + 63) / 64;
    size_t visited = 0;
    for (size_t s = 0; s < nchunks && visited < (size_t)max_cells; s++) {
        size_t p = 0;
        bool inr = false;
        if (s < GREEDY_ROWS) {
            const int j = j0 + ((s & 1) ? (int)(s + 1) / 2 : -(int)(s / 2));
            const int i = i0 - 32 + lane;
            inr = i >= 0 && i < A.nx_ && j >= 0 && j < A.ny_;
            if (inr) p = (size_t)j * A.nx_ + i;
        } else {
            const size_t q = (s - GREEDY_ROWS) * 64 + lane;
            inr = q < n2;
            if (inr) p = (p0 + q) % n2;
        }
        bool col = false;
        if (inr) {
            const int i = (int)(p % A.nx_), j = (int)(p / A.nx_);
            col = i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax && A.tmask[p] == 1;
        }
        const double a = col ? area[p] : 0.0;
        visited += (size_t)64 * A.nz;
        for (int k = 0; k < A.nz; k++) {
            const size_t pk = (size_t)k * n2 + p;
            float c = col ? ft[pk] : 0.f;
            const double m = col ? a * A.hn[pk] : 0.0;
            const int e = ilogbf(c);
            const bool normal = col && c != 0.f && e >= -126;
            const double ulp = normal ? exp2((double)(e - 23)) : 0.0;
            const double wm = m * ulp;  // mass of one ulp in this cell
            unsigned long long cand = __ballot(normal && wm > 0.0 && wm <= 2.0 * fabs(R));
            while (cand) {
                const int l = __ffsll((long long)cand) - 1;
                cand &= cand - 1;
                const double wl = __shfl(wm, l);
                const double n = fmin(fmax(rint(R / wl), -GREEDY_MAX_ULPS), GREEDY_MAX_ULPS);
                if (n == 0.0) continue;
                double dR = 0.0;
                if (lane == l) {
                    const float cn = c + (float)(n * ulp);
                    dR = m * ((double)cn - (double)c);
                    c = cn;
                    ft[pk] = cn;
                }
                R -= __shfl(dR, l);
                if (fabs(R) <= tol) return;
            }
        }
    }
}

// Compensated FP32 fixer: every wet cell of tracer t becomes x + ratio[t] |x|
// in FP64 (x (1 + ratio) for a tracer of one sign; ratio = D / I_abs), encoded
// again; the inventory then equals the target up to the storage rounding of
// each cell (2^-40 of its value).
__global__ void __launch_bounds__(INV_THREADS) k_fix_c(const StepArgs A, float* __restrict__ f,
                                                       short* __restrict__ lo, const double* __restrict__ ratio)
{
    const int i = blockIdx.x * IBX + threadIdx.x, j = blockIdx.y * IBY + threadIdx.y;
    const int t = blockIdx.z;
    if (i >= A.nx_ || j >= A.ny_) return;
    const size_t p = (size_t)j * A.nx_ + i;
    if (!(i >= A.imin && i <= A.imax && j >= A.jmin && j <= A.jmax && A.tmask[p] == 1)) return;
    // r |x| is below half the residual unit 2^-39 |x| (every cell keeps its
    // stored value) when |r| < 2^-40: the inventory is already at the target to
    // the storage precision
    const double r = ratio[t];
    if (!(fabs(r) >= 0x1p-40)) return;
    float* fc = f + (size_t)t * A.n3 + p;
    short* lc = lo + (size_t)t * A.n3 + p;
    for (int k = 0; k < A.nz; k++) {
        const size_t o = (size_t)k * A.n2;
        const double x = cdecode(fc[o], lc[o]);
        float hi;
        short l;
        cencode(fma(fabs(x), r, x), hi, l);
        fc[o] = hi;
        lc[o] = l;
    }
}

int launch_inventory_c(const StepArgs& A, const double* h, const double* area, int ntr, const float* f,
                       const short* lo, double* part, double* inv, const double* target, double* ratio, hipStream_t s)
{
    dim3 block(IBX, IBY, 1);
    dim3 grid((A.nx_ + IBX - 1) / IBX, (A.ny_ + IBY - 1) / IBY, ntr);
    const int nblk = grid.x * grid.y;
    const dim3 tgrid(xcd_grid((size_t)nblk * ((ntr + INV_TCH - 1) / INV_TCH)));
    ProfileScope ps(GETM_K_INVENTORY, s);
    hipLaunchKernelGGL((k_inventory_tr<float, true, INV_TCH>), tgrid, block, 0, s, A, h, area, ntr, f, lo, part);
    hipLaunchKernelGGL(k_inventory_reduce, dim3(ntr), dim3(64), 0, s, ntr, nblk, part, inv, target, ratio);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

int launch_fix_c(const StepArgs& A, int ntr, float* f, short* lo, const double* ratio, hipStream_t s)
{
    dim3 block(IBX, IBY, 1);
    dim3 grid((A.nx_ + IBX - 1) / IBX, (A.ny_ + IBY - 1) / IBY, ntr);
    ProfileScope ps(GETM_K_FIX, s);
    hipLaunchKernelGGL(k_fix_c, grid, block, 0, s, A, f, lo, ratio);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

template <typename R>
int launch_inventory(const StepArgs& A, const double* h, const double* area, int ntr, const R* f, double* part,
                     double* inv, const double* target, double* ratio, hipStream_t s)
{
    dim3 block(IBX, IBY, 1);
    dim3 grid((A.nx_ + IBX - 1) / IBX, (A.ny_ + IBY - 1) / IBY, ntr);
    const int nblk = grid.x * grid.y;
    ProfileScope ps(GETM_K_INVENTORY, s);
    if constexpr (std::is_same<R, double>::value) {
        // FP64: no c/ulp sums (no fixer), tracers in chunks
        const dim3 tgrid(xcd_grid((size_t)nblk * ((ntr + INV_TCH - 1) / INV_TCH)));
        hipLaunchKernelGGL((k_inventory_tr<double, false, INV_TCH>), tgrid, block, 0, s, A, h, area, ntr, f, nullptr,
                           part);
        hipLaunchKernelGGL(k_inventory_reduce, dim3(ntr), dim3(64), 0, s, ntr, nblk, part, inv, target, ratio);
    } else {
        hipLaunchKernelGGL((k_inventory<R>), grid, block, 0, s, A, h, area, ntr, f, part, nullptr);
        hipLaunchKernelGGL(k_inventory_reduce, dim3(ntr), dim3(64), 0, s, ntr, nblk, part, inv, target, ratio);
    }
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
    hipLaunchKernelGGL(k_residual, dim3(ntr), dim3(64), 0, s, ntr, nblk, part2, target, resid);
    hipLaunchKernelGGL(k_greedy_f32, dim3(ntr), dim3(64), 0, s, A, ntr, f, area, resid, target, 1 << 18);
    GETM_HIP_OK(hipGetLastError());
    return GETM_OK;
}

size_t inventory_part_doubles(const StepArgs& A, int ntr)
{
    const size_t nblk = (size_t)((A.nx_ + IBX - 1) / IBX) * ((A.ny_ + IBY - 1) / IBY);
    return nblk * INV_SLOTS * (size_t)ntr;
}

size_t fix_part_doubles(const StepArgs& A, int ntr)
{
    const size_t nblk = (size_t)((A.nx_ + IBX - 1) / IBX) * ((A.ny_ + IBY - 1) / IBY);
    return nblk * 4 * (size_t)ntr;
}

template int launch_inventory<float>(const StepArgs&, const double*, const double*, int, const float*, double*,
                                     double*, const double*, double*, hipStream_t);
template int launch_inventory<double>(const StepArgs&, const double*, const double*, int, const double*, double*,
                                      double*, const double*, double*, hipStream_t);

}  // namespace getm
