// Shared device helpers for the tracer-transport kernels (gfx950).
//
// The arithmetic restates the reference's FP64 operation order so the FP64
// kernels (built with -ffp-contract=off) are bit-identical to it:
//   limiters         src/operators/advection_schemes/limiters.h:1-44
//   face values      src/operators/advection_schemes/advection.F90.template:66-89
// MIN/MAX use the compare-and-select form (max(a,b) = a > b ? a : b, folded
// left to right) that flang gives Fortran's MIN/MAX, so NaN operands select
// the same operand as the reference does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace getm {

enum Scheme { HSIMT = 1, MUSCL = 2, P2_PDM = 3, SPLMAX13 = 4, SUPERBEE = 5, UPSTREAM = 6 };

constexpr int kItersMax = 200;  // advection.F90.template:239

template <typename R>
__device__ __forceinline__ R fmax2(R a, R b) { return a > b ? a : b; }
template <typename R>
__device__ __forceinline__ R fmin2(R a, R b) { return a < b ? a : b; }
template <typename R>
__device__ __forceinline__ R fmin3(R a, R b, R c) { return fmin2(fmin2(a, b), c); }

// Fortran CEILING to default integer as amdflang evaluates it: saturating, NaN -> 0.
__device__ __forceinline__ int fceiling(double x)
{
    if (x != x) return 0;
    double c = ceil(x);
    if (c >= 2147483647.0) return 2147483647;
    if (c <= -2147483648.0) return -2147483647 - 1;
    return (int)c;
}

template <int S, typename R>
__device__ __forceinline__ R limiter(R cfl, R ratio)
{
    if constexpr (S == HSIMT) {
        R kappa = R(1) - cfl;
        R x = R(0.25) * (kappa - (R(1) / R(3)) / kappa);
        R lim = (R(0.5) + x) + (R(0.5) - x) * ratio;
        lim = fmax2(R(0), lim);
        return fmin3(R(2) * ratio, lim, R(2));
    } else if constexpr (S == MUSCL) {
        return fmin3(R(2) * ratio, R(0.5) * (R(1) + ratio), R(2));
    } else if constexpr (S == P2_PDM) {
        R x = (R(1) / R(6)) * (R(1) - R(2) * cfl);
        R lim = (R(0.5) + x) + (R(0.5) - x) * ratio;
        return fmin3((R(2) * ratio) / (cfl + R(1.e-10)), lim, R(2) / (R(1) - cfl));
    } else if constexpr (S == SPLMAX13) {
        return fmin3(R(2) * ratio, (R(1) / R(3)) * fmax2(R(1) + R(2) * ratio, R(2) + ratio), R(2));
    } else if constexpr (S == SUPERBEE) {
        return fmax2(fmin2(R(2) * ratio, R(1)), fmin2(ratio, R(2)));
    } else {
        return R(0);
    }
}

// Flux-limited upstream face value from the upstream (fu), upstream-upstream (fuu)
// and downstream (fd) cell values.
template <int S, typename R>
__device__ __forceinline__ R face_value(R cfl, R fu, R fuu, R fd)
{
    R deltaf = fd - fu;
    R deltafu = fu - fuu;
    if (deltaf * deltafu > R(0)) {
        R ratio = deltafu / deltaf;
        R lim = limiter<S, R>(cfl, ratio);
        fu = fu + ((R(0.5) * lim) * (R(1) - cfl)) * deltaf;
    }
    return fu;
}

__device__ __forceinline__ bool open_face(int m) { return m == 1 || m == 2; }

// XCD-aware 1-D grid over nx * ny * nz logical blocks (x fastest).  Work-groups
// go to the 8 XCDs round-robin (block b to XCD b % 8) and each XCD has its own
// L2; block b takes logical block v = (b % 8) (gridDim.x / 8) + b / 8, so every
// XCD walks a contiguous range of the logical order and the blocks resident on
// one XCD are neighbours along x (then y): the cache lines that neighbouring
// blocks share (rows are not aligned to lines) come from that XCD's L2.
// gridDim.x is the block count rounded up to a multiple of 8; false: padding.
__device__ __forceinline__ bool xcd_block(unsigned nx, unsigned ny, unsigned nz, unsigned& bx, unsigned& by,
                                          unsigned& bz)
{
    const unsigned b = blockIdx.x, per = gridDim.x / 8;
    unsigned v = (b % 8) * per + b / 8;
    if (v >= nx * ny * nz) return false;
    bx = v % nx;
    v /= nx;
    by = v % ny;
    bz = v / ny;
    return true;
}

// ---------------------------------------------------------------------------
// Raw buffer access (CDNA buffer resources): the base and extent of an array
// in SGPRs, a 32-bit per-lane byte offset and a wave-uniform byte offset, so
// an access costs no 64-bit address arithmetic per lane and the per-level
// offsets of a column stay scalar.  Lane offsets at or beyond `bytes` read 0
// and write nothing (the hardware's range check); callers stay in range.
// ---------------------------------------------------------------------------
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc make_rsrc(const void* base, unsigned bytes)
{
    // stride 0 (raw), the gfx9 default third dword (32-bit data format)
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float ld_f32(Rsrc r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ short ld_i16(Rsrc r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(short, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
}
__device__ __forceinline__ double ld_f64(Rsrc r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void st_f32(Rsrc r, unsigned voff, unsigned soff, float x)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, voff, soff, 0);
}
__device__ __forceinline__ void st_i16(Rsrc r, unsigned voff, unsigned soff, short x)
{
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, x), r, voff, soff, 0);
}
__device__ __forceinline__ void st_f64(Rsrc r, unsigned voff, unsigned soff, double x)
{
    using u2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x), r, voff, soff, 0);
}
// Typed loads and stores of tracer values (FP32 or FP64).
template <typename T>
__device__ __forceinline__ T ld_val(Rsrc r, unsigned voff)
{
    if constexpr (sizeof(T) == 8) return ld_f64(r, voff, 0);
    else return ld_f32(r, voff, 0);
}
template <typename T>
__device__ __forceinline__ void st_val(Rsrc r, unsigned voff, T x)
{
    if constexpr (sizeof(T) == 8) st_f64(r, voff, 0, x);
    else st_f32(r, voff, 0, x);
}
// Lane offset that the range check of every resource of fewer than 2^32 - 1
// bytes rejects (with a zero wave-uniform offset): the load returns 0, the store
// writes nothing.  Lets a lane skip an access without an exec-mask branch.
constexpr unsigned RSRC_SKIP = 0xffffffffu;

// Size of the 1-D grid of xcd_block for n logical blocks.
inline unsigned xcd_grid(size_t n) { return (unsigned)((n + 7) / 8 * 8); }

// 1/x to about one ulp of double for x > 0: the hardware estimate refined by two
// Newton steps.  Used only where the result is rounded to FP32 afterwards (and
// x is a positive layer thickness), never on the bit-exact FP64 path.
__device__ __forceinline__ double rcp_nr(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// Element at a 32-bit byte offset from a wave-uniform base: lets the compiler
// use the scalar-base + 32-bit-offset form of global loads and stores instead
// of 64-bit address arithmetic per lane.  Callers keep byte offsets below 2^32
// (in-layer offsets: grids have nx_ ny_ * 8 < 2^32).
template <typename T>
__device__ __forceinline__ const T& at32(const T* base, unsigned byte_off)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <typename T>
__device__ __forceinline__ T& at32(T* base, unsigned byte_off)
{
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off);
}

// ---------------------------------------------------------------------------
// Unbiased (stochastic) rounding of FP32 tracer stores.  In a nearly steady flow
// the FP32 update of a cell sees nearly the same operands every step, so
// rounding to nearest errs the same way step after step: an increment below half
// an ulp is lost every time and the error grows linearly with the step count.
// Rounding f + d up or down with probability given by the exact residual makes
// the expected stored value exact, so rounding errors become a random walk.
// The random bits are a hash of (step, stage, tracer, cell): results are
// deterministic and do not depend on launch shapes, batching or graph replays.
// ---------------------------------------------------------------------------
// 32-bit integer hash (lowbias32 of the hash-prospector family)
__device__ __host__ __forceinline__ unsigned sr_hash(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Seed of one rounding site of one step: step count and stage within the step.
__device__ __host__ __forceinline__ unsigned sr_stage_seed(unsigned long long step, unsigned stage)
{
    return sr_hash(sr_hash((unsigned)step ^ sr_hash(stage * 0x85ebca6bu + 0x632be5abu)) ^ (unsigned)(step >> 32));
}

// Random bits of one cell at one rounding site (stage seed != 0).  Every tracer
// of the cell uses them, each XORed with its own key (sr_tracer_key): for one
// tracer the bits stay uniform and fresh at every step and stage, and the hash
// is computed once per cell rather than once per tracer.
__device__ __forceinline__ unsigned sr_cell_bits(unsigned stage_seed, unsigned cell)
{
    return sr_hash(cell ^ stage_seed);
}

// Key of a tracer (odd, so never 0).
__device__ __forceinline__ unsigned sr_tracer_key(int tracer)
{
    return sr_hash((unsigned)tracer * 0x9e3779b9u + 0x7f4a7c15u) | 1u;
}

// f + d in FP32, rounded stochastically with the 32 random bits r.  The FP64 sum
// of two floats is exact unless |d| < 2^-29 |f| (then its own rounding is far
// below an FP32 ulp); adding 29 uniform bits below the FP32 mantissa and
// truncating rounds away from zero with probability equal to the discarded
// fraction.  FP32-subnormal results take the conversion's rounding.
__device__ __forceinline__ float sr_add(float f, float d, unsigned r)
{
    const double x = (double)f + (double)d;
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    b += r >> 3;
    b &= ~((1ull << 29) - 1ull);
    return (float)__longlong_as_double((long long)b);
}

// Stochastic rounding of one tracer (key = its sr_tracer_key) or round to nearest (key 0).
struct SRound {
    unsigned key;
    __device__ __forceinline__ float add(float f, float d, unsigned cell_bits) const
    {
        return key ? sr_add(f, d, cell_bits ^ key) : f + d;
    }
};

// ---------------------------------------------------------------------------
// Compensated FP32 storage (precision GETM_F32C).  A tracer value x is stored
// as hi = (float)x, rounded to nearest, plus an int16 residual
//   lo = rint((x - hi) / 2^(e - 38)),
// with e the exponent of hi (hi in [2^e, 2^(e+1)); e = -126 for zero and FP32
// subnormals): x - hi in units of 2^-15 of hi's ulp, so |lo| <= 2^14 and the
// stored value carries 39 significant bits.  The kernels decode both into an
// FP64 value, work with it and encode the result once per launch; the store's
// rounding (at most 2^-16 of an FP32 ulp) is too small to accumulate over any
// practical number of steps, so no stochastic rounding is needed.
// ---------------------------------------------------------------------------
// 2^(e - 38) for the exponent e of hi, built from hi's exponent field eb
// (e = eb - 127, fields 0 and 1 both meaning -126): FP64 exponent field
// eb - 165 + 1023 = eb + 858.
__device__ __forceinline__ double res_unit(float hi)
{
    unsigned eb = (__float_as_uint(hi) >> 23) & 0xffu;
    eb = eb > 1u ? eb : 1u;
    return __hiloint2double((int)((eb + 858u) << 20), 0);
}

// x = hi + lo 2^(e - 38), exact (39 significant bits).
__device__ __forceinline__ double cdecode(float hi, short lo)
{
    return fma((double)lo, res_unit(hi), (double)hi);
}

// hi = (float)x, lo = rint((x - hi) 2^(38 - e)); x - hi is exact in FP64 and the
// scaling by 2^(38 - e) (FP64 exponent field 1023 + 165 - eb) is exact.
__device__ __forceinline__ void cencode(double x, float& hi, short& lo)
{
    hi = (float)x;
    const double r = x - (double)hi;
    unsigned eb = (__float_as_uint(hi) >> 23) & 0xffu;
    eb = eb > 1u ? eb : 1u;
    const double inv = __hiloint2double((int)((1188u - eb) << 20), 0);
    lo = (short)__double2int_rn(r * inv);
}

}  // namespace getm

// Instantiate a templated launcher for every scheme; `body` is called with a
// compile-time scheme constant.
#define GETM_SCHEME_SWITCH(scheme, ...)                                      \
    switch (scheme) {                                                        \
    case 1: { constexpr int S_ = 1; __VA_ARGS__; } break;                    \
    case 2: { constexpr int S_ = 2; __VA_ARGS__; } break;                    \
    case 3: { constexpr int S_ = 3; __VA_ARGS__; } break;                    \
    case 4: { constexpr int S_ = 4; __VA_ARGS__; } break;                    \
    case 5: { constexpr int S_ = 5; __VA_ARGS__; } break;                    \
    default: { constexpr int S_ = 6; __VA_ARGS__; } break;                   \
    }
