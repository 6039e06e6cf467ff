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
