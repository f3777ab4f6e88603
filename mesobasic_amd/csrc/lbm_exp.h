// lbm_exp.h -- double-precision exp() for the device that returns the same
// bits as the host's glibc 2.35 exp() (x86-64, FMA code path) for every
// |x| < 512, so that psi(rho) = 1 - exp(-rho) on the GPU equals the
// reference's psi (scmp3D.cpp:187-190), which calls libm.
//
// Why: a 1-ulp difference in psi at 0.1% of the nodes grows to a relative
// L2 of ~5e-12 in u over 10,000 steps of spinodal decomposition; the FP64
// path has to be bit-faithful to libm, not merely accurate.
//
// Algorithm (published, ARM optimized-routines math/exp.c, used by glibc
// since 2.28; Szabolcs Nagy 2018): x = k ln2/128 + r with |r| <= ln2/256,
// exp(x) = 2^(k/128) * exp(r); 2^(k/128) from a 128-entry table of
// (tail, scale bits) pairs (exp_table.inc, generated from its definition by
// gen_exp_table.py); exp(r) - 1 by a degree-5 polynomial.  glibc's x86-64
// build selects an FMA-compiled copy of this code on CPUs with FMA, whose
// contractions are written out below as explicit fma() calls.
// Checked bit-for-bit against the host libm on 1.5e8 random arguments.
// For x < -512 (results below 2^-738, psi == 1 exactly whenever rho > 37.5)
// the subnormal branch is computed to within 1 ulp of libm.
#pragma once

#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define LBM_EXP_FN __device__ __forceinline__
#define LBM_EXP_MEMBER __device__ __forceinline__
#define LBM_EXP_TAB __device__ __constant__ static const
#else  // host build of the same code, used to test the algorithm on a CPU
#define LBM_EXP_FN static inline
#define LBM_EXP_MEMBER inline
#define LBM_EXP_TAB static const
#endif

namespace lbm {

LBM_EXP_TAB uint64_t kExpTab[256] = {
#include "exp_table.inc"
};

// exp_libm with the table read through tab(i) (e.g. a copy in LDS, which a
// kernel fills once per workgroup); same bits as exp_libm.
template <typename Tab> LBM_EXP_FN double exp_libm_tab(double x, const Tab& tab)
{
    constexpr double kInvLn2N = 0x1.71547652b82fep0 * 128.0;
    constexpr double kShift = 0x1.8p52;
    constexpr double kNegLn2hiN = -0x1.62e42fefa0000p-8;
    constexpr double kNegLn2loN = -0x1.cf79abc9e3b3ap-47;
    constexpr double kC2 = 0x1.ffffffffffdbdp-2;
    constexpr double kC3 = 0x1.555555555543cp-3;
    constexpr double kC4 = 0x1.55555cf172b91p-5;
    constexpr double kC5 = 0x1.1111167a4d017p-7;

    const uint64_t ux = __builtin_bit_cast(uint64_t, x);
    uint32_t abstop = static_cast<uint32_t>(ux >> 52) & 0x7ffu;
    // top12(2^-54) = 0x3c9, top12(512) = 0x408, top12(1024) = 0x409
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
        if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;   // |x| < 2^-54
        if (abstop >= 0x409u) {
            if (ux == 0xfff0000000000000ull) return 0.0;      // -inf
            if (abstop >= 0x7ffu) return 1.0 + x;              // inf, nan
            return (ux >> 63) ? 0.0 : __builtin_bit_cast(double, uint64_t(0x7ff0000000000000ull));
        }
        abstop = 0;                                            // 512 <= |x| < 1024
    }
    double kd = __builtin_fma(kInvLn2N, x, kShift);
    const uint64_t ki = __builtin_bit_cast(uint64_t, kd);
    kd -= kShift;
    const double r = __builtin_fma(kd, kNegLn2loN, __builtin_fma(kd, kNegLn2hiN, x));
    const uint32_t idx = 2u * static_cast<uint32_t>(ki & 127u);
    const uint64_t top = ki << 45;
    const double tail = __builtin_bit_cast(double, tab(idx));
    uint64_t sbits = tab(idx + 1) + top;
    const double r2 = r * r;
    const double tmp = __builtin_fma(r2 * r2, __builtin_fma(r, kC5, kC4),
                                     __builtin_fma(r2, __builtin_fma(r, kC3, kC2), tail + r));
    if (abstop == 0) {
        if ((ki & 0x80000000ull) == 0) {
            sbits -= 1009ull << 52;
            const double scale = __builtin_bit_cast(double, sbits);
            return 0x1p1009 * __builtin_fma(scale, tmp, scale);
        }
        sbits += 1022ull << 52;
        const double scale = __builtin_bit_cast(double, sbits);
        double y = __builtin_fma(scale, tmp, scale);
        if (y < 1.0) {
            double lo = __builtin_fma(scale, tmp, scale - y);
            const double hi = 1.0 + y;
            lo = 1.0 - hi + y + lo;
            y = (hi + lo) - 1.0;
            if (y == 0.0) y = 0.0;
        }
        return 0x1p-1022 * y;
    }
    const double scale = __builtin_bit_cast(double, sbits);
    return __builtin_fma(scale, tmp, scale);
}

struct ExpTabConst {
    LBM_EXP_MEMBER uint64_t operator()(uint32_t i) const { return kExpTab[i]; }
};

LBM_EXP_FN double exp_libm(double x) { return exp_libm_tab(x, ExpTabConst{}); }

}  // namespace lbm
