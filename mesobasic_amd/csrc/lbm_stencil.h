// lbm_stencil.h -- compile-time D2Q9 / D3Q19 tables and exact term-skipping
// helpers shared by every kernel.
//
// Direction order and weights follow Stencil::setStencil
// (src/lattice_boltzmann/LBUtils/Stencil.cpp:5-120).  The order matters: it
// fixes the floating-point summation order of the moments and of the
// Shan-Chen force, and the bounce-back pairs.
//
// Zero-coefficient terms are skipped.  The reference adds f*e and
// (w*e)*psi with e in {0, +1, -1}; a product with e = 0 is +-0, and adding
// +-0 to an accumulator that started at +0.0 and only ever holds sums of
// nonzero terms leaves it bit-for-bit unchanged (such an accumulator is never
// -0 in round-to-nearest).  e = -1 turns an add into an exact subtract.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace lbm {

// Compile-time loop: fn(std::integral_constant<int, i>) for i in [0, N).
template <typename F, int... I>
__host__ __device__ __forceinline__ void static_for_impl(F&& fn, std::integer_sequence<int, I...>)
{
    (fn(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F> __host__ __device__ __forceinline__ void static_for(F&& fn)
{
    static_for_impl(static_cast<F&&>(fn), std::make_integer_sequence<int, N>{});
}

template <int Q> struct St;

template <> struct St<9> {
    static constexpr int D = 2;
    static constexpr int ex[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
    static constexpr int ey[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
    static constexpr int ez[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr int opp[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};
    // Stencil.cpp:32-40 (each a correctly rounded quotient, as there)
    static constexpr double w[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    // Walls on the last axis (y): directions leaving the low wall, i.e. the
    // ones replaced at row 1 (LBfluid2D.cpp:374-376), and their opposites,
    // replaced at row n-2 (LBfluid2D.cpp:382-384).
    static constexpr int nwall = 3;
    static constexpr int wall_up[3] = {3, 5, 7};
    static constexpr int wall_dn[3] = {4, 6, 8};
};

template <> struct St<19> {
    static constexpr int D = 3;
    static constexpr int ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
    static constexpr int ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
    static constexpr int ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
    static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    // Stencil.cpp:79-97
    static constexpr double w[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                     1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                     1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                     1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    // z walls: LBfluid3D.cpp:328-332 (row k=2) and :338-342 (row k=nz-1)
    static constexpr int nwall = 5;
    static constexpr int wall_up[5] = {5, 9, 11, 16, 18};
    static constexpr int wall_dn[5] = {6, 10, 12, 15, 17};
};

// Component of direction n along the last axis (z in 3-D, y in 2-D).
template <int Q> __host__ __device__ constexpr int e_last(int n)
{
    return St<Q>::D == 3 ? St<Q>::ez[n] : St<Q>::ey[n];
}

// acc (+)= c*t for c in {0,+1,-1}, exactly as acc += t*double(c).
template <int C, typename T> __device__ __forceinline__ void acc_e(T& acc, T t)
{
    if constexpr (C == 1) acc = acc + t;
    else if constexpr (C == -1) acc = acc - t;
}

// First two/three terms of e.u with e in {0,+1,-1}: ((ex*u + ey*v) + ez*w)
// with zero terms dropped (only the sign of a zero result can differ, and
// feq does not see it).
template <int CX, int CY, int CZ, typename T>
__device__ __forceinline__ T edot(T u, T v, T w)
{
    constexpr int nz = (CX != 0) + (CY != 0) + (CZ != 0);
    if constexpr (nz == 0) {
        return T(0);
    } else {
        T r;
        bool have = false;
        if constexpr (CX != 0) { r = (CX > 0) ? u : -u; have = true; }
        if constexpr (CY != 0) {
            if (have) r = (CY > 0) ? r + v : r - v;
            else { r = (CY > 0) ? v : -v; have = true; }
        }
        if constexpr (CZ != 0) {
            if (have) r = (CZ > 0) ? r + w : r - w;
            else r = (CZ > 0) ? w : -w;
        }
        return r;
    }
}

}  // namespace lbm
