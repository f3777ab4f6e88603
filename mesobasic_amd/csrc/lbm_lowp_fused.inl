// lbm_lowp_fused.inl -- launcher of the reduced-precision fused D3Q19
// kernel (lbm_fused_lp.h) for one storage format; included by one
// translation unit per format.
#pragma once

#include <type_traits>

#include "../../include/mesobasic_lbm.h"
#include "lbm_fused_lp.h"
#include "lbm_fused_pair.h"
#include "lbm_lowp.h"
#include "lbm_storage.h"

namespace lbm {
namespace lowp_fused_detail {

template <typename SP, bool W, bool NS, bool T1, typename TL, bool NTS>
void launch(const void* Pin, int kind, void* Pout, const Grid& g, const StepArgs& a, float rs,
            const float* rin, float* rout, const FusedTuning& t, hipStream_t s)
{
    using T = typename SP::T;
    const int ncx = (g.nx + t.chunk - 1) / t.chunk;
    const int nty = (g.ny + TL::TY - 1) / TL::TY, ntz = (g.nz + TL::TZ - 1) / TL::TZ;
    const unsigned nb = static_cast<unsigned>(static_cast<int64_t>(nty) * ntz * ncx);
    const bool partial = !TL::ZF && (g.ny % TL::TY != 0 || g.nz % TL::TZ != 0);
    auto go = [&](auto part) {
        constexpr bool P = decltype(part)::value;
        if constexpr (!NS && !T1) {
            // constant shift (FP32 storage: shift 0, f itself): the shared
            // march kernel in its FP32-arithmetic form, converting at load
            fused::k_fused<SP, float, W, false, false, TL, NTS, P><<<nb, TL::NT, 0, s>>>(
                static_cast<const T*>(Pin), kind, static_cast<T*>(Pout), g, (float)a.tau,
                (float)a.G, rs, t.chunk, t.order);
        } else {
            fused::k_fused_lp<SP, W, NS, T1, TL, NTS, P><<<nb, TL::NT, 0, s>>>(
                static_cast<const T*>(Pin), kind, static_cast<T*>(Pout), g, (float)a.tau,
                (float)a.G, rs, rin, rout, t.chunk, t.order);
        }
    };
    if constexpr (!TL::ZF) {
        if (partial) {
            go(std::true_type{});
            return;
        }
    }
    go(std::false_type{});
}

template <typename SP, bool W, bool NS, bool T1, bool NTS>
void by_shape(const void* Pin, int kind, void* Pout, const Grid& g, const StepArgs& a, float rs,
              const float* rin, float* rout, const FusedTuning& t, hipStream_t s)
{
    using namespace fused;
#define LBM_LP_SHAPE(k)                                                                        \
    case k:                                                                                    \
        launch<SP, W, NS, T1, Tile<kShapeTY[k], kShapeTZ[k], kShapeZF[k]>, NTS>(Pin, kind, Pout, g, a, \
                                                                          rs, rin, rout, t, s);  \
        break;
    switch (t.shape) {
        LBM_LP_SHAPE(0)
        LBM_LP_SHAPE(1)
        LBM_LP_SHAPE(2)
        LBM_LP_SHAPE(3)
        LBM_LP_SHAPE(4)
        default: LBM_LP_SHAPE(5)
    }
#undef LBM_LP_SHAPE
}

// Pair tiles (16-bit storage, tau = 1, no walls; checked by the caller
// through lowp_shape_ok).
template <typename SP, bool NS, int I, bool NTS>
void launch_pair(const void* Pin, int kind, void* Pout, const Grid& g, const StepArgs& a, float rs,
                 const float* rin, float* rout, const FusedTuning& t, hipStream_t s)
{
    using namespace fused;
    using T = typename SP::T;
    using TL = PairTile<kPairTY[I], kPairTZ[I]>;
    const int ncx = (g.nx + t.chunk - 1) / t.chunk;
    const unsigned nb = static_cast<unsigned>((g.ny / TL::TY) * (g.nz / TL::TZ) * ncx);
    k_fused_pair<SP, NS, TL, NTS><<<nb, TL::NT, 0, s>>>(
        static_cast<const T*>(Pin), kind, static_cast<T*>(Pout), g, (float)a.tau, (float)a.G, rs,
        rin, rout, t.chunk, t.order);
}

template <typename SP, bool NS, bool NTS>
void by_pair_shape(const void* Pin, int kind, void* Pout, const Grid& g, const StepArgs& a,
                   float rs, const float* rin, float* rout, const FusedTuning& t, hipStream_t s)
{
    static_assert(fused::kNumPairShapes == 2, "one case per pair shape");
    if (t.shape == fused::kNumShapes) launch_pair<SP, NS, 0, NTS>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
    else launch_pair<SP, NS, 1, NTS>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
}

template <typename SP, bool NS, bool T1>
void by_flags(const void* Pin, int kind, void* Pout, const Grid& g, const StepArgs& a, float rs,
              const float* rin, float* rout, const FusedTuning& t, hipStream_t s)
{
    if constexpr (!std::is_same_v<SP, StF32> && T1) {
        if (fused::is_pair_shape(t.shape) && !a.walls) {
            if (t.nts) by_pair_shape<SP, NS, true>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
            else by_pair_shape<SP, NS, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
            return;
        }
    }
    if (a.walls) {
        if (t.nts) by_shape<SP, true, NS, T1, true>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
        else by_shape<SP, true, NS, T1, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
    } else {
        if (t.nts) by_shape<SP, false, NS, T1, true>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
        else by_shape<SP, false, NS, T1, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
    }
}

template <typename SP, bool NS>
hipError_t run(const void* Pin, int kind, void* Pout, const Grid& g, const StepArgs& a, float rs,
               const float* rin, float* rout, const FusedTuning& t, hipStream_t s)
{
    // tau = 1 has its own kernel (a waiting plane keeps only its momentum)
    if (a.tau == 1.0) by_flags<SP, NS, true>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
    else by_flags<SP, NS, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
    return hipGetLastError();
}

}  // namespace lowp_fused_detail
}  // namespace lbm
