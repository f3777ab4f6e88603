// lbm_lowp_fused_f16.hip -- fused reduced-precision D3Q19 kernels, FP16 storage (constant and node shift).
#include "lbm_lowp_fused.inl"

namespace lbm {

template <>
hipError_t lowp_fused_launch<1, false>(const void* Pin, int kind, void* Pout, const Grid& g,
                                       const StepArgs& a, float rs, const float* rin, float* rout,
                                       const FusedTuning& t, hipStream_t s)
{
    return lowp_fused_detail::run<StF16, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
}

template <>
hipError_t lowp_fused_launch<1, true>(const void* Pin, int kind, void* Pout, const Grid& g,
                                       const StepArgs& a, float rs, const float* rin, float* rout,
                                       const FusedTuning& t, hipStream_t s)
{
    return lowp_fused_detail::run<StF16, true>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
}

}  // namespace lbm
