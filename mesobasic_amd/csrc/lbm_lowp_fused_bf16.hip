// lbm_lowp_fused_bf16.hip -- fused reduced-precision D3Q19 kernels, BF16 storage (constant and node shift).
#include "lbm_lowp_fused.inl"

namespace lbm {

template <>
hipError_t lowp_fused_launch<2, false>(const void* Pin, int kind, void* Pout, const Grid& g,
                                       const StepArgs& a, float rs, const float* rin, float* rout,
                                       const FusedTuning& t, hipStream_t s)
{
    return lowp_fused_detail::run<StBF16, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
}

template <>
hipError_t lowp_fused_launch<2, true>(const void* Pin, int kind, void* Pout, const Grid& g,
                                       const StepArgs& a, float rs, const float* rin, float* rout,
                                       const FusedTuning& t, hipStream_t s)
{
    return lowp_fused_detail::run<StBF16, true>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
}

}  // namespace lbm
