// lbm_lowp_fused_f32.hip -- fused reduced-precision D3Q19 kernels, FP32 storage of f.
#include "lbm_lowp_fused.inl"

namespace lbm {

template <>
hipError_t lowp_fused_launch<0, false>(const void* Pin, int kind, void* Pout, const Grid& g,
                                       const StepArgs& a, float rs, const float* rin, float* rout,
                                       const FusedTuning& t, hipStream_t s)
{
    return lowp_fused_detail::run<StF32, false>(Pin, kind, Pout, g, a, rs, rin, rout, t, s);
}
}  // namespace lbm
