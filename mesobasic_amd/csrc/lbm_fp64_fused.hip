// lbm_fp64_fused.hip -- FP64 instantiation of the single-pass D3Q19 kernel
// (lbm_fused.h).  Compiled with -ffp-contract=off like lbm_fp64.hip, so its
// results are bit-identical to the two-pass kernels and to the reference.
#pragma clang fp contract(off)

#include "lbm_fused.h"
#include "lbm_storage.h"

namespace lbm {

bool fp64_fused_supported(const StepArgs& a, const Grid& g) { return fused::supported(g, a.q); }

hipError_t fp64_fused_step(const StepArgs& a, const Grid& g, const double* Pin, int kind,
                           double* Pout, hipStream_t s)
{
    if (!fused::supported(g, a.q)) return hipErrorNotSupported;
    const int ncx = (g.nx + fused::kChunkX - 1) / fused::kChunkX;
    const unsigned nb = static_cast<unsigned>((g.ny / fused::TY) * (g.nz / fused::TZ) * ncx);
    const bool tau1 = (a.tau == 1.0);
    if (a.walls) {
        if (tau1) fused::k_fused<StF64, double, true, true, true><<<nb, fused::NT, 0, s>>>(Pin, kind, Pout, g, a.tau, a.G, 0.0);
        else fused::k_fused<StF64, double, true, true, false><<<nb, fused::NT, 0, s>>>(Pin, kind, Pout, g, a.tau, a.G, 0.0);
    } else {
        if (tau1) fused::k_fused<StF64, double, false, true, true><<<nb, fused::NT, 0, s>>>(Pin, kind, Pout, g, a.tau, a.G, 0.0);
        else fused::k_fused<StF64, double, false, true, false><<<nb, fused::NT, 0, s>>>(Pin, kind, Pout, g, a.tau, a.G, 0.0);
    }
    return hipGetLastError();
}

}  // namespace lbm

// ---- test hook (include/mesobasic_lbm_testing.h) ---------------------------

extern "C" int lbm_probe_tile_map(int nty, int ntz, int ncx, int* ty, int* tz, int* cx)
{
    if (nty < 1 || ntz < 1 || ncx < 1 || !ty || !tz || !cx) return -1;
    const int nb = nty * ntz * ncx;
    for (int b = 0; b < nb; b++) lbm::fused::tile_of_block(b, nty, ntz, ncx, ty[b], tz[b], cx[b]);
    return 0;
}
