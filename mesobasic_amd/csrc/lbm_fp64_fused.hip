// lbm_fp64_fused.hip -- FP64 instantiation of the single-pass D3Q19 kernel
// (lbm_fused.h).  Compiled with -ffp-contract=off like lbm_fp64.hip, so its
// results are bit-identical to the two-pass kernels and to the reference.
#pragma clang fp contract(off)

#include "lbm_fused.h"
#include "lbm_storage.h"

namespace lbm {

bool fp64_fused_supported(const StepArgs& a, const Grid& g) { return fused::default_shape(g, a.q) >= 0; }

namespace {

template <bool W, bool T1, typename TL, bool NTS>
void launch(const double* Pin, int kind, double* Pout, const Grid& g, const StepArgs& a,
            const FusedTuning& t, hipStream_t s)
{
    const int ncx = (g.nx + t.chunk - 1) / t.chunk;
    const unsigned nb = static_cast<unsigned>((g.ny / TL::TY) * (g.nz / TL::TZ) * ncx);
    fused::k_fused<StF64, double, W, true, T1, TL, NTS><<<nb, TL::NT, 0, s>>>(
        Pin, kind, Pout, g, a.tau, a.G, 0.0, t.chunk, t.order);
}

template <bool W, bool T1, bool NTS>
void launch_shape(const double* Pin, int kind, double* Pout, const Grid& g, const StepArgs& a,
                  const FusedTuning& t, hipStream_t s)
{
    using namespace fused;
    switch (t.shape) {
    case 0: launch<W, T1, Tile<kShapeTY[0], kShapeTZ[0], kShapeZF[0]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 1: launch<W, T1, Tile<kShapeTY[1], kShapeTZ[1], kShapeZF[1]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 2: launch<W, T1, Tile<kShapeTY[2], kShapeTZ[2], kShapeZF[2]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 3: launch<W, T1, Tile<kShapeTY[3], kShapeTZ[3], kShapeZF[3]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    case 4: launch<W, T1, Tile<kShapeTY[4], kShapeTZ[4], kShapeZF[4]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    default: launch<W, T1, Tile<kShapeTY[5], kShapeTZ[5], kShapeZF[5]>, NTS>(Pin, kind, Pout, g, a, t, s); break;
    }
}

template <bool W, bool T1>
void launch_nts(const double* Pin, int kind, double* Pout, const Grid& g, const StepArgs& a,
                const FusedTuning& t, hipStream_t s)
{
    if (t.nts) launch_shape<W, T1, true>(Pin, kind, Pout, g, a, t, s);
    else launch_shape<W, T1, false>(Pin, kind, Pout, g, a, t, s);
}

}  // namespace

hipError_t fp64_fused_step(const StepArgs& a, const Grid& g, const double* Pin, int kind,
                           double* Pout, const FusedTuning& t, hipStream_t s)
{
    if (!fused::supported(g, a.q, t.shape) || t.chunk < 1) return hipErrorNotSupported;
    const bool tau1 = (a.tau == 1.0);   // x / 1.0 == x: skip 19 divides, same bits
    if (a.walls) {
        if (tau1) launch_nts<true, true>(Pin, kind, Pout, g, a, t, s);
        else launch_nts<true, false>(Pin, kind, Pout, g, a, t, s);
    } else {
        if (tau1) launch_nts<false, true>(Pin, kind, Pout, g, a, t, s);
        else launch_nts<false, false>(Pin, kind, Pout, g, a, t, s);
    }
    return hipGetLastError();
}

}  // namespace lbm

// ---- test hook (include/mesobasic_lbm_testing.h) ---------------------------

extern "C" int lbm_probe_tile_map(int nty, int ntz, int ncx, int order, int* ty, int* tz, int* cx)
{
    if (nty < 1 || ntz < 1 || ncx < 1 || !ty || !tz || !cx) return -1;
    const int nb = nty * ntz * ncx;
    for (int b = 0; b < nb; b++)
        lbm::fused::tile_of_block(b, nty, ntz, ncx, order, ty[b], tz[b], cx[b]);
    return 0;
}
