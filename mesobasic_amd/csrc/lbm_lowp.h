// lbm_lowp.h -- reduced-precision forms of the Shan-Chen step (FP32
// compute; FP32 storage of f, or FP16/BF16 storage of d = f_i - w_i*R with
// R a constant shift (LBM_FP16S, LBM_BF16S) or the node shift, R = the
// reading node's density after the previous step (LBM_FP16N, LBM_BF16N)).
// Entry points used by the C-ABI layer.
//
// rin / rout: the node-shift density fields (n floats each) that the
// populations read / written by a call are relative to; nullptr for the
// constant-shift forms.
#pragma once

#include "lbm_internal.h"

namespace lbm {

inline bool lowp_node_shift(int precision)
{
    return precision == LBM_FP16N || precision == LBM_BF16N;
}

// Whether the fused pair kernel (two nodes per thread, lbm_fused_pair.h)
// serves this form: 16-bit storage, tau = 1, no walls.
inline bool lowp_pairs(int precision, const StepArgs& a)
{
    const bool b16 = precision == LBM_FP16S || precision == LBM_BF16S || lowp_node_shift(precision);
    return b16 && !a.walls && a.tau == 1.0;
}
// Whether a fused tile shape (one-node or pair id) can run this form and
// lattice; and the default shape (a pair tile where one fits), or -1.
bool lowp_shape_ok(int precision, const StepArgs& a, const Grid& g, int shape);
int lowp_default_shape(int precision, const StepArgs& a, const Grid& g);

// Kernel the handle will use (LBM_KERNEL_*), or -1 if `requested` cannot run.
int lowp_resolve_kernel(int precision, const StepArgs& a, const Grid& g, int requested);
// Initial equilibrium from rho0 into P (held as f at its own node); writes
// rout = rho0 with the node shift.
hipError_t lowp_init(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                     const double* rho0, void* P, float* psi, float* rout, hipStream_t s);
// rho0 != nullptr: the LBfluid2D first step (rho = rho0, u = 0).  psi is an
// n-float scratch field (unused by the fused kernel).
hipError_t lowp_step(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                     const void* Pin, int kind, void* Pout, const double* rho0, float* psi,
                     const float* rin, float* rout, int kernel, const FusedTuning& t,
                     hipStream_t s);
hipError_t lowp_moments(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                        const void* P, int kind, const float* rin, double* rho, double* u,
                        double* v, double* w, hipStream_t s);
hipError_t lowp_get_f(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                      const void* P, int kind, const float* rin, double* aos, hipStream_t s);
// Stores f (AoS doubles) as held at its own node; with the node shift also
// writes rout = sum f per node.
hipError_t lowp_set_f(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                      const double* aos, void* P, float* rout, hipStream_t s);
hipError_t lowp_conservation(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                             const void* P, int kind, const float* rin, double* partials,
                             int* nblocks, hipStream_t s);

// Fused D3Q19 kernel launch for one storage format (one translation unit
// per format, compiled in parallel).  SPI: 0 FP32, 1 FP16, 2 BF16.
template <int SPI, bool NS>
hipError_t lowp_fused_launch(const void* Pin, int kind, void* Pout, const Grid& g,
                             const StepArgs& a, float rs, const float* rin, float* rout,
                             const FusedTuning& t, hipStream_t s);
#define LBM_LOWP_FUSED_DECL(SPI, NS)                                                     \
    template <>                                                                          \
    hipError_t lowp_fused_launch<SPI, NS>(const void* Pin, int kind, void* Pout, const Grid& g, \
                                          const StepArgs& a, float rs, const float* rin,    \
                                          float* rout, const FusedTuning& t, hipStream_t s);
LBM_LOWP_FUSED_DECL(0, false)
LBM_LOWP_FUSED_DECL(1, false)
LBM_LOWP_FUSED_DECL(1, true)
LBM_LOWP_FUSED_DECL(2, false)
LBM_LOWP_FUSED_DECL(2, true)
#undef LBM_LOWP_FUSED_DECL

}  // namespace lbm
