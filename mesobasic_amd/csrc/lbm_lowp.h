// lbm_lowp.h -- reduced-precision forms of the Shan-Chen step (FP32
// compute; FP32 storage, or FP16/BF16 storage of the shifted populations
// f_i - w_i*rho_shift).  Entry points used by the C-ABI layer.
#pragma once

#include "lbm_internal.h"

namespace lbm {

// Kernel the handle will use (LBM_KERNEL_*), or -1 if `requested` cannot run.
int lowp_resolve_kernel(int precision, const StepArgs& a, const Grid& g, int requested);
hipError_t lowp_init(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                     const double* rho0, void* P, float* psi, hipStream_t s);
// rho0 != nullptr: the LBfluid2D first step (rho = rho0, u = 0).  psi is an
// n-float scratch field (unused by the fused kernel).
hipError_t lowp_step(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                     const void* Pin, int kind, void* Pout, const double* rho0, float* psi,
                     int kernel, const FusedTuning& t, hipStream_t s);
hipError_t lowp_moments(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                        const void* P, int kind, double* rho, double* u, double* v, double* w,
                        hipStream_t s);
hipError_t lowp_get_f(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                      const void* P, int kind, double* aos, hipStream_t s);
hipError_t lowp_set_f(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                      const double* aos, void* P, hipStream_t s);
hipError_t lowp_conservation(int precision, const StepArgs& a, const Grid& g, double rho_shift,
                             const void* P, int kind, double* partials, int* nblocks,
                             hipStream_t s);

}  // namespace lbm
