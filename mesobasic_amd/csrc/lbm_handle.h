// lbm_handle.h -- the lattice handle behind the C ABI (include/mesobasic_lbm.h).
// Shared by the translation units that implement its entry points; not part
// of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mesobasic_lbm.h"
#include "lbm_internal.h"

struct lbm_handle {
    lbm_desc d{};
    lbm::Grid g{};
    lbm::StepArgs a{};
    int elem = 8;             // bytes per stored population
    void* buf[2] = {nullptr, nullptr};
    int kind[2] = {lbm::kHoldsF, lbm::kHoldsF};
    int cur = 0;
    double* psi = nullptr;    // FP64 two-pass density field
    float* rref[2] = {nullptr, nullptr};  // node-shift reference density of buf[0] / buf[1]
    double* rho0 = nullptr;   // initial density (reference r array at N=0)
    double* tmp = nullptr;    // staging for fields / f / partials
    int64_t tmp_bytes = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t steps = 0;
    bool initialised = false;
    bool first_step_stored = false;  // LBfluid2D order: step 1 uses rho0, u=0
    bool prev_valid = false;         // buf[1-cur] holds f(N-1)
    bool ref_is_rho0 = false;        // reference arrays still hold rho0, u = 0
    int kernel = LBM_KERNEL_TWO_PASS;
    lbm::FusedTuning tune{-1, 64, 3, 1};  // fused kernel launch configuration
    int64_t bytes = 0;
    int device = 0;           // device current at lbm_create
    // asynchronous snapshot (lbm_snapshot_begin / lbm_snapshot_wait): fields
    // formed in snap_dev on the stream, copied to pinned snap_host on
    // copy_stream while the next steps run
    double* snap_dev = nullptr;
    double* snap_host = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_snap = nullptr, ev_copied = nullptr;
    bool snap_pending = false;
    int64_t snap_step = 0;
    // immersed-boundary capsule (lbm_capsule_attach): device arrays in one
    // allocation, nodes placed by lbm_capsule_set_nodes
    // D2Q9 FP64 temporal block in use (a fused::kFirstBlockShape id), -1: none
    int block = -1;
    bool has_capsule = false;
    bool capsule_nodes_set = false;
    double* capsule_mem = nullptr;
    lbm::Capsule cap{};
};

inline bool is_lowp(const lbm_handle* h) { return h->d.precision != LBM_FP64; }
