// BatchServer.hpp -- many input.dat runs served together on one GPU.
//
// The reference runs one input.dat per process (MesoExecute.cpp:12-152).
// The server takes a list of run directories, each holding an input.dat for
// scmp3D or scmp2D, reads each with the driver's own key semantics
// (read_common_params, ScmpKeys), and runs them all as jobs of one batch
// engine (lbm_batch_*): a run is admitted when it fits the HBM budget,
// advances in every tick with the other runs, stops at each output step of
// its own cadence (nstep/numOutputs, MesoExecute.cpp:107-152), where its
// fields are read (lbm_batch_read) and its VTK file is written into
// <dir>/vtkoutput by a writer thread, and continues (lbm_batch_extend) until
// its nstep.  The files are those the driver writes for the run, computed in
// the batch's precision (FP32 arithmetic; FP32 or 16-bit storage).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mesobasic {

struct ServeOptions {
    int precision = 1;          // enum lbm_precision, one of the batch forms (LBM_FP32)
    double rho_shift = -1.0;    // constant-shift storage; < 0: the first run's LBApp/rhoAi
    int steps_per_tick = 20;
    int64_t hbm_budget = 0;     // bytes; 0 = 90% of free HBM
    int64_t work_limit = 0;     // lbm_batch_set_work_limit
    int max_wait = 0;           // lbm_batch_set_max_wait
    bool quiet = false;
};

// Runs every directory's input.dat to its nstep; returns the number of runs
// that failed (bad input, or retired by the engine as non-finite).  Each
// run's result line goes to stdout, a summary to stderr.
int serve_runs(const std::vector<std::string>& dirs, const ServeOptions& opt);

}  // namespace mesobasic
