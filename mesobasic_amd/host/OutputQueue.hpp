// OutputQueue.hpp -- the output step of the models without stopping the run.
//
// The reference's outputLatticeBoltzmann (scmp3D.cpp:127-133,
// mcmp2D.cpp:131-137) copies r, u and fx to the host and formats a VTK file
// while the time loop waits.  Here each output takes an asynchronous snapshot
// of the lattice (lbm_snapshot_begin: the fields are formed and copied to
// pinned host memory while the next steps run); the snapshot is collected at
// the next output (or at finish()) and its VTK file is formatted and written
// by a writer thread while the time loop goes on enqueuing steps.  The files
// are the ones the synchronous path writes, byte for byte.
//
// Output/async = 0 (a key of this build) turns the overlap off: each output
// then waits for its snapshot and writes the file before returning.
#pragma once

#include <exception>
#include <thread>
#include <vector>

#include "../../include/mesobasic_lbm.h"
#include "CommonParams.h"

namespace mesobasic {

class OutputQueue {
public:
    // q = 19: density only (LBfluid3D::writeVTKFile); q = 9: density and the
    // Shan-Chen velocity u + F/(2 rho) (LBfluid2D::writeVTKFile).
    // capsule: the handle carries an IB capsule, whose mesh is written with
    // each output as capsule_<step>.vtk (capsule2D.cpp:158-164)
    OutputQueue(const CommonParams& p, int q, bool async, bool capsule = false)
        : p_(p), q_(q), async_(async), capsule_(capsule)
    {
    }
    ~OutputQueue();
    OutputQueue(const OutputQueue&) = delete;
    OutputQueue& operator=(const OutputQueue&) = delete;

    // The output of step `step` of component comp of handle h.
    void output(lbm_handle* h, int comp, int step);
    // Writes what is still pending, then waits for the handle's queued
    // device work; rethrows a writer's error.
    void finish(lbm_handle* h);

private:
    const CommonParams& p_;
    int q_;
    bool async_;
    bool capsule_;
    bool pending_ = false;
    std::vector<double> cap_x_, cap_y_;   // mesh of the pending output
    int pending_step_ = 0;
    std::thread writer_;
    std::exception_ptr writer_error_;

    void collect_and_write(lbm_handle* h);
    void join_writer();
};

}  // namespace mesobasic
