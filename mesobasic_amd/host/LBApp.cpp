#include "LBApp.hpp"

#include <stdexcept>
#include <string>

namespace mesobasic {

CommonParams read_common_params(const InputParams& in)
{
    // LBApp.cpp:18-31
    CommonParams p{};
    p.NX = in("Domain/nx", 1);
    p.NY = in("Domain/ny", 1);
    p.NZ = in("Domain/nz", 1);
    p.dx = in("Domain/dx", 1.0);
    p.dy = in("Domain/dy", 1.0);
    p.dz = in("Domain/dz", 1.0);
    p.dt = in("Time/dt", 1.0);
    p.nstep = in("Time/nstep", 1);
    p.iskip = in("Output/iskip", 1);
    p.jskip = in("Output/jskip", 1);
    p.kskip = in("Output/kskip", 1);
    p.numOutputs = in("Output/numOutputs", 1);
    p.LX = p.NX * p.dx;
    p.LY = p.NY * p.dy;
    p.LZ = p.NZ * p.dz;
    // One GPU holds the whole lattice: the slab split of LBApp.cpp:37-60
    // degenerates to rank 0 of 1 owning every x plane.
    p.np = 1;
    p.rank = 0;
    p.nbrL = 0;
    p.nbrR = 0;
    p.nx = p.NX;
    p.ny = p.NY;
    p.nz = p.NZ;
    p.xOff = 0;
    if (p.iskip < 1 || p.jskip < 1 || p.kskip < 1)
        throw std::runtime_error("Output/iskip, jskip, kskip must be >= 1");
    return p;
}

LBApp::LBApp(const InputParams& in) : p(read_common_params(in))
{
    lb_object.reset(LBBaseClass::LBFactory(p, in));
    if (!lb_object)
        throw std::runtime_error("LBApp/type '" + in("LBApp/type", "mcmp2D") +
                                 "' is not provided by this build (scmp3D, scmp2D, mcmp2D, mcmp3D, mcmp2DFilm, "
                                 "mcmp3DFilm, capsule2D)");
}

void LBApp::initSystem() { lb_object->initLatticeBoltzmann(); }

void LBApp::stepForward(int step)
{
    // LBApp.cpp:98-114
    current_step = step;
    lb_object->setTimeStep(current_step);
    lb_object->updateLatticeBoltzmann();
}

void LBApp::writeOutput(int) { lb_object->outputLatticeBoltzmann(); }

void LBApp::finishOutput() { lb_object->finishOutput(); }

}  // namespace mesobasic
