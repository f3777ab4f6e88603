// LBBaseClass.hpp -- the model interface the reference's LB application
// drives (src/lattice_boltzmann/LBBaseClass.hpp:14-41), unchanged: the
// driver calls setTimeStep(step) then updateLatticeBoltzmann() each step and
// outputLatticeBoltzmann() at the output cadence.
#pragma once

#include "CommonParams.h"
#include "InputParams.hpp"

namespace mesobasic {

class LBBaseClass {
public:
    // Factory keyed on "LBApp/type" (LBBaseClass.cpp:21-45).  Returns
    // nullptr for a type this build does not provide, as the reference does
    // for an unknown type; the caller reports it.
    static LBBaseClass* LBFactory(const CommonParams&, const InputParams&);

    virtual void initLatticeBoltzmann() = 0;
    virtual void updateLatticeBoltzmann() = 0;
    virtual void outputLatticeBoltzmann() = 0;
    virtual void setTimeStep(int) = 0;
    // Completes output still in flight at the end of the run and waits for
    // the device to finish the steps queued after it (the reference writes
    // each output before returning and steps synchronously).
    virtual void finishOutput() {}

    virtual ~LBBaseClass() {}
};

}  // namespace mesobasic
