// LBApp.hpp -- the lattice-Boltzmann application (src/lattice_boltzmann/
// LBApp.hpp:10-26): reads the Domain/Time/Output keys, builds CommonParams,
// owns the model from the factory and forwards init/step/output to it.
#pragma once

#include <memory>

#include "CommonParams.h"
#include "InputParams.hpp"
#include "LBBaseClass.hpp"

namespace mesobasic {

// Domain/Time/Output keys into CommonParams (LBApp.cpp:18-31), one GPU
// holding the whole lattice (rank 0 of 1).
CommonParams read_common_params(const InputParams& in);

class MesoBase {
public:
    virtual void initSystem() = 0;
    virtual void stepForward(int) = 0;
    virtual void writeOutput(int) = 0;
    virtual void finishOutput() {}
    virtual ~MesoBase() {}
};

class LBApp : public MesoBase {
public:
    explicit LBApp(const InputParams& in);  // throws if LBApp/type is not provided
    void initSystem() override;
    void stepForward(int step) override;
    void writeOutput(int step) override;
    void finishOutput() override;
    const CommonParams& params() const { return p; }
    LBBaseClass* model() { return lb_object.get(); }

private:
    int current_step = 0;
    CommonParams p{};
    std::unique_ptr<LBBaseClass> lb_object;
};

}  // namespace mesobasic
