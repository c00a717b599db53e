// solverBase.h — the reference's solver driver class (core/solverBase.h:4-47):
// case drivers derive from it, override conditionInitialize / outputData /
// handleDataAfterCalculateContact and call solve().
#pragma once

#include <string>

#include "dataBase.h"

class solverBase : public dataBase {
public:
    solverBase() : dataBase() {}
    ~solverBase() override {}

    // initialize, then steps 1 .. numSteps (solverBase.cpp:3-11)
    void solve();

protected:
    void setProblemName(const std::string& name) { dir = name + "Output"; }

    // ASCII VTU snapshots of the current state into the problem directory
    void outputFluidVTU();
    void outputSolidVTU();
    void outputFluid2SolidVTU();

    virtual void conditionInitialize() {}
    virtual void outputData() {}
    virtual void handleDataAfterCalculateContact() {}

    const std::string& outputDirectory() const { return dir; }

private:
    std::string dir = "problem";

    void initialize();
    virtual void update();
};
