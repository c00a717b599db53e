// solverBase.h — the reference's solver driver class (core/solverBase.h:4-47):
// case drivers derive from it, override conditionInitialize / outputData /
// handleDataAfterCalculateContact and call solve().
#pragma once

#include <cstdint>
#include <string>
#include <thread>

#include "dataBase.h"

class solverBase : public dataBase {
public:
    solverBase() : dataBase() {}
    ~solverBase() override { joinFrames(); }

    // initialize, then steps 1 .. numSteps (solverBase.cpp:3-11)
    void solve();

protected:
    void setProblemName(const std::string& name) { dir = name + "Output"; }

    // VTU files of the current state into the problem directory
    // (fluid_NNNN.vtu, solid_NNNN.vtu, fluid2Solid_NNNN.vtu): called from
    // outputData(), each adds its file to the frame, which is snapshot on the
    // device when outputData() returns and written by a background thread
    void outputFluidVTU();
    void outputSolidVTU();
    void outputFluid2SolidVTU();

    virtual void conditionInitialize() {}
    virtual void outputData() {}
    virtual void handleDataAfterCalculateContact() {}

    const std::string& outputDirectory() const { return dir; }

    // an addition rebuilds the device data: the frame writers finish first
    void beforeDeviceRebuild() override { joinFrames(); }

private:
    std::string dir = "problem";

    void initialize();
    virtual void update();

    // frames: the species asked for by this outputData(), and one writer
    // thread per frame slot
    uint32_t frameSpecies_ = 0;
    std::thread writers_[2];
    void flushFrame();
    void joinFrames();
};
