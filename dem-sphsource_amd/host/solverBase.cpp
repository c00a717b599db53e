// solverBase.cpp — the step loop of core/solverBase.cpp:3-67 over the C ABI,
// with the output hooks writing VTU frames on a background thread.
#include "solverBase.h"

#include <sys/stat.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

void solverBase::solve() {
    initialize();
    while (simPara.iStep <= simPara.numSteps) {
        update();
        simPara.iStep++;
    }
    joinFrames();
}

// solverBase::initialize (solverBase.cpp:13-47): the case sets up the scene,
// the step count follows from the maximum time, the device data is built and
// the initial search, bonds and wall normals are done before step 1
void solverBase::initialize() {
    std::cout << "Initializing..." << std::endl;
    hos.contactModels = HostSolidContactModel(10);
    conditionInitialize();
    std::cout << "Using GPU device " << gpuPara.deviceIndex << std::endl;
    simPara.numSteps = int(simPara.maximumTime / simPara.timeStep) + 1;
    simPara.frameInterval = simPara.numSteps / simPara.numFrames;
    if (simPara.frameInterval < 1) simPara.frameInterval = 1;
    simPara.iStep = 0;
    simPara.iFrame = 0;
    if (autotuneSteps_ > 0) autotune();
    buildDeviceData();
    check(dsph_initialize(dev, 1));
    ::mkdir(dir.c_str(), 0755);
    outputData();
    flushFrame();
    std::cout << "Initialization completed." << std::endl;
    simPara.iStep++;
}

// solverBase::update (solverBase.cpp:49-67): the five phases with the user
// hook between the contact forces and the second solid half-step
void solverBase::update() {
    const int i = simPara.iStep, gap = fluidIntegrateGap;
    const double dt = simPara.timeStep;
    check(dsph_neighbor_search(dev, i, gap));
    check(dsph_solid_before_contact(dev, dt));
    check(dsph_fluid_integrate(dev, dt, i, gap));
    check(dsph_contact_forces(dev, dt));
    handleDataAfterCalculateContact();
    check(dsph_solid_after_contact(dev, dt));
    if (simPara.iStep % simPara.frameInterval == 0) {
        simPara.iFrame++;
        std::cout << "Frame " << simPara.iFrame << " at time " << simPara.iStep * simPara.timeStep
                  << std::endl;
        outputData();
        flushFrame();
    }
}

// ------------------------------------------------------------------ output
// The reference's writers (solverBase.cpp:69-333) each copy the device state
// to the host mirrors and print one ASCII file while the GPU waits.  Here each
// adds its species to the frame; when outputData() returns the frame is
// packed on the device and its copy to pinned host memory queued beside the
// steps (dsph_frame_begin), and a writer thread writes the binary VTU files
// (dsph_frame_write) while the steps go on.  The two frame slots alternate;
// a slot's previous writer is joined before the slot is used again.
void solverBase::outputFluidVTU() { frameSpecies_ |= DSPH_FRAME_FLUID; }
void solverBase::outputSolidVTU() { frameSpecies_ |= DSPH_FRAME_SOLID; }
void solverBase::outputFluid2SolidVTU() { frameSpecies_ |= DSPH_FRAME_FLUID_SOLID; }

void solverBase::flushFrame() {
    if (!frameSpecies_ || !dev) return;
    const int slot = simPara.iFrame % DSPH_FRAME_SLOTS;
    if (writers_[slot].joinable()) writers_[slot].join();
    check(dsph_frame_begin(dev, slot, simPara.iStep, simPara.iStep * simPara.timeStep, frameSpecies_));
    dsph_ctx* const d = dev;
    const int frame = simPara.iFrame;
    const std::string out = dir;
    writers_[slot] = std::thread([d, slot, out, frame] {
        const int rc = dsph_frame_write(d, slot, out.c_str(), frame);
        if (rc != DSPH_OK) std::cerr << "frame " << frame << ": VTU files not written (code " << rc << ")\n";
    });
    frameSpecies_ = 0;
}

void solverBase::joinFrames() {
    for (std::thread& t : writers_)
        if (t.joinable()) t.join();
}
