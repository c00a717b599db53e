// dataBase.h — the reference's scene/parameter host API
// (core/dataBase.h:6-219) over the C ABI of include/dsph.h.  Case drivers
// derive from solverBase (which derives from this class) and call the same
// protected setters and builders as upstream; the device data behind them is
// an opaque dsph_ctx instead of a struct of raw device pointers.
//
// Differences from the reference, all additive:
//  - setPrecision() selects the FP64 or the mixed (FP32 pair loops) path and
//    setBehaviourFlags() the DSPH_FLAG_* alternatives; the defaults are FP64
//    and the reference's own behaviour;
//  - particles, walls and the domain can only be added or changed before the
//    run starts (the reference also re-uploads mid-run); gravity, contact
//    models and clump velocities can change at any time.
#pragma once

#include <string>
#include <vector>

#include "dsph.h"
#include "generatePoints.h"
#include "hostStructs.h"

class dataBase {
public:
    dataBase();
    virtual ~dataBase();
    dataBase(const dataBase&) = delete;
    dataBase& operator=(const dataBase&) = delete;

protected:
    void setGPUParameterDeviceIndex(int index) { gpuPara.deviceIndex = index; }
    void setSimulationParameterMaximumTime(double t);
    void setSimulationParameterNumFrames(int n);
    void setSimulationParameterTimeStep(double dt);
    void setFluidIntegrateGap(int gap);
    void setDomain(double3 origin, double3 size);
    void setGravity(double3 g);
    void setHertzianContactModel(int mat_i, int mat_j, double E, double G, double res, double k_r_k_s,
                                 double k_t_k_s, double mu_s, double mu_r, double mu_t);
    void setLinearContactModel(int mat_i, int mat_j, double k_n, double k_s, double k_r, double k_t,
                               double d_n, double d_s, double d_r, double d_t, double mu_s,
                               double mu_r, double mu_t);
    void setBondedContactModel(int mat_i, int mat_j, double E, double k_n_k_s, double gamma,
                               double sigma_s, double C, double mu);
    void setClumpVelocity(int index, double3 velocity);

    void addFluid(std::vector<double3> p, double3 velocity, double smoothLength, double density,
                  double soundSpeed, double kinematicViscosity);
    void addDummyParticleWall(std::vector<double3> p, double radius);
    void addSolid(std::vector<double3> p, double3 velocity, double radius, double density,
                  int materialID);
    void addInfiniteWall(double3 p, double3 normal, int materialID);
    void addCylinderWall(double3 p, double3 axis, double3 angularVelocity, double radius,
                         int materialID);
    void addCluster(std::vector<double3> p, std::vector<double> radius, std::vector<double> density,
                    double3 velocity, int materialID);
    void addClump(std::vector<double3> p, std::vector<double> radius, double3 centroidPosition,
                  double3 velocity, double3 angularVelocity, double mass, symMatrix inertiaTensor,
                  int materialID);

    void addExternalForce(int index, double3 force);
    void addGlobalDamping(int index, double C_d);

    double getTime() const { return simPara.iStep * simPara.timeStep; }
    const HostSolid& getHostSolidData();
    const HostFluid& getHostFluidData();
    const HostClump& getHostClumpData();

    // additions of this implementation
    void setPrecision(int precision) { precision_ = precision; }
    void setBehaviourFlags(unsigned flags) { flags_ = flags; }
    dsph_ctx* context() const { return dev; }
    // before solve(): time trialSteps steps per candidate launch shape and
    // Verlet skin on a throwaway copy of the scene and run with the fastest
    void setAutotune(int trialSteps) { autotuneSteps_ = trialSteps; }
    // pick launch shapes and skin now (used by solverBase::initialize)
    void autotune();
    // the configuration the next device context gets
    void setTuning(int pairBlock, int buildBlock, double skin, int group = 1) {
        tunePair_ = pairBlock;
        tuneBuild_ = buildBlock;
        tuneSkin_ = skin;
        tuneGroup_ = group;
    }

    struct simulationParameter {
        double maximumTime;
        double timeStep;
        int numSteps;
        int iStep;
        int numFrames;
        int iFrame;
        int frameInterval;
    };
    struct GPUParameter {
        int deviceIndex;
        int maxThreadsPerBlock;
    };

    HostData hos;
    dsph_ctx* dev = nullptr;
    simulationParameter simPara;
    GPUParameter gpuPara;
    double3 domainOrigin;
    double3 domainSize;
    double3 gravity;
    int fluidIntegrateGap;

    // uploads hos to a (new) device context; state arrays are taken as they
    // are in hos
    void buildDeviceData();
    // device -> hos for every state field (after the run started)
    void downloadState();
    // throws std::runtime_error with the library's message on failure
    void check(int rc) const;

private:
    void addFluidData(const HostFluid& f);
    void addSolidData(const HostSolid& s);
    // particles added after the start, to the device in one addition
    void appendToDevice(const HostFluid& f, const HostSolid& o, const HostClump& c);

protected:
    // called before the device data is rebuilt for an addition: work that
    // reads it from other threads (frame writers) must end first
    virtual void beforeDeviceRebuild() {}

private:
    void uploadContactModels();
    int precision_ = DSPH_PREC_FP64;
    unsigned flags_ = 0;
    int tunePair_ = 256, tuneBuild_ = 64, tuneGroup_ = 1;
    double tuneSkin_ = 0.3;

protected:
    int autotuneSteps_ = 0;

private:
    // flat copies handed to dsph_upload_scene
    std::vector<double> tables_[3];
};
