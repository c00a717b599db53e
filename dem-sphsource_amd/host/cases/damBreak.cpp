// damBreak.cpp — the reference's own case driver (driver/cases/damBreak.cpp:
// 14-58) on the solverBase/dataBase API of this implementation: a water
// column released in a long tank of wall dummies, an infinite floor plane,
// a cubic clump of touching pebbles downstream, two materials, the DEM time
// step from the pebble contact stiffness and fluid sub-cycling with
// gap = int(dt_f / dt).
//
//   damBreak [full|small] [steps] [precision] [state.bin]
//
// "full" is the reference's scene (8 x 0.7 x 0.8 m tank, 3.5 x 0.7 x 0.4 m
// column, 0.15 m cube of 216 pebbles at x = 5.2 m, 3 s of simulated time,
// 150 VTU frames); "small" keeps its spacing, pebbles, materials, time step
// and gap in a 1.0 x 0.35 x 0.4 m tank with a 0.3 m column and the cube
// 0.035 m from it.  With a step count the run stops after that many steps
// and writes one frame; with a path it also writes the final fluid and solid
// positions and velocities and the clump state (little-endian doubles,
// caller order).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "solverBase.h"

class damBreak : public solverBase {
public:
    struct Sizes {
        double3 tank, fluid, clumpOrigin;
    };

    damBreak(const Sizes& s, int steps, int precision) : size_(s), steps_(steps) {
        setPrecision(precision);
    }

    // a step at whose contact phase a block of 3 x 3 x 3 fluid particles is
    // poured in above the water (addFluid after the start), or -1
    int pourAt = -1;

    void handleDataAfterCalculateContact() override {
        if (simPara.iStep != pourAt) return;
        const double dx = particle_spacing;
        addFluid(getRegularPackedPoints(make_double3(2. * dx, 2. * dx, size_.fluid.z + 2. * dx),
                                        make_double3(3. * dx, 3. * dx, 3. * dx), dx),
                 make_double3(0, 0, 0), 1.3 * dx, 1000., 30., 1.e-6);
    }

    double particle_spacing = 0.025;
    double particle_radius = 0.5 * particle_spacing;
    double E = 0.3e9;
    double3 cubeSize = make_double3(0.15, 0.15, 0.15);
    double cubeDensity = 800;

    void conditionInitialize() override {
        setProblemName("damBreak");
        const double dx = particle_spacing;
        const double3 T = size_.tank;
        const double3 origin = make_double3(-3. * dx, -3. * dx, -3. * dx);
        const double3 domain = make_double3(T.x + 6. * dx, T.y + 6. * dx, T.z + 6. * dx);
        setDomain(origin, domain);

        addFluid(getRegularPackedPoints(make_double3(0, 0, 0), size_.fluid, dx), make_double3(0, 0, 0),
                 1.3 * dx, 1000., 30., 1.e-6);

        std::vector<double3> wall;
        for (const double3& p : getRegularPackedPoints(origin, domain, dx))
            if (p.x < 0 || p.y < 0 || p.z < 0 || p.x > T.x || p.y > T.y) wall.push_back(p);
        addDummyParticleWall(wall, 0.5 * dx);
        addInfiniteWall(make_double3(0.5 * T.x, 0.5 * T.y, 0), make_double3(0, 0, 1), 0);

        setHertzianContactModel(0, 1, E, 0.3 * E, 0.9, 0, 0, 0.35, 0, 0);
        setHertzianContactModel(1, 1, 0.5 * E, 0.15 * E, 0.9, 0, 0, 0.45, 0, 0);

        const std::vector<double3> cube = getRegularPackedPoints(size_.clumpOrigin, cubeSize, 2 * particle_radius);
        const std::vector<double> rad(cube.size(), particle_radius);
        const double mass = cubeDensity * cubeSize.x * cubeSize.y * cubeSize.z;
        const symMatrix inertia =
            make_symMatrix(mass / 12. * (cubeSize.y * cubeSize.y + cubeSize.z * cubeSize.z),
                           mass / 12. * (cubeSize.x * cubeSize.x + cubeSize.z * cubeSize.z),
                           mass / 12. * (cubeSize.x * cubeSize.x + cubeSize.y * cubeSize.y), 0., 0., 0.);
        const double3 centroid = size_.clumpOrigin + 0.5 * cubeSize;
        addClump(cube, rad, centroid, make_double3(0, 0, 0), make_double3(0, 0, 0), mass, inertia, 1);

        const double k = 0.5 * E * pi() * particle_radius;
        const double mij = 8 * particle_radius * particle_radius * particle_radius * cubeDensity;
        const double dt = 0.2 * std::sqrt(mij / k);
        const double dt_f = 0.6 * 1.3 * dx / (22. * std::sqrt(0.4 * 9.8));
        int gap = int(dt_f / dt);
        if (gap < 1) gap = 1;
        setSimulationParameterTimeStep(dt);
        setFluidIntegrateGap(gap);
        if (steps_ > 0) {
            // numSteps = int(T / dt) + 1 = steps_
            setSimulationParameterMaximumTime((steps_ - 0.5) * dt);
            setSimulationParameterNumFrames(1);
        } else {
            setSimulationParameterMaximumTime(3.);
            setSimulationParameterNumFrames(150);
        }
    }

    void outputData() override {
        outputFluidVTU();
        outputSolidVTU();
    }

    void writeState(const char* path) {
        const HostFluid& f = getHostFluidData();
        const HostSolid& s = getHostSolidData();
        const HostClump& c = getHostClumpData();
        FILE* out = std::fopen(path, "wb");
        if (!out) return;
        const int32_t n[4] = {f.points.num, s.points.num, c.num, fluidIntegrateGap};
        std::fwrite(n, sizeof(int32_t), 4, out);
        std::fwrite(f.points.position.data(), sizeof(double3), f.points.position.size(), out);
        std::fwrite(f.dyn.velocities.data(), sizeof(double3), f.dyn.velocities.size(), out);
        std::fwrite(s.points.position.data(), sizeof(double3), s.points.position.size(), out);
        std::fwrite(s.dyn.velocities.data(), sizeof(double3), s.dyn.velocities.size(), out);
        std::fwrite(c.centroidPosition.data(), sizeof(double3), c.centroidPosition.size(), out);
        std::fwrite(c.dyn.velocities.data(), sizeof(double3), c.dyn.velocities.size(), out);
        std::fwrite(c.angularVelocities.data(), sizeof(double3), c.angularVelocities.size(), out);
        std::fclose(out);
    }

    int particles() const { return hos.fluids.points.num + hos.solids.points.num; }
    int steps() const { return simPara.numSteps; }
    int gap() const { return fluidIntegrateGap; }

private:
    Sizes size_;
    int steps_;
};

int main(int argc, char** argv) {
    const bool small = argc > 1 && std::strcmp(argv[1], "small") == 0;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 0;
    const int prec = argc > 3 ? std::atoi(argv[3]) : DSPH_PREC_FP64;
    damBreak::Sizes sz;
    if (small) {
        sz = {make_double3(1.0, 0.35, 0.4), make_double3(0.3, 0.35, 0.3), make_double3(0.31, 0.1, 0.0)};
    } else {
        sz = {make_double3(8.0, 0.7, 0.8), make_double3(3.5, 0.7, 0.4), make_double3(5.2, 0.275, 0.0)};
    }
    try {
        damBreak problem(sz, steps, prec);
        if (argc > 5 && std::strcmp(argv[5], "pour") == 0) problem.pourAt = 2;
        const auto t0 = std::chrono::steady_clock::now();
        problem.solve();
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("damBreak: %d particles, %d steps (gap %d), %.3f s incl. set-up and output\n",
                    problem.particles(), problem.steps(), problem.gap(), sec);
        if (argc > 4) problem.writeState(argv[4]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "damBreak: %s\n", e.what());
        return 1;
    }
    return 0;
}
