// coupledDamBreak.cpp — a case driver written against the solverBase/dataBase
// API: the coupled DEM–SPH dam break (1.6 x 0.6 x 0.6 m tank, 0.4 m water
// column in the corner, three layers of wall dummies, a seeded-jitter lattice
// of DEM spheres downstream, Hertz–Mindlin contact, tank planes).
//
//   coupledDamBreak [dx] [steps] [nx ny nz] [precision] [state.bin]
//
// dx = 0.005 gives 512,000 fluid + 452,376 wall + 2,000 DEM spheres.  The run
// prints its step rate; with a path it also writes the final fluid and solid
// positions and velocities (little-endian doubles, caller order).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "seededRandom.h"
#include "solverBase.h"

class coupledDamBreak : public solverBase {
public:
    coupledDamBreak(double dx, int steps, int nx, int ny, int nz, int precision)
        : dx_(dx), steps_(steps), n_{nx, ny, nz} {
        setPrecision(precision);
    }

    void conditionInitialize() override {
        setProblemName("coupledDamBreak");
        const double L = 1.6, W = 0.6, H = 0.6, g = 9.81, column = 0.4;
        const double3 origin = make_double3(-3 * dx_, -3 * dx_, -3 * dx_);
        const double3 size = make_double3(L + 6 * dx_, W + 6 * dx_, H + 6 * dx_);
        setDomain(origin, size);

        const double h = 1.3 * dx_;
        const double c0 = 10.0 * std::sqrt(2.0 * g * column);
        addFluid(getRegularPackedPoints(make_double3(0, 0, 0), make_double3(column, column, column), dx_),
                 make_double3(0, 0, 0), h, 1000.0, c0, 1.0e-6);

        std::vector<double3> wall;
        for (const double3& p : getRegularPackedPoints(origin, size, dx_))
            if (p.x < 0 || p.y < 0 || p.z < 0 || p.x > L || p.y > W) wall.push_back(p);
        addDummyParticleWall(wall, 0.5 * dx_);

        // DEM spheres: lattice at 0.025 m from (0.9, 0.175, 0.012), x outermost,
        // each position jittered by U(-0.1 r, 0.1 r) per axis
        const double r = 0.01, rho_s = 2500.0, spacing = 0.025;
        const size_t nd = (size_t)n_[0] * n_[1] * n_[2];
        const std::vector<double> jit = uniform(kRootSeed, 1001, 3 * nd, -0.1 * r, 0.1 * r);
        std::vector<double3> sph;
        size_t k = 0;
        for (int i = 0; i < n_[0]; i++)
            for (int j = 0; j < n_[1]; j++)
                for (int l = 0; l < n_[2]; l++, k++)
                    sph.push_back(make_double3(0.9 + spacing * i + jit[3 * k], 0.175 + spacing * j + jit[3 * k + 1],
                                               0.012 + spacing * l + jit[3 * k + 2]));
        addSolid(sph, make_double3(0, 0, 0), r, rho_s, 0);

        // E = 1e7 Pa, Poisson 0.3, fed as effective moduli E* = E/(2(1-v^2)),
        // G* = G/(2(2-v)) with G = E/(2(1+v))
        const double E = 1.0e7, nu = 0.3;
        const double Gm = E / (2 * (1 + nu));
        setHertzianContactModel(0, 0, E / (2 * (1 - nu * nu)), Gm / (2 * (2 - nu)), 0.5, 0, 0, 0.4, 0, 0);
        // the floor first: the wall loop stops at the first wall out of reach
        addInfiniteWall(make_double3(0, 0, 0), make_double3(0, 0, 1), 0);
        addInfiniteWall(make_double3(0, 0, 0), make_double3(1, 0, 0), 0);
        addInfiniteWall(make_double3(L, 0, 0), make_double3(-1, 0, 0), 0);
        addInfiniteWall(make_double3(0, 0, 0), make_double3(0, 1, 0), 0);
        addInfiniteWall(make_double3(0, W, 0), make_double3(0, -1, 0), 0);

        const double kc = 0.5 * E * pi() * r, mij = 8 * std::pow(r, 3) * rho_s;
        const double dt = std::fmin(0.25 * h / c0, 0.2 * std::sqrt(mij / kc));
        setSimulationParameterTimeStep(dt);
        setFluidIntegrateGap(1);
        // numSteps = int(T / dt) + 1 = steps_
        setSimulationParameterMaximumTime((steps_ - 0.5) * dt);
        setSimulationParameterNumFrames(1);
    }

    void writeState(const char* path) {
        const HostFluid& f = getHostFluidData();
        const HostSolid& s = getHostSolidData();
        FILE* out = std::fopen(path, "wb");
        if (!out) return;
        const int32_t n[2] = {f.points.num, s.points.num};
        std::fwrite(n, sizeof(int32_t), 2, out);
        std::fwrite(f.points.position.data(), sizeof(double3), f.points.position.size(), out);
        std::fwrite(f.dyn.velocities.data(), sizeof(double3), f.dyn.velocities.size(), out);
        std::fwrite(s.points.position.data(), sizeof(double3), s.points.position.size(), out);
        std::fwrite(s.dyn.velocities.data(), sizeof(double3), s.dyn.velocities.size(), out);
        std::fclose(out);
    }

    int particles() const { return hos.fluids.points.num + hos.solids.points.num; }
    int steps() const { return simPara.numSteps; }

private:
    double dx_;
    int steps_;
    int n_[3];
};

int main(int argc, char** argv) {
    const double dx = argc > 1 ? std::atof(argv[1]) : 0.005;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 100;
    const int nx = argc > 5 ? std::atoi(argv[3]) : 20;
    const int ny = argc > 5 ? std::atoi(argv[4]) : 10;
    const int nz = argc > 5 ? std::atoi(argv[5]) : 10;
    const int prec = argc > 6 ? std::atoi(argv[6]) : DSPH_PREC_FP64;
    try {
        coupledDamBreak problem(dx, steps, nx, ny, nz, prec);
        const auto t0 = std::chrono::steady_clock::now();
        problem.solve();
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("coupledDamBreak: %d particles, %d steps, %.3f s (%.4g particle-steps/s incl. set-up)\n",
                    problem.particles(), problem.steps(), sec, (double)problem.particles() * problem.steps() / sec);
        if (argc > 7) problem.writeState(argv[7]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "coupledDamBreak: %s\n", e.what());
        return 1;
    }
    return 0;
}
