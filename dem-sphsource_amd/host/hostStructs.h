// hostStructs.h — the host-side scene containers of the solverBase/dataBase
// API, with the reference's member names (core/hostStructs.h:7-399) so that
// case drivers reading getHostFluidData()/getHostSolidData() keep compiling.
// Plain std::vector SoA; the device copy lives behind include/dsph.h.
#pragma once

#include <vector>

#include "hostTypes.h"

struct HostDynamicStateBase {
    std::vector<double3> velocities;
    std::vector<double3> accelerations;
    HostDynamicStateBase() = default;
    explicit HostDynamicStateBase(int n) : velocities(n, double3{0, 0, 0}), accelerations(n, double3{0, 0, 0}) {}
    void append(const HostDynamicStateBase& o) {
        velocities.insert(velocities.end(), o.velocities.begin(), o.velocities.end());
        accelerations.insert(accelerations.end(), o.accelerations.begin(), o.accelerations.end());
    }
};

struct HostPointCloud {
    int num{0};
    std::vector<double3> position;
    std::vector<double> effectiveRadii;  // search radius
    HostPointCloud() = default;
    explicit HostPointCloud(int n) : num(n), position(n, double3{0, 0, 0}), effectiveRadii(n, 0.0) {}
    void append(const HostPointCloud& o) {
        num += o.num;
        position.insert(position.end(), o.position.begin(), o.position.end());
        effectiveRadii.insert(effectiveRadii.end(), o.effectiveRadii.begin(), o.effectiveRadii.end());
    }
};

template <class T>
inline void append_vec(std::vector<T>& a, const std::vector<T>& b) {
    a.insert(a.end(), b.begin(), b.end());
}

struct HostFluid {
    HostPointCloud points;
    HostDynamicStateBase dyn;
    std::vector<double> rho0, dRho, c, nu;
    HostFluid() = default;
    explicit HostFluid(int n) : points(n), dyn(n), rho0(n, 0.0), dRho(n, 0.0), c(n, 0.0), nu(n, 0.0) {}
    void append(const HostFluid& o) {
        points.append(o.points);
        dyn.append(o.dyn);
        append_vec(rho0, o.rho0);
        append_vec(dRho, o.dRho);
        append_vec(c, o.c);
        append_vec(nu, o.nu);
    }
};

struct HostSolid {
    HostPointCloud points;
    HostDynamicStateBase dyn;
    std::vector<double3> normal, torques, angularVelocities;
    std::vector<double> radius, inverseMass;
    std::vector<int> materialID, clusterID, clumpID, isDummyParticle;
    HostSolid() = default;
    explicit HostSolid(int n)
        : points(n), dyn(n), normal(n, double3{0, 0, 0}), torques(n, double3{0, 0, 0}),
          angularVelocities(n, double3{0, 0, 0}), radius(n, 0.0), inverseMass(n, 0.0),
          materialID(n, 0), clusterID(n, -1), clumpID(n, -1), isDummyParticle(n, 0) {}
    void append(const HostSolid& o) {
        points.append(o.points);
        dyn.append(o.dyn);
        append_vec(normal, o.normal);
        append_vec(torques, o.torques);
        append_vec(angularVelocities, o.angularVelocities);
        append_vec(radius, o.radius);
        append_vec(inverseMass, o.inverseMass);
        append_vec(materialID, o.materialID);
        append_vec(clusterID, o.clusterID);
        append_vec(clumpID, o.clumpID);
        append_vec(isDummyParticle, o.isDummyParticle);
    }
};

struct HostClump {
    int num{0};
    HostDynamicStateBase dyn;
    std::vector<double3> centroidPosition, torques, angularVelocities;
    std::vector<quaternion> orientations;
    std::vector<symMatrix> inverseInertiaTensor;
    std::vector<double> inverseMass;
    std::vector<int> pebbleStartIndex, pebbleEndIndex;
    HostClump() = default;
    explicit HostClump(int n)
        : num(n), dyn(n), centroidPosition(n, double3{0, 0, 0}), torques(n, double3{0, 0, 0}),
          angularVelocities(n, double3{0, 0, 0}), orientations(n, quaternion{1, 0, 0, 0}),
          inverseInertiaTensor(n, symMatrix{0, 0, 0, 0, 0, 0}), inverseMass(n, 0.0),
          pebbleStartIndex(n, 0), pebbleEndIndex(n, 0) {}
    void append(const HostClump& o) {
        num += o.num;
        dyn.append(o.dyn);
        append_vec(centroidPosition, o.centroidPosition);
        append_vec(torques, o.torques);
        append_vec(angularVelocities, o.angularVelocities);
        append_vec(orientations, o.orientations);
        append_vec(inverseInertiaTensor, o.inverseInertiaTensor);
        append_vec(inverseMass, o.inverseMass);
        append_vec(pebbleStartIndex, o.pebbleStartIndex);
        append_vec(pebbleEndIndex, o.pebbleEndIndex);
    }
};

struct HostInfiniteWall {
    int num{0};
    std::vector<double3> pointInPlane, normal;
    std::vector<int> materialID;
};

// kept for API compatibility: cylinder walls are stored but, as in the
// reference, never reach the device step
struct HostCylinderWall {
    int num{0};
    std::vector<double3> pointInAxis, axis, angularVelocity;
    std::vector<double> radius;
    std::vector<int> materialID;
};

// contact parameters per material pair (core/hostStructs.h:339-375)
struct HostSolidContactModel {
    int nMaterial{0};
    int nCombinations{0};
    struct {
        std::vector<double> E, G, res, k_r_k_s, k_t_k_s, mu_s, mu_r, mu_t;
    } hertzian;
    struct {
        std::vector<double> k_n, k_s, k_r, k_t, d_n, d_s, d_r, d_t, mu_s, mu_r, mu_t;
    } linear;
    struct {
        std::vector<double> E, k_n_k_s, gamma, sigma_s, C, mu;
    } bonded;
    HostSolidContactModel() = default;
    explicit HostSolidContactModel(int m) : nMaterial(m), nCombinations(m * (m + 1) / 2) {
        const size_t n = (size_t)nCombinations;
        for (auto* v : {&hertzian.E, &hertzian.G, &hertzian.res, &hertzian.k_r_k_s, &hertzian.k_t_k_s,
                        &hertzian.mu_s, &hertzian.mu_r, &hertzian.mu_t, &linear.k_n, &linear.k_s,
                        &linear.k_r, &linear.k_t, &linear.d_n, &linear.d_s, &linear.d_r, &linear.d_t,
                        &linear.mu_s, &linear.mu_r, &linear.mu_t, &bonded.E, &bonded.k_n_k_s,
                        &bonded.gamma, &bonded.sigma_s, &bonded.C, &bonded.mu})
            v->assign(n, 0.0);
        hertzian.res.assign(n, 1.0);  // the reference's constructor defaults
        bonded.gamma.assign(n, 1.0);
        bonded.k_n_k_s.assign(n, 2.0);
    }
    // index of the unordered pair (i, j) in the upper triangle, row by row:
    // i(2M - i + 1)/2 + j - i for i <= j; -1 when a material is out of range
    int getCombinedIndex(int i, int j) const {
        if (i < 0 || j < 0 || i >= nMaterial || j >= nMaterial) return -1;
        if (i > j) {
            const int t = i;
            i = j;
            j = t;
        }
        return i * (2 * nMaterial - i + 1) / 2 + j - i;
    }
};

struct HostData {
    HostFluid fluids;
    HostSolid solids;
    HostClump clumps;
    HostInfiniteWall infiniteWalls;
    HostCylinderWall cylinderWalls;
    HostSolidContactModel contactModels{10};
};
