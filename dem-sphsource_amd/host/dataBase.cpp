// dataBase.cpp — setters, scene builders and state access of the host API
// (core/dataBase.h:33-219, core/dataBase.cpp:3-346) over include/dsph.h.
#include "dataBase.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>

static_assert(sizeof(double3) == 3 * sizeof(double), "double3 must be three packed doubles");
static_assert(sizeof(symMatrix) == 6 * sizeof(double), "symMatrix must be six packed doubles");
static_assert(sizeof(quaternion) == 4 * sizeof(double), "quaternion must be four packed doubles");

namespace {
const double* xyz(const std::vector<double3>& v) {
    return v.empty() ? nullptr : reinterpret_cast<const double*>(v.data());
}
double* xyz(std::vector<double3>& v) { return v.empty() ? nullptr : reinterpret_cast<double*>(v.data()); }
}  // namespace

dataBase::dataBase() {
    simPara = simulationParameter{1.0, 1.0, 1, 0, 1, 0, 1};
    gpuPara = GPUParameter{0, 256};
    domainOrigin = make_double3(0, 0, 0);
    domainSize = make_double3(1, 1, 1);
    gravity = make_double3(0, 0, -9.81);
    fluidIntegrateGap = 1;
}

dataBase::~dataBase() {
    if (dev) dsph_destroy(dev);
}

void dataBase::check(int rc) const {
    if (rc == DSPH_OK) return;
    const char* msg = dev ? dsph_last_error(dev)
                          : (rc == DSPH_E_NODEVICE ? "no HIP device" : "context creation failed");
    throw std::runtime_error(std::string("dsph error ") + std::to_string(rc) + ": " + (msg ? msg : ""));
}

// ----------------------------------------------------------------- setters
void dataBase::setSimulationParameterMaximumTime(double t) {
    if (t <= 0.) {
        std::cout << "Error: maximum time must be positive." << std::endl;
        return;
    }
    if (simPara.iStep > 0) {
        std::cout << "Warning: cannot change the maximum time during simulation." << std::endl;
        return;
    }
    simPara.maximumTime = t;
}

void dataBase::setSimulationParameterNumFrames(int n) {
    if (n < 1) {
        std::cout << "Error: number of frames must be at least 1." << std::endl;
        return;
    }
    if (simPara.iStep > 0) {
        std::cout << "Warning: cannot change the number of frames during simulation." << std::endl;
        return;
    }
    simPara.numFrames = n;
}

void dataBase::setSimulationParameterTimeStep(double dt) {
    if (dt <= 0.) {
        std::cout << "Error: time step must be positive." << std::endl;
        return;
    }
    if (simPara.iStep > 0) {
        std::cout << "Warning: cannot change the time step during simulation." << std::endl;
        return;
    }
    simPara.timeStep = dt;
}

void dataBase::setFluidIntegrateGap(int gap) {
    if (gap < 1) {
        std::cout << "Error: fluid integrate gap must be at least 1." << std::endl;
        return;
    }
    fluidIntegrateGap = gap;
}

void dataBase::setDomain(double3 origin, double3 size) {
    if (size.x <= 0. || size.y <= 0. || size.z <= 0.) {
        std::cout << "Error: domain size must be positive in all directions." << std::endl;
        return;
    }
    if (simPara.iStep > 0) {
        std::cout << "Error: the domain can only be set before the simulation starts." << std::endl;
        return;
    }
    domainOrigin = origin;
    domainSize = size;
}

void dataBase::setGravity(double3 g) {
    gravity = g;
    if (dev) {
        const double v[3] = {g.x, g.y, g.z};
        check(dsph_set_gravity(dev, v));
    }
}

void dataBase::setHertzianContactModel(int i, int j, double E, double G, double res, double k_r_k_s,
                                       double k_t_k_s, double mu_s, double mu_r, double mu_t) {
    const int c = hos.contactModels.getCombinedIndex(i, j);
    if (c < 0) {
        std::cout << "Error: material index exceeds the number of materials." << std::endl;
        return;
    }
    auto& h = hos.contactModels.hertzian;
    h.E[c] = E;
    h.G[c] = G;
    h.res[c] = res;
    h.k_r_k_s[c] = k_r_k_s;
    h.k_t_k_s[c] = k_t_k_s;
    h.mu_s[c] = mu_s;
    h.mu_r[c] = mu_r;
    h.mu_t[c] = mu_t;
    if (dev) uploadContactModels();
}

void dataBase::setLinearContactModel(int i, int j, double k_n, double k_s, double k_r, double k_t,
                                     double d_n, double d_s, double d_r, double d_t, double mu_s,
                                     double mu_r, double mu_t) {
    const int c = hos.contactModels.getCombinedIndex(i, j);
    if (c < 0) {
        std::cout << "Error: material index exceeds the number of materials." << std::endl;
        return;
    }
    auto& l = hos.contactModels.linear;
    l.k_n[c] = k_n;
    l.k_s[c] = k_s;
    l.k_r[c] = k_r;
    l.k_t[c] = k_t;
    l.d_n[c] = d_n;
    l.d_s[c] = d_s;
    l.d_r[c] = d_r;
    l.d_t[c] = d_t;
    l.mu_s[c] = mu_s;
    l.mu_r[c] = mu_r;
    l.mu_t[c] = mu_t;
    if (dev) uploadContactModels();
}

void dataBase::setBondedContactModel(int i, int j, double E, double k_n_k_s, double gamma,
                                     double sigma_s, double C, double mu) {
    const int c = hos.contactModels.getCombinedIndex(i, j);
    if (c < 0) {
        std::cout << "Error: material index exceeds the number of materials." << std::endl;
        return;
    }
    auto& b = hos.contactModels.bonded;
    b.E[c] = E;
    b.k_n_k_s[c] = k_n_k_s;
    b.gamma[c] = gamma;
    b.sigma_s[c] = sigma_s;
    b.C[c] = C;
    b.mu[c] = mu;
    if (dev) uploadContactModels();
}

void dataBase::setClumpVelocity(int index, double3 v) {
    if (index < 0 || index >= hos.clumps.num) return;
    hos.clumps.dyn.velocities[index] = v;
    if (dev) {
        const double a[3] = {v.x, v.y, v.z};
        check(dsph_set_clump_velocity(dev, index, a));
    }
}

// the contact tables flattened field-major, as dsph_scene carries them
void dataBase::uploadContactModels() {
    const auto& m = hos.contactModels;
    const auto& h = m.hertzian;
    const auto& l = m.linear;
    const auto& b = m.bonded;
    tables_[0].clear();
    tables_[1].clear();
    tables_[2].clear();
    for (const auto* v : {&h.E, &h.G, &h.res, &h.k_r_k_s, &h.k_t_k_s, &h.mu_s, &h.mu_r, &h.mu_t})
        tables_[0].insert(tables_[0].end(), v->begin(), v->end());
    for (const auto* v : {&l.k_n, &l.k_s, &l.k_r, &l.k_t, &l.d_n, &l.d_s, &l.d_r, &l.d_t, &l.mu_s,
                          &l.mu_r, &l.mu_t})
        tables_[1].insert(tables_[1].end(), v->begin(), v->end());
    for (const auto* v : {&b.gamma, &b.E, &b.k_n_k_s, &b.sigma_s, &b.C, &b.mu})
        tables_[2].insert(tables_[2].end(), v->begin(), v->end());
    if (dev && simPara.iStep > 0)
        check(dsph_set_contact_models(dev, tables_[0].data(), tables_[1].data(), tables_[2].data()));
}

// ----------------------------------------------------------------- builders
// Particles added after the start (dataBase.cpp:113-117, 226-268) go to
// the device in one dsph_append_scene: the running particles keep their
// state, the lists are searched again and an added cluster is bonded.
// The host mirrors grow with them; their state is read back on demand.
void dataBase::appendToDevice(const HostFluid& f, const HostSolid& o, const HostClump& c) {
    if (!dev) {
        std::cout << "Error: no device data to add particles to." << std::endl;
        return;
    }
    beforeDeviceRebuild();
    dsph_scene e{};
    e.n_fluid = f.points.num;
    e.fluid_pos = xyz(f.points.position);
    e.fluid_vel = xyz(f.dyn.velocities);
    e.fluid_h = f.points.effectiveRadii.data();
    e.fluid_rho0 = f.rho0.data();
    e.fluid_drho = f.dRho.data();
    e.fluid_c = f.c.data();
    e.fluid_nu = f.nu.data();
    e.n_solid = o.points.num;
    e.solid_pos = xyz(o.points.position);
    e.solid_vel = xyz(o.dyn.velocities);
    e.solid_angvel = xyz(o.angularVelocities);
    e.solid_normal = xyz(o.normal);
    e.solid_radius = o.radius.data();
    e.solid_eff_radius = o.points.effectiveRadii.data();
    e.solid_inv_mass = o.inverseMass.data();
    e.solid_material = o.materialID.data();
    e.solid_cluster = o.clusterID.data();
    e.solid_clump = o.clumpID.data();
    e.solid_dummy = o.isDummyParticle.data();
    e.n_clump = c.num;
    e.clump_centroid = xyz(c.centroidPosition);
    e.clump_vel = xyz(c.dyn.velocities);
    e.clump_angvel = xyz(c.angularVelocities);
    e.clump_orientation = c.orientations.empty() ? nullptr : &c.orientations[0].q0;
    e.clump_inv_inertia = c.inverseInertiaTensor.empty() ? nullptr : &c.inverseInertiaTensor[0].xx;
    e.clump_inv_mass = c.inverseMass.data();
    e.clump_pebble_start = c.pebbleStartIndex.data();
    e.clump_pebble_end = c.pebbleEndIndex.data();
    check(dsph_append_scene(dev, &e));
}

void dataBase::addFluidData(const HostFluid& f) {
    if (simPara.iStep > 0) appendToDevice(f, HostSolid(0), HostClump(0));
    hos.fluids.append(f);
}

void dataBase::addSolidData(const HostSolid& s) {
    if (simPara.iStep > 0) appendToDevice(HostFluid(0), s, HostClump(0));
    hos.solids.append(s);
}

void dataBase::addFluid(std::vector<double3> p, double3 velocity, double smoothLength, double density,
                        double soundSpeed, double kinematicViscosity) {
    HostFluid f((int)p.size());
    for (size_t i = 0; i < p.size(); i++) {
        f.points.position[i] = p[i];
        f.points.effectiveRadii[i] = smoothLength;
        f.dyn.velocities[i] = velocity;
        f.rho0[i] = density;
        f.c[i] = soundSpeed;
        f.nu[i] = kinematicViscosity;
    }
    addFluidData(f);
}

void dataBase::addDummyParticleWall(std::vector<double3> p, double radius) {
    HostSolid s((int)p.size());
    for (size_t i = 0; i < p.size(); i++) {
        s.points.position[i] = p[i];
        s.points.effectiveRadii[i] = 2.6 * radius;  // = 1.3 dx, the fluid h, for spacing dx = 2r
        s.radius[i] = radius;
        s.isDummyParticle[i] = 1;
    }
    addSolidData(s);
}

void dataBase::addSolid(std::vector<double3> p, double3 velocity, double radius, double density,
                        int materialID) {
    if (materialID >= hos.contactModels.nMaterial) {
        std::cout << "Error: material index exceeds the number of materials." << std::endl;
        return;
    }
    HostSolid s((int)p.size());
    for (size_t i = 0; i < p.size(); i++) {
        s.points.position[i] = p[i];
        s.dyn.velocities[i] = velocity;
        s.points.effectiveRadii[i] = 1.1 * radius;
        s.radius[i] = radius;
        s.materialID[i] = materialID;
        if (density > 0) s.inverseMass[i] = 1. / (4. / 3. * std::pow(radius, 3) * pi() * density);
    }
    addSolidData(s);
}

void dataBase::addInfiniteWall(double3 p, double3 normal, int materialID) {
    if (simPara.iStep > 0) {
        std::cout << "Error: infinite wall can only be added before the simulation starts." << std::endl;
        return;
    }
    hos.infiniteWalls.num++;
    hos.infiniteWalls.pointInPlane.push_back(p);
    hos.infiniteWalls.normal.push_back(normal);
    hos.infiniteWalls.materialID.push_back(materialID);
}

void dataBase::addCylinderWall(double3 p, double3 axis, double3 angularVelocity, double radius,
                               int materialID) {
    if (simPara.iStep > 0) {
        std::cout << "Error: cylinder wall can only be added before the simulation starts." << std::endl;
        return;
    }
    auto& c = hos.cylinderWalls;
    c.num++;
    c.pointInAxis.push_back(p);
    c.axis.push_back(axis);
    c.angularVelocity.push_back(angularVelocity);
    c.radius.push_back(radius);
    c.materialID.push_back(materialID);
}

void dataBase::addCluster(std::vector<double3> p, std::vector<double> radius, std::vector<double> density,
                          double3 velocity, int materialID) {
    if (p.size() != radius.size() || p.size() != density.size()) {
        std::cout << "Error: position, radius and density vectors must have the same size." << std::endl;
        return;
    }
    if (materialID >= hos.contactModels.nMaterial) {
        std::cout << "Error: material index exceeds the number of materials." << std::endl;
        return;
    }
    int clusterID = 0;
    if (hos.solids.points.num > 0)
        clusterID = *std::max_element(hos.solids.clusterID.begin(), hos.solids.clusterID.end()) + 1;
    HostSolid s((int)p.size());
    for (size_t i = 0; i < p.size(); i++) {
        s.points.position[i] = p[i];
        s.points.effectiveRadii[i] = 1.1 * radius[i];
        s.dyn.velocities[i] = velocity;
        s.radius[i] = radius[i];
        s.materialID[i] = materialID;
        s.clusterID[i] = clusterID;
        if (density[i] > 0) s.inverseMass[i] = 1. / (4. / 3. * std::pow(radius[i], 3) * pi() * density[i]);
    }
    addSolidData(s);
}

void dataBase::addClump(std::vector<double3> p, std::vector<double> radius, double3 centroidPosition,
                        double3 velocity, double3 angularVelocity, double mass, symMatrix inertiaTensor,
                        int materialID) {
    if (p.size() != radius.size()) {
        std::cout << "Error: position and radius vectors must have the same size." << std::endl;
        return;
    }
    if (materialID >= hos.contactModels.nMaterial) {
        std::cout << "Error: material index exceeds the number of materials." << std::endl;
        return;
    }
    HostClump c(1);
    c.centroidPosition[0] = centroidPosition;
    c.dyn.velocities[0] = velocity;
    c.angularVelocities[0] = angularVelocity;
    if (mass > 0.) c.inverseMass[0] = 1. / mass;
    c.inverseInertiaTensor[0] = inverse(inertiaTensor);
    c.pebbleStartIndex[0] = hos.solids.points.num;
    // pebble mass from the clump density over the pebbles' bounding cubes
    double volume = 0;
    for (double r : radius) volume += std::pow(2. * r, 3);
    const double density = volume > 0. ? mass / volume : 0.;
    HostSolid s((int)p.size());
    const int clumpID = hos.clumps.num;
    for (size_t i = 0; i < p.size(); i++) {
        s.points.position[i] = p[i];
        s.points.effectiveRadii[i] = 1.1 * radius[i];
        s.dyn.velocities[i] = velocity + cross(angularVelocity, p[i] - centroidPosition);
        s.angularVelocities[i] = angularVelocity;
        s.radius[i] = radius[i];
        s.materialID[i] = materialID;
        s.clumpID[i] = clumpID;
        if (density > 0. && radius[i] > 0) s.inverseMass[i] = 1. / (std::pow(2. * radius[i], 3) * density);
    }
    if (simPara.iStep > 0) {
        // pebbles and clump in one addition, the pebble range within it
        HostClump r = c;
        r.pebbleStartIndex[0] = 0;
        r.pebbleEndIndex[0] = s.points.num;
        for (size_t i = 0; i < p.size(); i++) s.clumpID[i] = 0;
        appendToDevice(HostFluid(0), s, r);
        for (size_t i = 0; i < p.size(); i++) s.clumpID[i] = clumpID;
    }
    hos.solids.append(s);
    c.pebbleEndIndex[0] = hos.solids.points.num;
    hos.clumps.append(c);
}

// ------------------------------------------------------- runtime mutation
void dataBase::addExternalForce(int index, double3 force) {
    if (index < 0 || index >= hos.solids.points.num) {
        std::cout << "Error: solid index out of range when adding an external force." << std::endl;
        return;
    }
    if (simPara.iStep == 0 || !dev) {
        std::cout << "Error: external forces can only be added after the simulation starts." << std::endl;
        return;
    }
    const double f[3] = {force.x, force.y, force.z};
    check(dsph_add_external_force(dev, index, f));
}

void dataBase::addGlobalDamping(int index, double C_d) {
    if (index < 0 || index >= hos.solids.points.num) {
        std::cout << "Error: solid index out of range when adding global damping." << std::endl;
        return;
    }
    if (simPara.iStep == 0 || !dev) {
        std::cout << "Error: global damping can only be added after the simulation starts." << std::endl;
        return;
    }
    check(dsph_add_global_damping(dev, index, C_d));
}

// ------------------------------------------------------------- device data
void dataBase::buildDeviceData() {
    if (dev) {
        dsph_destroy(dev);
        dev = nullptr;
    }
    check(dsph_create(gpuPara.deviceIndex, precision_, &dev));
    check(dsph_set_flags(dev, flags_));
    check(dsph_set_tuning(dev, DSPH_TUNE_PAIR_BLOCK, tunePair_));
    check(dsph_set_tuning(dev, DSPH_TUNE_BUILD_BLOCK, tuneBuild_));
    check(dsph_set_tuning(dev, DSPH_TUNE_GROUP, tuneGroup_));
    check(dsph_set_skin(dev, tuneSkin_));
    uploadContactModels();
    dsph_scene s{};
    s.domain_origin[0] = domainOrigin.x;
    s.domain_origin[1] = domainOrigin.y;
    s.domain_origin[2] = domainOrigin.z;
    s.domain_size[0] = domainSize.x;
    s.domain_size[1] = domainSize.y;
    s.domain_size[2] = domainSize.z;
    s.gravity[0] = gravity.x;
    s.gravity[1] = gravity.y;
    s.gravity[2] = gravity.z;
    const HostFluid& f = hos.fluids;
    s.n_fluid = f.points.num;
    s.fluid_pos = xyz(f.points.position);
    s.fluid_vel = xyz(f.dyn.velocities);
    s.fluid_h = f.points.effectiveRadii.data();
    s.fluid_rho0 = f.rho0.data();
    s.fluid_drho = f.dRho.data();
    s.fluid_c = f.c.data();
    s.fluid_nu = f.nu.data();
    const HostSolid& o = hos.solids;
    s.n_solid = o.points.num;
    s.solid_pos = xyz(o.points.position);
    s.solid_vel = xyz(o.dyn.velocities);
    s.solid_angvel = xyz(o.angularVelocities);
    s.solid_radius = o.radius.data();
    s.solid_eff_radius = o.points.effectiveRadii.data();
    s.solid_inv_mass = o.inverseMass.data();
    s.solid_material = o.materialID.data();
    s.solid_cluster = o.clusterID.data();
    s.solid_clump = o.clumpID.data();
    s.solid_dummy = o.isDummyParticle.data();
    const HostClump& c = hos.clumps;
    s.n_clump = c.num;
    s.clump_centroid = xyz(c.centroidPosition);
    s.clump_vel = xyz(c.dyn.velocities);
    s.clump_angvel = xyz(c.angularVelocities);
    s.clump_orientation = c.orientations.empty() ? nullptr : &c.orientations[0].q0;
    s.clump_inv_inertia = c.inverseInertiaTensor.empty() ? nullptr : &c.inverseInertiaTensor[0].xx;
    s.clump_inv_mass = c.inverseMass.data();
    s.clump_pebble_start = c.pebbleStartIndex.data();
    s.clump_pebble_end = c.pebbleEndIndex.data();
    const HostInfiniteWall& w = hos.infiniteWalls;
    s.n_wall = w.num;
    s.wall_point = xyz(w.pointInPlane);
    s.wall_normal = xyz(w.normal);
    s.wall_material = w.materialID.data();
    const HostCylinderWall& y = hos.cylinderWalls;
    s.n_cylinder = y.num;
    s.cylinder_point = xyz(y.pointInAxis);
    s.cylinder_axis = xyz(y.axis);
    s.cylinder_angvel = xyz(y.angularVelocity);
    s.cylinder_radius = y.radius.data();
    s.cylinder_material = y.materialID.data();
    s.n_material = hos.contactModels.nMaterial;
    s.hertz = tables_[0].data();
    s.linear = tables_[1].data();
    s.bonded = tables_[2].data();
    check(dsph_upload_scene(dev, &s));
}

void dataBase::downloadState() {
    if (!dev) return;
    HostFluid& f = hos.fluids;
    if (f.points.num > 0)
        check(dsph_download_fluid(dev, xyz(f.points.position), xyz(f.dyn.velocities),
                                  xyz(f.dyn.accelerations), f.dRho.data(), nullptr));
    HostSolid& o = hos.solids;
    if (o.points.num > 0)
        check(dsph_download_solid(dev, xyz(o.points.position), xyz(o.dyn.velocities),
                                  xyz(o.dyn.accelerations), xyz(o.angularVelocities), xyz(o.torques),
                                  xyz(o.normal)));
    HostClump& c = hos.clumps;
    if (c.num > 0)
        check(dsph_download_clump(dev, xyz(c.centroidPosition), xyz(c.dyn.velocities),
                                  xyz(c.dyn.accelerations), xyz(c.angularVelocities), xyz(c.torques),
                                  &c.orientations[0].q0));
}

const HostSolid& dataBase::getHostSolidData() {
    if (simPara.iStep >= 1 && dev && hos.solids.points.num > 0) {
        HostSolid& o = hos.solids;
        check(dsph_download_solid(dev, xyz(o.points.position), xyz(o.dyn.velocities),
                                  xyz(o.dyn.accelerations), xyz(o.angularVelocities), xyz(o.torques),
                                  xyz(o.normal)));
    }
    return hos.solids;
}

const HostFluid& dataBase::getHostFluidData() {
    if (simPara.iStep >= 1 && dev && hos.fluids.points.num > 0) {
        HostFluid& f = hos.fluids;
        check(dsph_download_fluid(dev, xyz(f.points.position), xyz(f.dyn.velocities),
                                  xyz(f.dyn.accelerations), f.dRho.data(), nullptr));
    }
    return hos.fluids;
}

const HostClump& dataBase::getHostClumpData() {
    if (simPara.iStep >= 1 && dev && hos.clumps.num > 0) {
        HostClump& c = hos.clumps;
        check(dsph_download_clump(dev, xyz(c.centroidPosition), xyz(c.dyn.velocities),
                                  xyz(c.dyn.accelerations), xyz(c.angularVelocities), xyz(c.torques),
                                  &c.orientations[0].q0));
    }
    return hos.clumps;
}

// ---------------------------------------------------------------- tuning
// One coordinate at a time (skin, pair-loop workgroup, build block): each
// candidate runs warm-up steps (growing lists if needed) and then trial
// steps replayed from HIP graphs; the best value of a coordinate is kept
// before the next one is searched.  Results do not depend on the choice.
void dataBase::autotune() {
    const int trial = autotuneSteps_ > 0 ? autotuneSteps_ : 20, warm = 5;
    const double dt = simPara.timeStep;
    const int gap = fluidIntegrateGap;
    buildDeviceData();
    check(dsph_initialize(dev, 1));
    int step = 1;
    auto measure = [&](int pair, int build, double skin, int group) {
        check(dsph_set_tuning(dev, DSPH_TUNE_PAIR_BLOCK, pair));
        check(dsph_set_tuning(dev, DSPH_TUNE_BUILD_BLOCK, build));
        check(dsph_set_tuning(dev, DSPH_TUNE_GROUP, group));
        check(dsph_set_skin(dev, skin));
        // eager warm-up steps grow any list the new skin overflows
        for (int k = 0; k < warm; k++, step++) check(dsph_step(dev, step, dt, gap, 1));
        auto timed = [&]() {
            check(dsph_synchronize(dev));
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = dsph_step_n(dev, step, trial, dt, gap);
            step += trial;
            return std::make_pair(
                rc, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / trial);
        };
        auto r = timed();
        if (r.first == DSPH_E_CAPACITY) {  // grow, then time again
            check(dsph_step(dev, step++, dt, gap, 1));
            r = timed();
        }
        check(r.first);
        return r.second;
    };
    int pair = 256, build = 64, group = 1;
    double skin = 0.3, best = 1e300;
    for (int v : {1, 2, 4, 8}) {
        const double ms = measure(pair, build, skin, v);
        if (ms < best) best = ms, group = v;
    }
    best = 1e300;
    for (double v : {0.1, 0.2, 0.3, 0.5}) {
        const double ms = measure(pair, build, v, group);
        if (ms < best) best = ms, skin = v;
    }
    best = 1e300;
    for (int v : {64, 128, 256, 512}) {
        const double ms = measure(v, build, skin, group);
        if (ms < best) best = ms, pair = v;
    }
    best = 1e300;
    for (int v : {32, 48, 64}) {
        const double ms = measure(pair, v, skin, group);
        if (ms < best) best = ms, build = v;
    }
    std::cout << "Autotuned: row group " << group << ", pair-loop workgroup " << pair << ", build block "
              << build << ", skin " << skin << " h (" << best << " ms/step)" << std::endl;
    setTuning(pair, build, skin, group);
    dsph_destroy(dev);
    dev = nullptr;
}
