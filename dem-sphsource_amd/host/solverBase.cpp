// solverBase.cpp — the step loop of core/solverBase.cpp:3-67 over the C ABI,
// and plain ASCII VTU snapshots for the output hooks.
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
    }
}

// ------------------------------------------------------------------ output
namespace {
void vtu_points(std::ofstream& o, const std::vector<double3>& p) {
    o << "<Points><DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (const auto& q : p) o << q.x << ' ' << q.y << ' ' << q.z << '\n';
    o << "</DataArray></Points>\n<Cells>\n"
      << "<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
    for (size_t i = 0; i < p.size(); i++) o << i << '\n';
    o << "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
    for (size_t i = 0; i < p.size(); i++) o << i + 1 << '\n';
    o << "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (size_t i = 0; i < p.size(); i++) o << "1\n";
    o << "</DataArray>\n</Cells>\n";
}
void vtu_vec(std::ofstream& o, const char* name, const std::vector<double3>& v) {
    o << "<DataArray type=\"Float64\" Name=\"" << name << "\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (const auto& q : v) o << q.x << ' ' << q.y << ' ' << q.z << '\n';
    o << "</DataArray>\n";
}
template <class T>
void vtu_scalar(std::ofstream& o, const char* name, const char* type, const std::vector<T>& v) {
    o << "<DataArray type=\"" << type << "\" Name=\"" << name << "\" format=\"ascii\">\n";
    for (const auto& q : v) o << q << '\n';
    o << "</DataArray>\n";
}
bool vtu_open(std::ofstream& o, const std::string& dir, const char* stem, int frame, size_t n) {
    char name[64];
    std::snprintf(name, sizeof(name), "/%s%04d.vtu", stem, frame);
    o.open(dir + name);
    if (!o) return false;
    o.precision(10);
    o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\">\n"
      << "<UnstructuredGrid><Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << n << "\">\n";
    return true;
}
void vtu_close(std::ofstream& o) { o << "</Piece></UnstructuredGrid></VTKFile>\n"; }
}  // namespace

void solverBase::outputFluidVTU() {
    const HostFluid& f = getHostFluidData();
    std::ofstream o;
    if (!vtu_open(o, dir, "fluid", simPara.iFrame, f.points.position.size())) return;
    vtu_points(o, f.points.position);
    o << "<PointData>\n";
    vtu_vec(o, "velocity", f.dyn.velocities);
    std::vector<double> rho(f.rho0.size()), p(f.rho0.size());
    for (size_t i = 0; i < rho.size(); i++) {
        rho[i] = f.rho0[i] + f.dRho[i];
        p[i] = f.c[i] * f.c[i] * f.dRho[i];  // the solver's linear equation of state
    }
    vtu_scalar(o, "density", "Float64", rho);
    vtu_scalar(o, "pressure", "Float64", p);
    o << "</PointData>\n";
    vtu_close(o);
}

void solverBase::outputSolidVTU() {
    const HostSolid& s = getHostSolidData();
    std::vector<double3> pos;
    std::vector<double3> vel, w;
    std::vector<double> r;
    std::vector<int> mat, clump;
    for (int i = 0; i < s.points.num; i++) {
        if (s.isDummyParticle[i]) continue;
        pos.push_back(s.points.position[i]);
        vel.push_back(s.dyn.velocities[i]);
        w.push_back(s.angularVelocities[i]);
        r.push_back(s.radius[i]);
        mat.push_back(s.materialID[i]);
        clump.push_back(s.clumpID[i]);
    }
    std::ofstream o;
    if (!vtu_open(o, dir, "solid", simPara.iFrame, pos.size())) return;
    vtu_points(o, pos);
    o << "<PointData>\n";
    vtu_vec(o, "velocity", vel);
    vtu_vec(o, "angularVelocity", w);
    vtu_scalar(o, "radius", "Float64", r);
    vtu_scalar(o, "materialID", "Int32", mat);
    vtu_scalar(o, "clumpID", "Int32", clump);
    o << "</PointData>\n";
    vtu_close(o);
}

// fluid -> solid forces summed per solid, at the solids that receive any
void solverBase::outputFluid2SolidVTU() {
    if (!dev) return;
    const HostSolid& s = getHostSolidData();
    int64_t n = 0;
    check(dsph_download_list(dev, DSPH_LIST_FLUID_SOLID, nullptr, nullptr, 0, &n, nullptr));
    std::vector<int32_t> fi((size_t)n), sj((size_t)n);
    std::vector<double> F(3 * (size_t)n);
    check(dsph_download_list(dev, DSPH_LIST_FLUID_SOLID, fi.data(), sj.data(), n, &n, nullptr));
    check(dsph_download_fluid_solid_force(dev, F.data(), n, &n));
    std::vector<double3> sum((size_t)s.points.num, double3{0, 0, 0});
    std::vector<char> hit((size_t)s.points.num, 0);
    for (int64_t k = 0; k < n; k++) {
        sum[sj[k]] += double3{F[3 * k], F[3 * k + 1], F[3 * k + 2]};
        hit[sj[k]] = 1;
    }
    std::vector<double3> pos, force;
    for (int i = 0; i < s.points.num; i++)
        if (hit[i]) {
            pos.push_back(s.points.position[i]);
            force.push_back(sum[i]);
        }
    std::ofstream o;
    if (!vtu_open(o, dir, "fluid2solid", simPara.iFrame, pos.size())) return;
    vtu_points(o, pos);
    o << "<PointData>\n";
    vtu_vec(o, "fluidForce", force);
    o << "</PointData>\n";
    vtu_close(o);
}
