/*
 * vtu_ascii.cpp — the reference's ASCII VTU frame writers, restated.
 *
 * TEST INFRASTRUCTURE ONLY: the frames tests parse these files as the
 * reference's output of the same host arrays, and the benchmark's frame leg
 * times them (with the FP64 read-back they need) as the reference's way of
 * writing a frame.  Nothing in the product links it.
 *
 * Format (solverBase.cpp:69-333): an UnstructuredGrid of one VTK_VERTEX cell
 * per point, every value printed as text with std::fixed and 10 decimals
 * (solverBase.cpp:81, 181, 270), FieldData TIME and STEP, the Points, the
 * cells (connectivity, offsets, types), then the PointData arrays declared
 * Float32 / Int32.  The writers print from FP64 host mirrors; a reader turns
 * each value into the declared type.
 */
#include "oracle.h"

#include <cmath>
#include <fstream>
#include <iomanip>
#include <string>

namespace {

struct Ascii {
    std::ofstream out;
    explicit Ascii(const char* path) : out(path) { out << std::fixed << std::setprecision(10); }

    // header, FieldData and the opening of the Piece (solverBase.cpp:83-99)
    void head(double time, int32_t step, int64_t n) {
        out << "<?xml version=\"1.0\"?>\n"
               "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
               "  <UnstructuredGrid>\n"
               "    <FieldData>\n"
               "      <DataArray type=\"Float32\" Name=\"TIME\"  NumberOfTuples=\"1\" format=\"ascii\"> "
            << time << " </DataArray>\n"
               "      <DataArray type=\"Int32\"   Name=\"STEP\"  NumberOfTuples=\"1\" format=\"ascii\"> "
            << step << " </DataArray>\n"
               "    </FieldData>\n"
               "    <Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << n << "\">\n";
    }
    // Points from a per-point function p(i, k) (solverBase.cpp:101-109)
    template <class P>
    void points(int64_t n, P p) {
        out << "      <Points>\n"
               "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n";
        for (int64_t i = 0; i < n; i++) out << ' ' << p(i, 0) << ' ' << p(i, 1) << ' ' << p(i, 2);
        out << "\n        </DataArray>\n      </Points>\n";
    }
    // one VTK_VERTEX (type 1) cell per point (solverBase.cpp:111-123)
    void cells(int64_t n) {
        out << "      <Cells>\n        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n";
        for (int64_t i = 0; i < n; i++) out << ' ' << i;
        out << "\n        </DataArray>\n        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n";
        for (int64_t i = 1; i <= n; i++) out << ' ' << i;
        out << "\n        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
        for (int64_t i = 0; i < n; i++) out << " 1";
        out << "\n        </DataArray>\n      </Cells>\n";
    }
    // a PointData array of `comps` values per point, value v(i, k)
    template <class V>
    void array(const char* type, const char* name, int comps, int64_t n, V v) {
        out << "        <DataArray type=\"" << type << "\" Name=\"" << name << "\"";
        if (comps > 1) out << " NumberOfComponents=\"" << comps << "\"";
        out << " format=\"ascii\">\n";
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < comps; k++) out << ' ' << v(i, k);
        out << "\n        </DataArray>\n";
    }
    int tail() {
        out << "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
        out.close();
        return out.fail() ? -1 : 0;
    }
};

}  // namespace

extern "C" {

/* solverBase::outputFluidVTU (solverBase.cpp:69-170): smooth length,
 * velocity, density = rho0 + dRho, pressure = dRho c^2 */
int orc_vtu_fluid_ascii(const char* path, double time, int32_t step, int64_t n, const double* pos,
                        const double* h, const double* vel, const double* rho0, const double* drho,
                        const double* c) {
    Ascii a(path);
    if (!a.out) return -1;
    a.head(time, step, n);
    a.points(n, [&](int64_t i, int k) { return pos[3 * i + k]; });
    a.cells(n);
    a.out << "      <PointData Scalars=\"smooth length\">\n";
    a.array("Float32", "smooth length", 1, n, [&](int64_t i, int) { return h[i]; });
    a.array("Float32", "velocity", 3, n, [&](int64_t i, int k) { return vel[3 * i + k]; });
    a.array("Float32", "density", 1, n, [&](int64_t i, int) { return rho0[i] + drho[i]; });
    a.array("Float32", "pressure", 1, n, [&](int64_t i, int) { return drho[i] * std::pow(c[i], 2); });
    return a.tail();
}

/* solverBase::outputSolidVTU (solverBase.cpp:172-257): radius, clusterID,
 * clumpID, velocity, angularVelocity, normal */
int orc_vtu_solid_ascii(const char* path, double time, int32_t step, int64_t n, const double* pos,
                        const double* radius, const int32_t* cluster, const int32_t* clump,
                        const double* vel, const double* angvel, const double* normal) {
    Ascii a(path);
    if (!a.out) return -1;
    a.head(time, step, n);
    a.points(n, [&](int64_t i, int k) { return pos[3 * i + k]; });
    a.cells(n);
    a.out << "      <PointData Scalars=\"radius\">\n";
    a.array("Float32", "radius", 1, n, [&](int64_t i, int) { return radius[i]; });
    a.array("Int32", "clusterID", 1, n, [&](int64_t i, int) { return cluster[i]; });
    a.array("Int32", "clumpID", 1, n, [&](int64_t i, int) { return clump[i]; });
    a.array("Float32", "velocity", 3, n, [&](int64_t i, int k) { return vel[3 * i + k]; });
    a.array("Float32", "angularVelocity", 3, n, [&](int64_t i, int k) { return angvel[3 * i + k]; });
    a.array("Float32", "normal", 3, n, [&](int64_t i, int k) { return normal[3 * i + k]; });
    return a.tail();
}

/* solverBase::outputFluid2SolidVTU (solverBase.cpp:259-333): one point per
 * fluid–solid pair at the midpoint of the two particles, with the pair's
 * force; fluid[k], solid[k] index fluid_pos and solid_pos */
int orc_vtu_fluid_solid_ascii(const char* path, double time, int32_t step, int64_t n, const int32_t* fluid,
                              const int32_t* solid, const double* fluid_pos, const double* solid_pos,
                              const double* force) {
    Ascii a(path);
    if (!a.out) return -1;
    a.head(time, step, n);
    a.points(n, [&](int64_t i, int k) {
        return 0.5 * (fluid_pos[3 * (int64_t)fluid[i] + k] + solid_pos[3 * (int64_t)solid[i] + k]);
    });
    a.cells(n);
    a.out << "      <PointData Scalars=\"force\">\n";
    a.array("Float32", "force", 3, n, [&](int64_t i, int k) { return force[3 * i + k]; });
    return a.tail();
}

}  // extern "C"
