// generatePoints.h — the lattice generator case drivers build scenes with.
// Same lattice as core/generatePoints.h:5-34: numP = int((L + 1e-10)/spacing)
// points per axis spread evenly over the box (cell-centred), x outermost and
// z innermost; a box thinner than one spacing on any axis yields its centre.
#pragma once

#include <vector>

#include "hostTypes.h"

inline std::vector<double3> getRegularPackedPoints(double3 origin, double3 size, double spacing) {
    const int nx = int((size.x + 1.e-10) / spacing);
    const int ny = int((size.y + 1.e-10) / spacing);
    const int nz = int((size.z + 1.e-10) / spacing);
    if (nx == 0 || ny == 0 || nz == 0) return std::vector<double3>(1, origin + 0.5 * size);
    const double dx = size.x / double(nx), dy = size.y / double(ny), dz = size.z / double(nz);
    std::vector<double3> out;
    out.reserve((size_t)nx * ny * nz);
    for (int i = 1; i <= nx; i++) {
        const double px = origin.x + dx * (double(i) - 0.5);
        for (int j = 1; j <= ny; j++) {
            const double py = origin.y + dy * (double(j) - 0.5);
            for (int k = 1; k <= nz; k++) out.push_back({px, py, origin.z + dz * (double(k) - 0.5)});
        }
    }
    return out;
}
