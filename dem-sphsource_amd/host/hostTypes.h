// hostTypes.h — the small value types the solverBase/dataBase host API is
// written in: double3 (core/myVec.h:15-110), symMatrix (core/myMat.h:11-142)
// and quaternion (core/myQua.h:11-30), as plain host structs.  The host side
// never includes HIP headers; device work goes through include/dsph.h.
#pragma once

#include <cmath>

struct double3 {
    double x, y, z;
};

inline double3 make_double3(double x, double y, double z) { return double3{x, y, z}; }
inline double3 operator+(const double3& a, const double3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline double3 operator-(const double3& a, const double3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double3 operator-(const double3& a) { return {-a.x, -a.y, -a.z}; }
inline double3 operator*(double s, const double3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline double3 operator*(const double3& a, double s) { return {s * a.x, s * a.y, s * a.z}; }
inline double3 operator/(const double3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double3& operator+=(double3& a, const double3& b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
inline double dot(const double3& a, const double3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double3 cross(const double3& a, const double3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double lengthSquared(const double3& a) { return dot(a, a); }
inline double length(const double3& a) { return std::sqrt(dot(a, a)); }
// a zero vector is returned unchanged (core/myVec.h:92-96)
inline double3 normalize(const double3& a) {
    const double l = length(a);
    return l > 0 ? a / l : a;
}

inline double pi() { return 3.14159265358979323846; }

// symmetric 3x3 matrix: xx, yy, zz, xy, xz, yz
struct symMatrix {
    double xx, yy, zz, xy, xz, yz;
};
inline symMatrix make_symMatrix(double xx, double yy, double zz, double xy, double xz, double yz) {
    return symMatrix{xx, yy, zz, xy, xz, yz};
}
// inverse by cofactors (core/myMat.h:129-142); a singular matrix gives inf/nan
inline symMatrix inverse(const symMatrix& m) {
    const double det = 1.0 / (m.xx * (m.yy * m.zz - m.yz * m.yz) + m.xy * (m.yz * m.xz - m.xy * m.zz) +
                              m.xz * (m.xy * m.yz - m.yy * m.xz));
    return symMatrix{(m.yy * m.zz - m.yz * m.yz) * det, (m.xx * m.zz - m.xz * m.xz) * det,
                     (m.xx * m.yy - m.xy * m.xy) * det, -(m.xy * m.zz - m.xz * m.yz) * det,
                     (m.xy * m.yz - m.xz * m.yy) * det, -(m.xx * m.yz - m.xz * m.xy) * det};
}

struct quaternion {
    double q0, q1, q2, q3;
};
