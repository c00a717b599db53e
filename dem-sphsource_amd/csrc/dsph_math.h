// dsph_math.h — host/device vector algebra and the SPH / DEM formulas used by
// the HIP kernels.
//
// The formulas follow the reference solver's core/SPHKernelDEMContactModel.h,
// core/myVec.h, core/myQua.h and core/myMat.h (cited per function); the code
// is this project's own, written for CDNA4: explicit powers instead of pow(),
// reciprocal-square-root based normalisation in the FP32 path, and the
// precision of the fluid pair loops as a template parameter.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define DSPH_HD __host__ __device__ __forceinline__

namespace dsph {

constexpr double kPi = 3.1415926535897931;  // myVec.h:10-13

// ---------------------------------------------------------------- vectors
template <class T>
struct V3 {
    T x, y, z;
};
using d3 = V3<double>;
using f3 = V3<float>;

template <class T> DSPH_HD V3<T> mk3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> DSPH_HD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> DSPH_HD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> DSPH_HD V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> DSPH_HD V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> DSPH_HD V3<T> operator*(T s, V3<T> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> DSPH_HD V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> DSPH_HD V3<T>& operator+=(V3<T>& a, V3<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <class T> DSPH_HD V3<T>& operator-=(V3<T>& a, V3<T> b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
template <class T> DSPH_HD V3<T>& operator*=(V3<T>& a, T s) { a.x *= s; a.y *= s; a.z *= s; return a; }
template <class T> DSPH_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> DSPH_HD V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
DSPH_HD double vlen(d3 a) { return sqrt(dot(a, a)); }
DSPH_HD float vlen(f3 a) { return sqrtf(dot(a, a)); }
// normalize: zero vector returned unchanged (myVec.h:92-96)
DSPH_HD d3 unit(d3 a) { double l = vlen(a); return l > 0 ? a / l : a; }

DSPH_HD d3 ld3(const double* p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
DSPH_HD void st3(double* p, long i, d3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
DSPH_HD d3 xyz(double4 v) { return {v.x, v.y, v.z}; }
DSPH_HD f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
DSPH_HD f3 tof(d3 a) { return {(float)a.x, (float)a.y, (float)a.z}; }

// Rodrigues rotation, identity below 1e-10 rad (myVec.h:98-110)
DSPH_HD d3 rodrigues(d3 v, d3 ang) {
    double a = vlen(ang);
    if (a < 1.e-10) return v;
    d3 k = ang / a;
    double ca = cos(a), sa = sin(a);
    return v * ca + cross(k, v) * sa + k * (dot(k, v) * (1.0 - ca));
}

// ------------------------------------------------------------- quaternions
struct Quat { double q0, q1, q2, q3; };

// q <- normalize(q + 1/2 (0, w dt) (x) q), myQua.h:46-55
DSPH_HD Quat quat_rotate(Quat q, d3 w, double dt) {
    d3 v = w * dt;
    Quat r{q.q0 + 0.5 * (-q.q1 * v.x - q.q2 * v.y - q.q3 * v.z),
           q.q1 + 0.5 * (q.q0 * v.x + q.q3 * v.y - q.q2 * v.z),
           q.q2 + 0.5 * (-q.q3 * v.x + q.q0 * v.y + q.q1 * v.z),
           q.q3 + 0.5 * (q.q2 * v.x - q.q1 * v.y + q.q0 * v.z)};
    double l = sqrt(r.q0 * r.q0 + r.q1 * r.q1 + r.q2 * r.q2 + r.q3 * r.q3);
    if (l > 0) { r.q0 /= l; r.q1 /= l; r.q2 /= l; r.q3 /= l; }
    return r;
}

// symmetric 3x3 tensor: xx, yy, zz, xy, xz, yz (myMat.h:11-31)
struct Sym { double xx, yy, zz, xy, xz, yz; };
DSPH_HD d3 mul(const Sym& m, d3 v) {
    return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
            m.xz * v.x + m.yz * v.y + m.zz * v.z};
}
// world-frame inverse inertia R I^-1 R^T, R = rotation of q (myMat.h:102-127)
DSPH_HD Sym rotate_inv_inertia(Quat q, const Sym& I) {
    const double R[3][3] = {
        {1 - 2 * q.q2 * q.q2 - 2 * q.q3 * q.q3, 2 * q.q1 * q.q2 - 2 * q.q0 * q.q3, 2 * q.q1 * q.q3 + 2 * q.q0 * q.q2},
        {2 * q.q1 * q.q2 + 2 * q.q0 * q.q3, 1 - 2 * q.q1 * q.q1 - 2 * q.q3 * q.q3, 2 * q.q2 * q.q3 - 2 * q.q0 * q.q1},
        {2 * q.q1 * q.q3 - 2 * q.q0 * q.q2, 2 * q.q2 * q.q3 + 2 * q.q0 * q.q1, 1 - 2 * q.q1 * q.q1 - 2 * q.q2 * q.q2}};
    // rows of R times I
    double RI[3][3];
    for (int r = 0; r < 3; r++) {
        d3 row{R[r][0], R[r][1], R[r][2]};
        d3 t = mul(I, row);
        RI[r][0] = t.x; RI[r][1] = t.y; RI[r][2] = t.z;
    }
    auto m = [&](int a, int b) { return R[b][0] * RI[a][0] + R[b][1] * RI[a][1] + R[b][2] * RI[a][2]; };
    return Sym{m(0, 0), m(1, 1), m(2, 2), m(0, 1), m(0, 2), m(1, 2)};
}

// ------------------------------------------------------- Wendland C2 (3-D)
// W(r,h) = sigma (1 - q/2)^4 (1 + 2q), sigma = 21/(16 pi h^3), support 2h
// (SPHKernelDEMContactModel.h:5-13)
DSPH_HD double wendland(double r, double h) {
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double sigma = 21.0 / (16.0 * kPi * h * h * h);
    double t = 1.0 - 0.5 * q;
    double t2 = t * t;
    return sigma * (t2 * t2) * (1.0 + 2.0 * q);
}
// dW/dr / r: grad W = (dW/dr / r) r_ij, zero for r < 1e-10 or r >= 2h
// (SPHKernelDEMContactModel.h:15-32).  d/dq[(1-q/2)^4(1+2q)] = -5 q (1-q/2)^3.
DSPH_HD double grad_wendland_factor(double r, double h) {
    if (r < 1.e-10 || r >= 2.0 * h) return 0.0;
    double q = r / h;
    double sigma = 21.0 / (16.0 * kPi * h * h * h);
    double t = 1.0 - 0.5 * q;
    double dwdq = -2.0 * (t * t * t) * (1.0 + 2.0 * q) + 2.0 * (t * t) * (t * t);
    return sigma * dwdq / h / r;
}

// ------------------------------------------------------------ DEM springs
// sliding / rolling spring (SPHKernelDEMContactModel.h:170-193)
DSPH_HD d3 spring_tangential(d3 prev, d3 sv, d3 n, d3 fn, double mu, double k, double d, double dt) {
    d3 s{0, 0, 0};
    if (mu > 0) {
        d3 p1 = prev - dot(prev, n) * n;
        double a1 = vlen(p1);
        if (a1 > 1.e-10) p1 *= vlen(prev) / a1;
        s = p1 + sv * dt;
        d3 f = -k * s - d * sv;
        double af = vlen(f), an = vlen(fn);
        if (af > mu * an) {
            f *= mu * an / af;
            s = -(f + d * sv) / k;
        }
    }
    return s;
}
// torsion spring (SPHKernelDEMContactModel.h:195-212)
DSPH_HD d3 spring_torsion(d3 prev, d3 tv, d3 n, d3 fn, double mu, double k, double d, double dt) {
    d3 s{0, 0, 0};
    if (mu > 0) {
        s = dot(prev + tv * dt, n) * n;
        d3 f = -k * s - d * tv;
        double af = vlen(f), an = vlen(fn);
        if (af > mu * an) {
            f *= mu * an / af;
            s = -(f + d * tv) / k;
        }
    }
    return s;
}

struct Springs { d3 s, r, t; };

// contact force/torque once stiffness and damping are fixed; shared by the
// linear (SPHKernelDEMContactModel.h:214-268) and Hertz–Mindlin (:270-326)
// models
DSPH_HD void contact_law(d3& F, d3& T, Springs& sp, d3 v, d3 w, d3 n, double R, double dt,
                         double kn, double ks, double kr, double kt, double dn, double ds,
                         double dr, double dtor, double overlap, double mus, double mur,
                         double mut) {
    d3 vn = dot(v, n) * n;
    d3 fn = kn * overlap * n - vn * dn;
    d3 vs = v - vn;
    sp.s = spring_tangential(sp.s, vs, n, fn, mus, ks, ds, dt);
    d3 fs = -ks * sp.s - ds * vs;
    d3 vr = -R * cross(n, w);
    sp.r = spring_tangential(sp.r, vr, n, fn, mur, kr, dr, dt);
    d3 fr = -kr * sp.r - dr * vr;
    double D = 2 * R;
    d3 vt = D * dot(w, n) * n;
    sp.t = spring_torsion(sp.t, vt, n, fn, mut, kt, dtor, dt);
    d3 ft = -kt * sp.t - dtor * vt;
    F = fn + fs;
    T = R * cross(n, fr) + D * ft;
}

// Contact tables on device: field-major arrays of n_comb entries
struct ContactTables {
    int nmat, ncomb;
    const double* hertz;   // 8 * ncomb
    const double* linear;  // 11 * ncomb
    const double* bonded;  // 6 * ncomb
};

// material-pair index (deviceStructs.h:958-973); materials are validated on
// upload so the result is always in range
DSPH_HD int combined_index(int nmat, int a, int b) {
    int i = a < b ? a : b, j = a < b ? b : a;
    return (i * (2 * nmat - i + 1)) / 2 + j - i;
}

// One contact-law evaluation with the pair's model: linear when k_n > 1e-10,
// else Hertz–Mindlin with D = -ln e / sqrt(ln^2 e + pi^2)
// (integrate.cu:459-505)
DSPH_HD void contact_model(const ContactTables& tb, int c, d3& F, d3& T, Springs& sp, d3 v,
                           d3 w, d3 n, double overlap, double m, double R, double dt) {
    const int nc = tb.ncomb;
    if (overlap <= 0) {  // both models reset the springs when not touching
        sp.s = sp.r = sp.t = d3{0, 0, 0};
        return;
    }
    if (tb.linear[c] > 1.e-10) {
        const double* p = tb.linear;
        double kn = p[c], ks = p[nc + c], kr = p[2 * nc + c], kt = p[3 * nc + c];
        contact_law(F, T, sp, v, w, n, R, dt, kn, ks, kr, kt, 2. * p[4 * nc + c] * sqrt(m * kn),
                    2. * p[5 * nc + c] * sqrt(m * ks), 2. * p[6 * nc + c] * sqrt(m * kr),
                    2. * p[7 * nc + c] * sqrt(m * kt), overlap, p[8 * nc + c], p[9 * nc + c],
                    p[10 * nc + c]);
    } else {
        const double* p = tb.hertz;
        double lr = log(p[2 * nc + c]);
        double D = -lr / sqrt(lr * lr + kPi * kPi);
        double sRd = sqrt(R * overlap);
        double kn = 4. / 3. * p[c] * sRd;
        double ks = 8. * p[nc + c] * sRd;
        double kr = ks * p[3 * nc + c], kt = ks * p[4 * nc + c];
        double c56 = 2. * sqrt(5. / 6.) * D;
        contact_law(F, T, sp, v, w, n, R, dt, kn, ks, kr, kt, c56 * sqrt(m * kn),
                    c56 * sqrt(m * ks), c56 * sqrt(m * kr), c56 * sqrt(m * kt), overlap,
                    p[5 * nc + c], p[6 * nc + c], p[7 * nc + c]);
    }
}

// parallel bond increment and break test (SPHKernelDEMContactModel.h:115-168)
// p: gamma, E, kn/ks, sigma_t, C, mu (field-major table, entry c)
DSPH_HD int parallel_bond(double& Fn, double& Tt, d3& Fs, d3& Tb, d3 n0, d3 n, d3 v, d3 wa,
                          d3 wb, double ra, double rb, double dt, const double* p, int nc, int c) {
    d3 th1 = cross(n0, n);
    Fs = rodrigues(Fs, th1);
    Tb = rodrigues(Tb, th1);
    d3 th2 = dot(0.5 * (wa + wb) * dt, n) * n;
    Fs = rodrigues(Fs, th2);
    Tb = rodrigues(Tb, th2);
    double rmin = ra < rb ? ra : rb;
    double R = p[c] * rmin;
    double A = R * R * kPi;
    double I = R * R * R * R / 4. * kPi;
    double J = 2 * I;
    double knA = p[nc + c] / (ra + rb);
    double ksA = knA / p[2 * nc + c];
    d3 dn = dot(v, n) * n * dt;
    d3 dtan = v * dt - dn;
    Fn -= dot(dn * knA * A, n);
    Fs -= dtan * ksA * A;
    d3 wr = wa - wb;
    dn = dot(wr, n) * n * dt;
    dtan = wr * dt - dn;
    Tt -= dot(dn * ksA * J, n);
    Tb -= dtan * knA * I;
    double sig = -Fn / A + vlen(Tb) / I * R;
    double tau = vlen(Fs) / A + fabs(Tt) / J * R;
    double st = p[3 * nc + c], C = p[4 * nc + c], mu = p[5 * nc + c];
    if (st > 0 && sig > st) return 0;
    if (C > 0 && tau > C - mu * sig) return 0;
    return 1;
}

}  // namespace dsph
