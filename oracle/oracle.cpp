/*
 * oracle.cpp — FP64 OpenMP CPU oracle of the DEM–SPH per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ and by the benchmark's
 * CPU-baseline leg as the checker.  Nothing in the product links it.
 *
 * It restates, in plain FP64 C++ with no FMA contraction
 * (built with -ffp-contract=off), the algorithms of the reference solver
 * kaiqideng/DEM-SPHSource.  Each routine cites the reference file:line it
 * follows.  Reference behaviours that are races or quirks are resolved as
 * follows (flag names from include/dsph.h):
 *  - cell end of the last key is n (neighborSearch.cu:63-68, race-free);
 *  - the continuity update reads a snapshot of dRho (Jacobi), the
 *    reference updates in place while neighbours read (integrate.cu:160);
 *  - the wall loop returns on the first non-touching wall unless
 *    DSPH_FLAG_WALL_CONTINUE (integrate.cu:535-541);
 *  - clump orientation is frozen unless DSPH_FLAG_CLUMP_QUATERNION
 *    (integrate.cu:931-938, myQua.h:46-55);
 *  - solid–solid spring history carries over from the previous step's list
 *    (neighborSearch.cu:287-301), without the capacity-growth reset;
 *  - with no fluid the normal-estimation smoothing length is 0
 *    (dataBase.cpp:359 reads max_element of an empty vector);
 *  - DSPH_FLAG_ART_VISCOSITY adds the artificial viscosity of
 *    SPHKernelDEMContactModel.h:74-77 (a helper never called upstream) in
 *    the dissipative orientation, with the updated particle's alpha.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ---------------- vector algebra: myVec.h:10-110 ------------------------ */
struct V3 { double x, y, z; };
inline V3 mk(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(double s, V3 a) { return mk(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline V3& operator*=(V3& a, double s) { a.x *= s; a.y *= s; a.z *= s; return a; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double len(V3 a) { return std::sqrt(dot(a, a)); }
/* myVec.h:92-96: zero vector returned unchanged */
inline V3 unit(V3 a) { return len(a) > 0 ? a / std::sqrt(dot(a, a)) : a; }
const double kPi = 3.1415926535897931; /* myVec.h:10-13 */

/* Rodrigues rotation, identity below 1e-10 rad (myVec.h:98-110) */
V3 rodrigues(V3 v, V3 ang) {
    double a = len(ang);
    if (a < 1.e-10) return v;
    V3 k = ang / a;
    return v * std::cos(a) + cross(k, v) * std::sin(a) +
           k * (dot(k, v) * (1 - std::cos(a)));
}

struct Quat { double q0, q1, q2, q3; };
/* myQua.h:46-55: q <- normalize(q + 1/2 (0,w dt) (x) q) */
Quat quat_rotate(Quat q, V3 w, double dt) {
    V3 v = w * dt;
    Quat d{0.5 * (-q.q1 * v.x - q.q2 * v.y - q.q3 * v.z),
           0.5 * (q.q0 * v.x + q.q3 * v.y - q.q2 * v.z),
           0.5 * (-q.q3 * v.x + q.q0 * v.y + q.q1 * v.z),
           0.5 * (q.q2 * v.x - q.q1 * v.y + q.q0 * v.z)};
    Quat r{q.q0 + d.q0, q.q1 + d.q1, q.q2 + d.q2, q.q3 + d.q3};
    double l = std::sqrt(r.q0 * r.q0 + r.q1 * r.q1 + r.q2 * r.q2 + r.q3 * r.q3);
    if (l > 0) { r.q0 /= l; r.q1 /= l; r.q2 /= l; r.q3 /= l; }
    return r;
}

struct Sym { double xx, yy, zz, xy, xz, yz; };
inline V3 mul(const Sym& m, V3 v) {
    return mk(m.xx * v.x + m.xy * v.y + m.xz * v.z,
              m.xy * v.x + m.yy * v.y + m.yz * v.z,
              m.xz * v.x + m.yz * v.y + m.zz * v.z);
}
/* quadratic form r^T I r of one row r=(a,b,c) of R (myMat.h:110-121) */
double invIq(double a, double b, double c, const Sym& I) {
    return I.xx * a * a + 2 * I.xy * a * b + 2 * I.xz * a * c + I.yy * b * b +
           2 * I.yz * b * c + I.zz * c * c;
}
/* R invI R^T with R the rotation matrix of q (myMat.h:102-127) */
Sym rotate_inv_inertia(Quat q, const Sym& I) {
    double a = 1 - 2 * q.q2 * q.q2 - 2 * q.q3 * q.q3;
    double b = 2 * q.q1 * q.q2 - 2 * q.q0 * q.q3;
    double c = 2 * q.q1 * q.q3 + 2 * q.q0 * q.q2;
    double d = 2 * q.q1 * q.q2 + 2 * q.q0 * q.q3;
    double e = 1 - 2 * q.q1 * q.q1 - 2 * q.q3 * q.q3;
    double f = 2 * q.q2 * q.q3 - 2 * q.q0 * q.q1;
    double g = 2 * q.q1 * q.q3 - 2 * q.q0 * q.q2;
    double h = 2 * q.q2 * q.q3 + 2 * q.q0 * q.q1;
    double i = 1 - 2 * q.q1 * q.q1 - 2 * q.q2 * q.q2;
    /* rows of R: (a b c), (d e f), (g h i); M = R I R^T */
    double ra[3] = {I.xx * a + I.xy * b + I.xz * c, I.xy * a + I.yy * b + I.yz * c,
                    I.xz * a + I.yz * b + I.zz * c};
    double rd[3] = {I.xx * d + I.xy * e + I.xz * f, I.xy * d + I.yy * e + I.yz * f,
                    I.xz * d + I.yz * e + I.zz * f};
    Sym m;
    m.xx = invIq(a, b, c, I);
    m.xy = d * ra[0] + e * ra[1] + f * ra[2];
    m.xz = g * ra[0] + h * ra[1] + i * ra[2];
    m.yy = invIq(d, e, f, I);
    m.yz = g * rd[0] + h * rd[1] + i * rd[2];
    m.zz = invIq(g, h, i, I);
    return m;
}

/* ---------------- SPH kernel: SPHKernelDEMContactModel.h:5-32 ---------- */
double wendland(double r, double h) {
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double sigma = 21.0 / (16.0 * kPi * h * h * h);
    double t = 1.0 - 0.5 * q;
    return sigma * std::pow(t, 4) * (1.0 + 2.0 * q);
}
V3 grad_wendland(V3 rij, double h) {
    double r = len(rij);
    if (r < 1.e-10 || r >= 2.0 * h) return mk(0, 0, 0);
    double q = r / h;
    double sigma = 21.0 / (16.0 * kPi * h * h * h);
    double t = 1.0 - 0.5 * q;
    double dwdq = -2.0 * std::pow(t, 3) * (1.0 + 2.0 * q) + 2.0 * std::pow(t, 4);
    double dwdr = sigma * dwdq / h;
    return (dwdr / r) * rij;
}

/* ---------------- DEM contact laws: SPHKernelDEMContactModel.h:170-326 - */
/* sliding/rolling spring: project on the tangent plane keeping the norm,
 * add v dt, cap at mu |Fn| with the spring back-solved (:170-193) */
V3 spring_tangential(V3 prev, V3 sv, V3 n, V3 fn, double mu, double k, double d,
                     double dt) {
    V3 s = mk(0, 0, 0);
    if (mu > 0) {
        V3 p1 = prev - dot(prev, n) * n;
        double a1 = len(p1);
        if (a1 > 1.e-10) p1 *= len(prev) / a1;
        s = p1 + sv * dt;
        V3 f = -k * s - d * sv;
        double af = len(f), an = len(fn);
        if (af > mu * an) {
            double ratio = mu * an / af;
            f *= ratio;
            s = -(f + d * sv) / k;
        }
    }
    return s;
}
/* torsion spring (:195-212) */
V3 spring_torsion(V3 prev, V3 tv, V3 n, V3 fn, double mu, double k, double d,
                  double dt) {
    V3 s = mk(0, 0, 0);
    if (mu > 0) {
        s = dot(prev + tv * dt, n) * n;
        V3 f = -k * s - d * tv;
        double af = len(f), an = len(fn);
        if (af > mu * an) {
            double ratio = mu * an / af;
            f *= ratio;
            s = -(f + d * tv) / k;
        }
    }
    return s;
}

struct Springs { V3 s, r, t; };

/* shared tail of LinearContact (:214-268) and HertzianMindlinContact
 * (:270-326) once stiffness and damping are known */
void contact_tail(V3& F, V3& T, Springs& sp, V3 v, V3 w, V3 n, double R,
                  double dt, double kn, double ks, double kr, double kt,
                  double dn, double ds, double dr, double dtor, double overlap,
                  double mus, double mur, double mut) {
    V3 vn = dot(v, n) * n;
    V3 fn = kn * overlap * n - vn * dn;
    V3 vs = v - vn;
    sp.s = spring_tangential(sp.s, vs, n, fn, mus, ks, ds, dt);
    V3 fs = -ks * sp.s - ds * vs;
    V3 vr = -R * cross(n, w);
    sp.r = spring_tangential(sp.r, vr, n, fn, mur, kr, dr, dt);
    V3 fr = -kr * sp.r - dr * vr;
    V3 tr = R * cross(n, fr);
    double D = 2 * R;
    V3 vt = D * dot(w, n) * n;
    sp.t = spring_torsion(sp.t, vt, n, fn, mut, kt, dtor, dt);
    V3 ft = -kt * sp.t - dtor * vt;
    V3 tt = D * ft;
    F = fn + fs;
    T = tr + tt;
}

void linear_contact(V3& F, V3& T, Springs& sp, V3 v, V3 w, V3 n, double overlap,
                    double m, double R, double dt, const double* p) {
    if (overlap > 0) {
        double kn = p[0], ks = p[1], kr = p[2], kt = p[3];
        double dn = 2. * p[4] * std::sqrt(m * kn);
        double ds = 2. * p[5] * std::sqrt(m * ks);
        double dr = 2. * p[6] * std::sqrt(m * kr);
        double dtt = 2. * p[7] * std::sqrt(m * kt);
        contact_tail(F, T, sp, v, w, n, R, dt, kn, ks, kr, kt, dn, ds, dr, dtt,
                     overlap, p[8], p[9], p[10]);
    } else {
        sp.s = sp.r = sp.t = mk(0, 0, 0);
    }
}

/* p = E*, G*, D, kr/ks, kt/ks, mu_s, mu_r, mu_t */
void hertz_contact(V3& F, V3& T, Springs& sp, V3 v, V3 w, V3 n, double overlap,
                   double m, double R, double dt, const double* p) {
    if (overlap > 0) {
        double kn = 4. / 3. * p[0] * std::sqrt(R * overlap);
        double ks = 8. * p[1] * std::sqrt(R * overlap);
        double c56 = 2. * std::sqrt(5. / 6.) * p[2];
        double dn = c56 * std::sqrt(m * kn);
        double ds = c56 * std::sqrt(m * ks);
        double kr = ks * p[3], kt = ks * p[4];
        double dr = c56 * std::sqrt(m * kr);
        double dtt = c56 * std::sqrt(m * kt);
        contact_tail(F, T, sp, v, w, n, R, dt, kn, ks, kr, kt, dn, ds, dr, dtt,
                     overlap, p[5], p[6], p[7]);
    } else {
        sp.s = sp.r = sp.t = mk(0, 0, 0);
    }
}

/* parallel bond increment + break test (SPHKernelDEMContactModel.h:115-168)
 * p = gamma, E, kn/ks, sigma_t, C, mu */
int parallel_bond(double& Fn, double& Tt, V3& Fs, V3& Tb, V3 n0, V3 n, V3 v,
                  V3 wa, V3 wb, double ra, double rb, double dt, const double* p) {
    V3 th1 = cross(n0, n);
    Fs = rodrigues(Fs, th1);
    Tb = rodrigues(Tb, th1);
    V3 th2 = dot(0.5 * (wa + wb) * dt, n) * n;
    Fs = rodrigues(Fs, th2);
    Tb = rodrigues(Tb, th2);
    double rmin = ra < rb ? ra : rb;
    double R = p[0] * rmin;
    double A = R * R * kPi;
    double I = R * R * R * R / 4. * kPi;
    double J = 2 * I;
    double knA = p[1] / (ra + rb);
    double ksA = knA / p[2];
    V3 dn = dot(v, n) * n * dt;
    V3 dtan = v * dt - dn;
    Fn -= dot(dn * knA * A, n);
    Fs -= dtan * ksA * A;
    V3 wr = wa - wb;
    dn = dot(wr, n) * n * dt;
    dtan = wr * dt - dn;
    Tt -= dot(dn * ksA * J, n);
    Tb -= dtan * knA * I;
    double sig = -Fn / A + len(Tb) / I * R;
    double tau = len(Fs) / A + std::fabs(Tt) / J * R;
    int bonded = 1;
    if (p[3] > 0 && sig > p[3]) bonded = 0;
    else if (p[4] > 0 && tau > p[4] - p[5] * sig) bonded = 0;
    return bonded;
}

/* deviceStructs.h:958-973 */
int combined_index(int nmat, int a, int b) {
    if (a >= nmat || b >= nmat) return -1;
    int i = a, j = b;
    if (a > b) { i = b; j = a; }
    return (i * (2 * nmat - i + 1)) / 2 + j - i;
}

/* stable counting sort of keys in [0,nkeys): perm[slot] = original index
 * (thrust::sort_by_key on int keys with values 0..n-1 is a stable radix
 * sort, neighborSearch.cu:3-9, 38-44) */
void stable_sort_keys(const std::vector<int>& key, int nkeys, std::vector<int>& perm,
                      std::vector<int>& sorted) {
    int n = (int)key.size();
    std::vector<int> cnt(nkeys + 1, 0);
    for (int i = 0; i < n; i++) cnt[key[i] + 1]++;
    for (int k = 0; k < nkeys; k++) cnt[k + 1] += cnt[k];
    perm.assign(n, -1);
    sorted.assign(n, -1);
    for (int i = 0; i < n; i++) {
        int p = cnt[key[i]]++;
        perm[p] = i;
        sorted[p] = key[i];
    }
}
/* findStartAndEnd with end[last] = n (neighborSearch.cu:56-69) */
void spans(const std::vector<int>& sorted, std::vector<int>& start, std::vector<int>& end) {
    int n = (int)sorted.size();
    for (int i = 0; i < n; i++) {
        if (i == 0 || sorted[i] != sorted[i - 1]) start[sorted[i]] = i;
        if (i == n - 1 || sorted[i] != sorted[i + 1]) end[sorted[i]] = i + 1;
    }
}

struct PairList {              /* interactionBase (deviceStructs.h:444-510) */
    std::vector<int> count, prefix; /* per pointed particle, inclusive scan */
    std::vector<int> pointed, pointing;
    std::vector<int> tindex;        /* pair ids sorted stably by pointing  */
    std::vector<int> rstart, rend;  /* per pointing particle range in tindex */
    int num() const { return (int)pointed.size(); }
    int seg_begin(int i) const { return i > 0 ? prefix[i - 1] : 0; }
    int seg_end(int i) const { return prefix[i]; }
    void transpose(int n_pointing) {
        std::vector<int> sk;
        stable_sort_keys(pointing, n_pointing, tindex, sk);
        rstart.assign(n_pointing, -1);
        rend.assign(n_pointing, -1);
        spans(sk, rstart, rend);
    }
};

} // namespace

struct orc_ctx {
    uint32_t flags = 0;
    int threads = 0;
    bool initialized = false;
    V3 g = mk(0, 0, -9.81);
    /* spatialGrid (deviceStructs.h:976-1014) */
    V3 minB, maxB, cell;
    int gx = 1, gy = 1, gz = 1, ncell = 1;
    std::vector<int> fstart, fend, sstart, send;
    /* fluid */
    int nf = 0;
    std::vector<V3> fpos, fvel, facc;
    std::vector<double> fh, rho0, drho, c, nu, sumw0, alpha;
    std::vector<int> fhash, fperm, fsorted;
    /* solid */
    int ns = 0;
    std::vector<V3> spos, svel, sacc, sw, storq, snorm;
    std::vector<double> srad, seff, sinvm;
    std::vector<int> smat, sclus, sclump, sdummy;
    std::vector<int> shash, sperm, ssorted;
    /* clumps */
    int nc = 0;
    std::vector<V3> cpos, cvel, cacc, cw, ctorq;
    std::vector<Quat> cq;
    std::vector<Sym> cinvI;
    std::vector<double> cinvm;
    std::vector<int> cstart, cend;
    /* infinite walls */
    int nw = 0;
    std::vector<V3> wpt, wn;
    std::vector<int> wmat;
    std::vector<Springs> wallHist; /* [ns * nw] */
    /* cylinder walls (hostStructs.h:176-200) */
    int ny = 0;
    std::vector<V3> ypt, yax, yw;
    std::vector<double> yrad;
    std::vector<int> ymat;
    std::vector<Springs> cylHist; /* [ns * ny] */
    /* contact tables */
    int nmat = 0;
    std::vector<double> hertz, linear, bonded; /* field-major */
    int ncomb = 0;
    /* lists */
    PairList ff, fs, ss;
    std::vector<V3> fsForce;
    std::vector<V3> ssForce, ssTorque;
    std::vector<Springs> ssSpring;
    /* bonds (interactionBonded, deviceStructs.h:714-789) */
    std::vector<V3> bN, bPt, bFs, bTb;
    std::vector<double> bFn, bTt;
    std::vector<int> bI, bJ, bOn;
};

namespace {

inline double hz(const orc_ctx* o, int field, int c) { return o->hertz[field * o->ncomb + c]; }

V3 ld3(const double* p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
void st3(double* p, int i, V3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

/* setSpatialGrids (dataBase.cpp:415-432) */
void set_grid(orc_ctx* o, V3 origin, V3 size) {
    double mf = 0, ms = 0;
    if (o->nf > 0) mf = *std::max_element(o->fh.begin(), o->fh.end());
    if (o->ns > 0) ms = *std::max_element(o->seff.begin(), o->seff.end());
    double cs = 2 * std::max(mf, ms);
    o->minB = origin;
    o->maxB = origin + size;
    o->gx = size.x > cs ? int(size.x / cs) : 1;
    o->gy = size.y > cs ? int(size.y / cs) : 1;
    o->gz = size.z > cs ? int(size.z / cs) : 1;
    o->cell = mk(size.x / double(o->gx), size.y / double(o->gy), size.z / double(o->gz));
    o->ncell = o->gx * o->gy * o->gz + 1;
}

/* calculateGridPosition (neighborSearch.h:44-49) */
inline void grid_pos(const orc_ctx* o, V3 p, int& x, int& y, int& z) {
    x = int((p.x - o->minB.x) / o->cell.x);
    y = int((p.y - o->minB.y) / o->cell.y);
    z = int((p.z - o->minB.z) / o->cell.z);
}
/* calculateHash (neighborSearch.h:51-54) */
inline int lin_hash(const orc_ctx* o, int x, int y, int z) {
    return z * o->gy * o->gx + y * o->gx + x;
}
/* calculatePointHash (neighborSearch.cu:19-36); a point within rounding of
 * the upper bound may truncate to index G on an axis (the reference then
 * indexes past its cell arrays): such a hash goes to the spare cell */
int point_hash(const orc_ctx* o, V3 p) {
    if (o->minB.x <= p.x && p.x < o->maxB.x && o->minB.y <= p.y && p.y < o->maxB.y &&
        o->minB.z <= p.z && p.z < o->maxB.z) {
        int x, y, z;
        grid_pos(o, p, x, y, z);
        int h = lin_hash(o, x, y, z);
        return (h >= 0 && h < o->ncell - 1) ? h : o->ncell - 1;
    }
    return o->ncell - 1;
}

/* updateGridCellStartEnd (neighborSearch.cu:89-113) */
void update_cells(orc_ctx* o) {
    o->fstart.assign(o->ncell, -1); o->fend.assign(o->ncell, -1);
    o->sstart.assign(o->ncell, -1); o->send.assign(o->ncell, -1);
    o->fhash.resize(o->nf);
    o->shash.resize(o->ns);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->nf; i++) o->fhash[i] = point_hash(o, o->fpos[i]);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->ns; i++) o->shash[i] = point_hash(o, o->spos[i]);
    stable_sort_keys(o->fhash, o->ncell, o->fperm, o->fsorted);
    spans(o->fsorted, o->fstart, o->fend);
    stable_sort_keys(o->shash, o->ncell, o->sperm, o->ssorted);
    spans(o->ssorted, o->sstart, o->send);
}

/* the 27-cell stencil in z,y,x order over linear hashes, skipping hashes
 * outside [0, num) (neighborSearch.cu:135-146) */
template <class F>
void for_stencil(const orc_ctx* o, V3 p, F&& body) {
    int ax, ay, az;
    grid_pos(o, p, ax, ay, az);
    for (int zz = -1; zz <= 1; zz++)
        for (int yy = -1; yy <= 1; yy++)
            for (int xx = -1; xx <= 1; xx++) {
                int h = lin_hash(o, ax + xx, ay + yy, az + zz);
                if (h < 0 || h >= o->ncell) continue;
                body(h);
            }
}

/* setFluidInteractionsKernel (neighborSearch.cu:115-220) as one pass per
 * particle appending in loop order (pass 1 order == pass 0 order) */
void build_fluid_lists(orc_ctx* o) {
    int n = o->nf;
    std::vector<std::vector<int>> lf(n), ls(n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int a = 0; a < n; a++) {
        V3 pa = o->fpos[a];
        double ra = o->fh[a];
        for_stencil(o, pa, [&](int h) {
            int st = o->fstart[h], en = o->fend[h];
            if (st != -1)
                for (int t = st; t < en; t++) {
                    int b = o->fperm[t];
                    if (a >= b) continue;
                    double rad = std::fmax(ra, o->fh[b]);
                    double dis = len(pa - o->fpos[b]);
                    if (2 * rad - dis >= 0.) lf[a].push_back(b);
                }
            st = o->sstart[h]; en = o->send[h];
            if (st == -1) return;
            for (int t = st; t < en; t++) {
                int b = o->sperm[t];
                double dis = len(pa - o->spos[b]);
                if (2 * ra - dis >= 0.) ls[a].push_back(b);
            }
        });
    }
    auto pack = [&](PairList& L, std::vector<std::vector<int>>& src, int npointing) {
        L.count.assign(n, 0); L.prefix.assign(n, 0);
        long tot = 0;
        for (int a = 0; a < n; a++) { L.count[a] = (int)src[a].size(); tot += L.count[a]; L.prefix[a] = (int)tot; }
        L.pointed.resize(tot); L.pointing.resize(tot);
        for (int a = 0; a < n; a++) {
            int b0 = L.seg_begin(a);
            for (size_t k = 0; k < src[a].size(); k++) { L.pointed[b0 + k] = a; L.pointing[b0 + k] = src[a][k]; }
        }
        L.transpose(npointing);
    };
    pack(o->ff, lf, n);
    pack(o->fs, ls, o->ns);
    o->fsForce.assign(o->fs.num(), mk(0, 0, 0));
}

/* setSolid2SolidInteractionsKernel with history carry-over
 * (neighborSearch.cu:222-312; interactionSolid2Solid::save,
 * deviceStructs.h:693-706) */
void build_solid_lists(orc_ctx* o) {
    int n = o->ns;
    /* history = previous list, its springs and its transpose */
    PairList prev = o->ss;
    std::vector<Springs> prevSp = o->ssSpring;
    std::vector<std::vector<int>> lst(n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int a = 0; a < n; a++) {
        if (o->sdummy[a] == 1) continue;
        V3 pa = o->spos[a];
        double ra = o->seff[a];
        for_stencil(o, pa, [&](int h) {
            int st = o->sstart[h], en = o->send[h];
            if (st == -1) return;
            for (int t = st; t < en; t++) {
                int b = o->sperm[t];
                if (o->sdummy[b] == 1) continue;
                if (a >= b) continue;
                if (o->sclump[a] >= 0 && o->sclump[a] == o->sclump[b]) continue;
                if (o->sinvm[a] < 1.e-10 && o->sinvm[b] < 1.e-10) continue;
                double dis = len(pa - o->spos[b]);
                if (o->seff[b] + ra - dis >= 0.) lst[a].push_back(b);
            }
        });
    }
    PairList& L = o->ss;
    L.count.assign(n, 0); L.prefix.assign(n, 0);
    long tot = 0;
    for (int a = 0; a < n; a++) { L.count[a] = (int)lst[a].size(); tot += L.count[a]; L.prefix[a] = (int)tot; }
    L.pointed.resize(tot); L.pointing.resize(tot);
    o->ssForce.assign(tot, mk(0, 0, 0));
    o->ssTorque.assign(tot, mk(0, 0, 0));
    o->ssSpring.assign(tot, Springs{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)});
    bool havePrev = prev.num() > 0 && (int)prev.rstart.size() == n;
#pragma omp parallel for schedule(dynamic, 64)
    for (int a = 0; a < n; a++) {
        int b0 = L.seg_begin(a);
        for (size_t k = 0; k < lst[a].size(); k++) {
            int b = lst[a][k];
            L.pointed[b0 + k] = a;
            L.pointing[b0 + k] = b;
            if (!havePrev || prev.rstart[b] == -1) continue;
            for (int j = prev.rstart[b]; j < prev.rend[b]; j++) {
                int j1 = prev.tindex[j];
                if (prev.pointed[j1] == a) { o->ssSpring[b0 + k] = prevSp[j1]; break; }
            }
        }
    }
    L.transpose(n);
}

} // namespace

/* =========================== C API ====================================== */
extern "C" {

orc_ctx* orc_create(void) { return new orc_ctx(); }
void orc_destroy(orc_ctx* o) { delete o; }

int orc_set_flags(orc_ctx* o, uint32_t flags) { o->flags = flags; return 0; }
int orc_set_threads(orc_ctx* o, int n) {
    o->threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
    return 0;
}

int orc_upload_scene(orc_ctx* o, const dsph_scene* s) {
    if (!s || s->n_fluid < 0 || s->n_solid < 0 || s->n_clump < 0 || s->n_wall < 0) return DSPH_E_INVALID;
    o->g = mk(s->gravity[0], s->gravity[1], s->gravity[2]);
    int n = o->nf = s->n_fluid;
    o->fpos.resize(n); o->fvel.resize(n); o->facc.assign(n, mk(0, 0, 0));
    o->fh.resize(n); o->rho0.resize(n); o->drho.assign(n, 0); o->c.resize(n); o->nu.resize(n);
    o->alpha.assign(n, 0);
    o->sumw0.assign(n, 0);
    for (int i = 0; i < n; i++) {
        o->fpos[i] = ld3(s->fluid_pos, i);
        o->fvel[i] = s->fluid_vel ? ld3(s->fluid_vel, i) : mk(0, 0, 0);
        o->fh[i] = s->fluid_h[i];
        o->rho0[i] = s->fluid_rho0[i];
        if (s->fluid_drho) o->drho[i] = s->fluid_drho[i];
        o->c[i] = s->fluid_c[i];
        o->nu[i] = s->fluid_nu[i];
        if (s->fluid_alpha) o->alpha[i] = s->fluid_alpha[i];
    }
    n = o->ns = s->n_solid;
    o->spos.resize(n); o->svel.resize(n); o->sacc.assign(n, mk(0, 0, 0)); o->sw.resize(n);
    o->storq.assign(n, mk(0, 0, 0)); o->snorm.assign(n, mk(0, 0, 0));
    o->srad.resize(n); o->seff.resize(n); o->sinvm.resize(n);
    o->smat.resize(n); o->sclus.resize(n); o->sclump.resize(n); o->sdummy.resize(n);
    for (int i = 0; i < n; i++) {
        o->spos[i] = ld3(s->solid_pos, i);
        o->svel[i] = s->solid_vel ? ld3(s->solid_vel, i) : mk(0, 0, 0);
        o->sw[i] = s->solid_angvel ? ld3(s->solid_angvel, i) : mk(0, 0, 0);
        if (s->solid_normal) o->snorm[i] = ld3(s->solid_normal, i);
        o->srad[i] = s->solid_radius[i];
        o->seff[i] = s->solid_eff_radius[i];
        o->sinvm[i] = s->solid_inv_mass[i];
        o->smat[i] = s->solid_material ? s->solid_material[i] : 0;
        o->sclus[i] = s->solid_cluster ? s->solid_cluster[i] : -1;
        o->sclump[i] = s->solid_clump ? s->solid_clump[i] : -1;
        o->sdummy[i] = s->solid_dummy ? s->solid_dummy[i] : 0;
    }
    n = o->nc = s->n_clump;
    o->cpos.resize(n); o->cvel.resize(n); o->cacc.assign(n, mk(0, 0, 0)); o->cw.resize(n);
    o->ctorq.assign(n, mk(0, 0, 0)); o->cq.resize(n); o->cinvI.resize(n); o->cinvm.resize(n);
    o->cstart.resize(n); o->cend.resize(n);
    for (int i = 0; i < n; i++) {
        o->cpos[i] = ld3(s->clump_centroid, i);
        o->cvel[i] = s->clump_vel ? ld3(s->clump_vel, i) : mk(0, 0, 0);
        o->cw[i] = s->clump_angvel ? ld3(s->clump_angvel, i) : mk(0, 0, 0);
        if (s->clump_orientation) {
            const double* q = s->clump_orientation + 4 * i;
            o->cq[i] = Quat{q[0], q[1], q[2], q[3]};
        } else {
            o->cq[i] = Quat{1, 0, 0, 0};
        }
        const double* I = s->clump_inv_inertia + 6 * i;
        o->cinvI[i] = Sym{I[0], I[1], I[2], I[3], I[4], I[5]};
        o->cinvm[i] = s->clump_inv_mass[i];
        o->cstart[i] = s->clump_pebble_start[i];
        o->cend[i] = s->clump_pebble_end[i];
        if (o->cstart[i] < 0 || o->cend[i] > o->ns || o->cstart[i] > o->cend[i]) return DSPH_E_INVALID;
    }
    n = o->nw = s->n_wall;
    o->wpt.resize(n); o->wn.resize(n); o->wmat.resize(n);
    for (int i = 0; i < n; i++) {
        o->wpt[i] = ld3(s->wall_point, i);
        o->wn[i] = ld3(s->wall_normal, i);
        o->wmat[i] = s->wall_material ? s->wall_material[i] : 0;
    }
    o->wallHist.assign((size_t)o->ns * o->nw, Springs{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)});
    n = o->ny = s->n_cylinder;
    o->ypt.resize(n); o->yax.resize(n); o->yw.resize(n); o->yrad.resize(n); o->ymat.resize(n);
    for (int i = 0; i < n; i++) {
        o->ypt[i] = ld3(s->cylinder_point, i);
        o->yax[i] = unit(ld3(s->cylinder_axis, i));
        o->yw[i] = s->cylinder_angvel ? ld3(s->cylinder_angvel, i) : mk(0, 0, 0);
        o->yrad[i] = s->cylinder_radius[i];
        o->ymat[i] = s->cylinder_material ? s->cylinder_material[i] : 0;
    }
    o->cylHist.assign((size_t)o->ns * o->ny, Springs{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)});
    o->nmat = s->n_material;
    o->ncomb = o->nmat * (o->nmat + 1) / 2;
    o->hertz.assign(8 * o->ncomb, 0);
    o->linear.assign(11 * o->ncomb, 0);
    o->bonded.assign(6 * o->ncomb, 0);
    if (s->hertz) std::copy(s->hertz, s->hertz + 8 * o->ncomb, o->hertz.begin());
    if (s->linear) std::copy(s->linear, s->linear + 11 * o->ncomb, o->linear.begin());
    if (s->bonded) std::copy(s->bonded, s->bonded + 6 * o->ncomb, o->bonded.begin());
    /* invalid material pairs would index out of the tables (the reference reads
       out of bounds there, deviceStructs.h:960-963): rejected */
    for (int i = 0; i < o->ns; i++)
        if (!o->sdummy[i] && (o->smat[i] < 0 || o->smat[i] >= o->nmat)) return DSPH_E_INVALID;
    for (int i = 0; i < o->nw; i++)
        if (o->wmat[i] < 0 || o->wmat[i] >= o->nmat) return DSPH_E_INVALID;
    set_grid(o, ld3(s->domain_origin, 0), ld3(s->domain_size, 0));
    o->ff = PairList(); o->fs = PairList(); o->ss = PairList();
    o->ssSpring.clear(); o->ssForce.clear(); o->ssTorque.clear();
    o->bN.clear(); o->bPt.clear(); o->bFs.clear(); o->bTb.clear(); o->bFn.clear(); o->bTt.clear();
    o->bI.clear(); o->bJ.clear(); o->bOn.clear();
    o->initialized = false;
    return 0;
}

/* neighborSearch (neighborSearch.cu:314-381) */
int orc_neighbor_search(orc_ctx* o, int32_t i_step, int32_t gap) {
    update_cells(o);
    if (i_step % gap == 0 && o->nf > 0) build_fluid_lists(o);
    if (o->ns > 0) build_solid_lists(o);
    return 0;
}

/* addBondData, initial branch (dataBase.cpp:270-308) */
static void add_bonds(orc_ctx* o) {
    const PairList& L = o->ss;
    for (int k = 0; k < L.num(); k++) {
        int i = L.pointed[k], j = L.pointing[k];
        V3 pi = o->spos[i], pj = o->spos[j];
        double ri = o->srad[i], rj = o->srad[j];
        if (ri + rj + 1.e-10 <= len(pi - pj)) continue;
        int cij = combined_index(o->nmat, o->smat[i], o->smat[j]);
        double Eb = o->bonded[1 * o->ncomb + cij];
        if (Eb > 1.e-10 && o->sclus[i] >= 0 && o->sclus[i] == o->sclus[j]) {
            double d = len(pi - pj);
            V3 n = (pi - pj) / d;
            double delta = ri + rj - d;
            o->bN.push_back(n);
            o->bPt.push_back(pj + (rj - 0.5 * delta) * n);
            o->bFs.push_back(mk(0, 0, 0)); o->bTb.push_back(mk(0, 0, 0));
            o->bFn.push_back(0); o->bTt.push_back(0);
            o->bI.push_back(i); o->bJ.push_back(j); o->bOn.push_back(1);
        }
    }
}

/* setSolidNormal (dataBase.cpp:348-413) */
static void solid_normals(orc_ctx* o) {
    double h = 0;
    if (o->nf > 0) h = *std::max_element(o->fh.begin(), o->fh.end());
#pragma omp parallel for schedule(dynamic, 256)
    for (int a = 0; a < o->ns; a++) {
        V3 th = mk(0, 0, 0);
        int ca = o->sclump[a];
        double ima = o->sinvm[a];
        V3 pa = o->spos[a];
        for_stencil(o, pa, [&](int hh) {
            int st = o->sstart[hh], en = o->send[hh];
            if (st == -1) return;
            for (int t = st; t < en; t++) {
                int b = o->sperm[t];
                V3 r = pa - o->spos[b];
                if (len(r) > 2 * h) continue;
                if (ima <= 1.e-10 && o->sinvm[b] <= 1.e-10) {
                    th += -std::pow(h / 1.3, 3) * grad_wendland(r, h);
                } else if (ca >= 0 && ca == o->sclump[b]) {
                    th += -std::pow(2. * o->srad[b], 3.) * grad_wendland(r, h);
                }
            }
        });
        o->snorm[a] = len(th) > 1.e-10 ? unit(th) : mk(0, 0, 0);
    }
}

int orc_initialize(orc_ctx* o, int compute_normals) {
    orc_neighbor_search(o, 0, 1);
    add_bonds(o);
    if (compute_normals) solid_normals(o);
    o->initialized = true;
    return 0;
}

/* dataBase::addFluid / addSolid / addCluster / addClump called after the run
 * started (dataBase.cpp:113-117, 226-268, 270-346).  The particles of e go
 * after the existing ones of each species, whose state stays as it is (the
 * reference's uploadState + insertData + copy); e's clump pebble ranges and
 * solid clump indices are relative to e.  Then the search of step 0, which
 * carries the contact springs over by (i, j) (neighborSearch(dev, 0, 1)),
 * and, when solids were added, addBondData's later branch (:309-343): a
 * contact of the new list becomes a bond when its pointed solid comes after
 * the last bonded one and the initial-bond conditions hold.  No normals are
 * estimated (setSolidNormal runs only at initialisation); a new solid's
 * normal is e's.  Resolved differently from the reference:
 *  - sumW0 of the existing fluid is kept, and a new particle's is its
 *    kernel sum just after the append (the reference reallocates sumW0 as
 *    zeros, deviceStructs.h:167, so its next re-initialisation divides by
 *    zero);
 *  - the infinite- and cylinder-wall spring histories of the existing solids
 *    are kept (the reference rebuilds them empty, dataBase.cpp:259-260);
 *  - with no bond yet, every contact may become one (the reference reads the
 *    last element of an empty vector there, :313);
 *  - walls, materials, contact tables, domain and gravity stay those of the
 *    scene (e must hold no walls).                                       */
static double shepard_sum(const orc_ctx* o, int i);
int orc_append_scene(orc_ctx* o, const dsph_scene* e) {
    if (!o || !e || !o->initialized) return DSPH_E_STATE;
    if (e->n_fluid < 0 || e->n_solid < 0 || e->n_clump < 0 || e->n_wall != 0 || e->n_cylinder != 0)
        return DSPH_E_INVALID;
    const int nf0 = o->nf, ns0 = o->ns, nc0 = o->nc;
    const int nf1 = e->n_fluid, ns1 = e->n_solid, nc1 = e->n_clump;
    for (int k = 0; k < nc1; k++)
        if (e->clump_pebble_start[k] < 0 || e->clump_pebble_end[k] > ns1 ||
            e->clump_pebble_start[k] > e->clump_pebble_end[k])
            return DSPH_E_INVALID;
    for (int i = 0; i < ns1; i++) {
        const int m = e->solid_material ? e->solid_material[i] : 0;
        const bool dummy = e->solid_dummy && e->solid_dummy[i];
        if (!dummy && (m < 0 || m >= o->nmat)) return DSPH_E_INVALID;
    }
    for (int i = 0; i < nf1; i++) {
        o->fpos.push_back(ld3(e->fluid_pos, i));
        o->fvel.push_back(e->fluid_vel ? ld3(e->fluid_vel, i) : mk(0, 0, 0));
        o->facc.push_back(mk(0, 0, 0));
        o->fh.push_back(e->fluid_h[i]);
        o->rho0.push_back(e->fluid_rho0[i]);
        o->drho.push_back(e->fluid_drho ? e->fluid_drho[i] : 0.0);
        o->c.push_back(e->fluid_c[i]);
        o->nu.push_back(e->fluid_nu[i]);
        o->alpha.push_back(e->fluid_alpha ? e->fluid_alpha[i] : 0.0);
        o->sumw0.push_back(0.0);
    }
    o->nf = nf0 + nf1;
    for (int i = 0; i < ns1; i++) {
        o->spos.push_back(ld3(e->solid_pos, i));
        o->svel.push_back(e->solid_vel ? ld3(e->solid_vel, i) : mk(0, 0, 0));
        o->sacc.push_back(mk(0, 0, 0));
        o->sw.push_back(e->solid_angvel ? ld3(e->solid_angvel, i) : mk(0, 0, 0));
        o->storq.push_back(mk(0, 0, 0));
        o->snorm.push_back(e->solid_normal ? ld3(e->solid_normal, i) : mk(0, 0, 0));
        o->srad.push_back(e->solid_radius[i]);
        o->seff.push_back(e->solid_eff_radius[i]);
        o->sinvm.push_back(e->solid_inv_mass[i]);
        o->smat.push_back(e->solid_material ? e->solid_material[i] : 0);
        o->sclus.push_back(e->solid_cluster ? e->solid_cluster[i] : -1);
        const int cl = e->solid_clump ? e->solid_clump[i] : -1;
        o->sclump.push_back(cl >= 0 ? cl + nc0 : -1);
        o->sdummy.push_back(e->solid_dummy ? e->solid_dummy[i] : 0);
    }
    o->ns = ns0 + ns1;
    for (int k = 0; k < nc1; k++) {
        o->cpos.push_back(ld3(e->clump_centroid, k));
        o->cvel.push_back(e->clump_vel ? ld3(e->clump_vel, k) : mk(0, 0, 0));
        o->cacc.push_back(mk(0, 0, 0));
        o->cw.push_back(e->clump_angvel ? ld3(e->clump_angvel, k) : mk(0, 0, 0));
        o->ctorq.push_back(mk(0, 0, 0));
        if (e->clump_orientation) {
            const double* q = e->clump_orientation + 4 * k;
            o->cq.push_back(Quat{q[0], q[1], q[2], q[3]});
        } else {
            o->cq.push_back(Quat{1, 0, 0, 0});
        }
        const double* I = e->clump_inv_inertia + 6 * k;
        o->cinvI.push_back(Sym{I[0], I[1], I[2], I[3], I[4], I[5]});
        o->cinvm.push_back(e->clump_inv_mass[k]);
        o->cstart.push_back(e->clump_pebble_start[k] + ns0);
        o->cend.push_back(e->clump_pebble_end[k] + ns0);
    }
    o->nc = nc0 + nc1;
    /* per-solid wall histories are solid-major: the new solids' start empty */
    const Springs zero{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
    o->wallHist.resize((size_t)o->ns * o->nw, zero);
    o->cylHist.resize((size_t)o->ns * o->ny, zero);
    /* the contact list's springs carry over by (i, j): its ranges per
       pointing solid cover the new solids too, with no pairs */
    o->ss.rstart.resize(o->ns, -1);
    o->ss.rend.resize(o->ns, -1);
    orc_neighbor_search(o, 0, 1);
    if (ns1 > 0) {
        int last = -1;  /* the last bonded pointed solid */
        for (int i : o->bI) last = std::max(last, i);
        const size_t nb0 = o->bI.size();
        add_bonds(o);
        /* keep the old bonds and the new ones past the last bonded solid */
        size_t w = nb0;
        for (size_t k = nb0; k < o->bI.size(); k++) {
            if (o->bI[k] <= last) continue;
            o->bN[w] = o->bN[k]; o->bPt[w] = o->bPt[k]; o->bFs[w] = o->bFs[k]; o->bTb[w] = o->bTb[k];
            o->bFn[w] = o->bFn[k]; o->bTt[w] = o->bTt[k];
            o->bI[w] = o->bI[k]; o->bJ[w] = o->bJ[k]; o->bOn[w] = o->bOn[k];
            w++;
        }
        o->bN.resize(w); o->bPt.resize(w); o->bFs.resize(w); o->bTb.resize(w);
        o->bFn.resize(w); o->bTt.resize(w); o->bI.resize(w); o->bJ.resize(w); o->bOn.resize(w);
    }
    for (int i = nf0; i < o->nf; i++) o->sumw0[i] = shepard_sum(o, i);
    return 0;
}

/* ---------------- fluid phase: integrate.cu:3-421 ---------------------- */

/* densityReinitialization (integrate.cu:3-53) */
/* the Shepard kernel sum of fluid particle i over its lists
 * (integrate.cu:22-44) */
static double shepard_sum(const orc_ctx* o, int i) {
    V3 ri = o->fpos[i];
    double hi = o->fh[i];
    double V0 = std::pow(hi / 1.3, 3);
    double sw = wendland(0, hi);
    for (int k = o->ff.seg_begin(i); k < o->ff.seg_end(i); k++)
        sw += wendland(len(ri - o->fpos[o->ff.pointing[k]]), hi);
    if (o->ff.rstart[i] != -1)
        for (int k = o->ff.rstart[i]; k < o->ff.rend[i]; k++)
            sw += wendland(len(ri - o->fpos[o->ff.pointed[o->ff.tindex[k]]]), hi);
    for (int k = o->fs.seg_begin(i); k < o->fs.seg_end(i); k++) {
        int a = o->fs.pointing[k];
        double Va = std::pow(2. * o->srad[a], 3);
        sw += wendland(len(ri - o->spos[a]), hi) * Va / V0;
    }
    return sw;
}

static void density_reinit(orc_ctx* o, int i_step) {
    std::vector<double> out(o->drho);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->nf; i++) {
        double rhos = o->rho0[i] + o->drho[i];
        const double sw = shepard_sum(o, i);
        if (i_step == 1) {
            o->sumw0[i] = sw;
        } else {
            double rho1 = std::fmax(rhos, o->rho0[i] * sw / o->sumw0[i]);
            out[i] = rho1 - o->rho0[i];
        }
    }
    if (i_step != 1) o->drho.swap(out);
}

/* solveMassConservationEquationDensityIntegrate (integrate.cu:55-161),
 * Jacobi: neighbours' dRho from the snapshot */
static void continuity(orc_ctx* o, double dt) {
    std::vector<double> out(o->nf);
    const V3 g = o->g;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->nf; i++) {
        V3 ri = o->fpos[i], vi = o->fvel[i];
        double hi = o->fh[i], rho0i = o->rho0[i], dri = o->drho[i];
        double rhoi = rho0i + dri, ci = o->c[i], Pi = dri * ci * ci;
        double acc = 0;
        auto pair = [&](int j) {
            V3 rj = o->fpos[j], vj = o->fvel[j];
            double hj = o->fh[j], rho0j = o->rho0[j], drj = o->drho[j];
            double rhoj = rho0j + drj, cj = o->c[j], Pj = drj * cj * cj;
            double mj = std::pow(hj / 1.3, 3) * o->rho0[j];
            double Vj = mj / rhoj;
            V3 rij = ri - rj;
            double h = 0.5 * (hi + hj), c = 0.5 * (ci + cj);
            V3 e = unit(rij);
            V3 dW = grad_wendland(rij, h);
            double UL = dot(vi, -e), UR = dot(vj, -e);
            double Us = (rhoi * UL + rhoj * UR + (Pi - Pj) / c) / (rhoi + rhoj);
            V3 vs = (Us - (rhoi * UL + rhoj * UR) / (rhoi + rhoj)) * (-e) +
                    (rhoi * vi + rhoj * vj) / (rhoi + rhoj);
            acc += 2 * rhoi * Vj * dot(vi - vs, dW);
        };
        for (int k = o->ff.seg_begin(i); k < o->ff.seg_end(i); k++) pair(o->ff.pointing[k]);
        if (o->ff.rstart[i] != -1)
            for (int k = o->ff.rstart[i]; k < o->ff.rend[i]; k++) pair(o->ff.pointed[o->ff.tindex[k]]);
        for (int k = o->fs.seg_begin(i); k < o->fs.seg_end(i); k++) {
            int a = o->fs.pointing[k];
            V3 ra = o->spos[a], va = o->svel[a];
            double Va = std::pow(2. * o->srad[a], 3);
            double ha = 2.6 * o->srad[a];
            double h = 0.5 * (hi + ha), c = ci;
            V3 ria = ri - ra;
            V3 e = unit(ri - ra);
            V3 dW = grad_wendland(ria, h);
            V3 aa = o->sacc[a], na = o->snorm[a];
            double Pa = Pi + rhoi * std::fmax(0., dot(g - aa, na)) * dot(ria, na);
            double rhoa = Pa / (c * c) + rho0i;
            V3 vs = (Pi - Pa) / c / (rhoi + rhoa) * (-e) +
                    (rhoi * vi + rhoa * (2 * va - vi)) / (rhoi + rhoa);
            acc += 2 * rhoi * Va * dot(vi - vs, dW);
        }
        out[i] = dri + dt * acc;
    }
    o->drho.swap(out);
}

/* wall/DEM term of the momentum equation, per fluid–solid pair
 * (integrate.cu:251-277 and 282-330) */
static V3 fluid_solid_term(const orc_ctx* o, int i, int a, double Pi, double rhoi,
                           double gammai, double mi, double Vi) {
    V3 ri = o->fpos[i], vi = o->fvel[i];
    double hi = o->fh[i], ci = o->c[i], rho0i = o->rho0[i];
    V3 ra = o->spos[a], va = o->svel[a];
    double Va = std::pow(2. * o->srad[a], 3);
    double ha = 2.6 * o->srad[a];
    double h = 0.5 * (hi + ha), c = ci;
    V3 ria = ri - ra;
    V3 e = unit(ria);
    V3 dW = grad_wendland(ria, h);
    V3 aa = o->sacc[a], na = o->snorm[a];
    double UL = dot(vi, -na);
    double UR = dot((2 * va - vi), -na);
    double Pa = Pi + rhoi * std::fmax(0., dot(o->g - aa, na)) * dot(ria, na);
    double rhoa = Pa / (c * c) + rho0i;
    double beta = std::fmin(3. * std::fmax(UL - UR, 0.), c);
    double Ps = (rhoi * Pa + rhoa * Pi + rhoi * rhoa * beta * (UL - UR)) / (rhoi + rhoa);
    return -2 / mi * Vi * Va * Ps * dW +
           2 / mi * Vi * Va * gammai * (vi - (2 * va - vi)) / len(ria) * dot(e, dW);
}

/* calFluid2SolidForce (integrate.cu:282-330) */
static void fluid_solid_force(orc_ctx* o) {
    int np = o->fs.num();
#pragma omp parallel for schedule(static)
    for (int k = 0; k < np; k++) {
        int i = o->fs.pointed[k], a = o->fs.pointing[k];
        double rho0i = o->rho0[i], dri = o->drho[i];
        double rhoi = rho0i + dri, ci = o->c[i];
        double Pi = dri * ci * ci;
        double gammai = o->nu[i] * rhoi;
        double mi = std::pow(o->fh[i] / 1.3, 3) * rho0i;
        double Vi = mi / rhoi;
        o->fsForce[k] = fluid_solid_term(o, i, a, Pi, rhoi, gammai, mi, Vi) * mi;
    }
}

/* solveMomentumConservationEquation (integrate.cu:163-280) */
static void momentum(orc_ctx* o) {
    const bool art = (o->flags & DSPH_FLAG_ART_VISCOSITY) != 0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->nf; i++) {
        V3 ri = o->fpos[i], vi = o->fvel[i];
        double hi = o->fh[i], rho0i = o->rho0[i], dri = o->drho[i];
        double rhoi = rho0i + dri, ci = o->c[i], Pi = dri * ci * ci;
        double gi = o->nu[i] * rhoi;
        double mi = std::pow(hi / 1.3, 3) * rho0i;
        double Vi = mi / rhoi;
        V3 dv = mk(0, 0, 0);
        auto pair = [&](int j) {
            V3 rj = o->fpos[j], vj = o->fvel[j];
            double hj = o->fh[j], rho0j = o->rho0[j], drj = o->drho[j];
            double rhoj = rho0j + drj, cj = o->c[j], Pj = drj * cj * cj;
            double gj = o->nu[j] * rhoj;
            double mj = std::pow(hj / 1.3, 3) * o->rho0[j];
            double Vj = mj / rhoj;
            V3 rij = ri - rj;
            double h = 0.5 * (hi + hj), c = 0.5 * (ci + cj);
            V3 dW = grad_wendland(rij, h);
            V3 e = unit(rij);
            double UL = dot(vi, -e), UR = dot(vj, -e);
            double beta = std::fmin(3. * std::fmax(UL - UR, 0.), c);
            double Ps = (rhoi * Pj + rhoj * Pi + rhoi * rhoj * beta * (UL - UR)) / (rhoi + rhoj);
            dv += -2 / mi * Vi * Vj * Ps * dW +
                  2 / mi * Vi * Vj * 2 * gi * gj / (gi + gj) * (vi - vj) / len(rij) * dot(e, dW);
            if (art) {
                /* getArtificialViscosityAccelerationFromParticleJ
                 * (SPHKernelDEMContactModel.h:74-77): m_j alpha h_ij c_ij
                 * (v_ij . r_ij) / (rho_ij r^2 + 0.01 h_ij^2) grad W_ij, signed
                 * so that the pair term removes kinetic energy */
                double rho_ij = 0.5 * (rhoi + rhoj);
                dv += mj * o->alpha[i] * h * c * dot(vi - vj, rij) /
                      (rho_ij * dot(rij, rij) + 0.01 * h * h) * dW;
            }
        };
        for (int k = o->ff.seg_begin(i); k < o->ff.seg_end(i); k++) pair(o->ff.pointing[k]);
        if (o->ff.rstart[i] != -1)
            for (int k = o->ff.rstart[i]; k < o->ff.rend[i]; k++) pair(o->ff.pointed[o->ff.tindex[k]]);
        for (int k = o->fs.seg_begin(i); k < o->fs.seg_end(i); k++)
            dv += fluid_solid_term(o, i, o->fs.pointing[k], Pi, rhoi, gi, mi, Vi);
        o->facc[i] = dv + o->g;
    }
}

/* fluidIntegrate (integrate.cu:385-421) */
int orc_fluid_integrate(orc_ctx* o, double dt, int32_t i_step, int32_t gap) {
    if (o->nf == 0 && o->fs.num() == 0) return 0;
    double dtf = dt * double(gap);
    if (o->nf > 0 && (i_step % (5 * gap) == 0 || i_step == 1)) density_reinit(o, i_step);
    if (i_step % gap == 0 && o->nf > 0) {
        continuity(o, 0.5 * dtf);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < o->nf; i++) o->fpos[i] += 0.5 * dtf * o->fvel[i];
    }
    fluid_solid_force(o);
    if (i_step % gap == 0 && o->nf > 0) {
        momentum(o);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < o->nf; i++) o->fvel[i] += dtf * o->facc[i];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < o->nf; i++) o->fpos[i] += 0.5 * dtf * o->fvel[i];
        continuity(o, 0.5 * dtf);
    }
    return 0;
}

/* ---------------- solid phase: integrate.cu:423-998 -------------------- */

static void run_contact_model(const orc_ctx* o, int cij, V3& F, V3& T, Springs& sp,
                              V3 v, V3 w, V3 n, double delta, double m, double R,
                              double dt) {
    if (o->linear[0 * o->ncomb + cij] > 1.e-10) {
        double p[11];
        for (int f = 0; f < 11; f++) p[f] = o->linear[f * o->ncomb + cij];
        linear_contact(F, T, sp, v, w, n, delta, m, R, dt, p);
    } else {
        double lr = std::log(hz(o, 2, cij));
        double D = -lr / std::sqrt(lr * lr + kPi * kPi);
        double p[8] = {hz(o, 0, cij), hz(o, 1, cij), D, hz(o, 3, cij),
                       hz(o, 4, cij), hz(o, 5, cij), hz(o, 6, cij), hz(o, 7, cij)};
        hertz_contact(F, T, sp, v, w, n, delta, m, R, dt, p);
    }
}

/* calSolidContactForceTorque (integrate.cu:423-512) */
static void solid_contacts(orc_ctx* o, double dt) {
    int np = o->ss.num();
#pragma omp parallel for schedule(static)
    for (int k = 0; k < np; k++) {
        int i = o->ss.pointed[k], j = o->ss.pointing[k];
        double ri = o->srad[i], rj = o->srad[j];
        V3 pi = o->spos[i], pj = o->spos[j];
        V3 n = unit(pi - pj);
        double delta = ri + rj - len(pi - pj);
        V3 rc = pj + (rj - 0.5 * delta) * n;
        double R = ri * rj / (ri + rj);
        double m = 1. / (o->sinvm[i] + o->sinvm[j]);
        V3 vi = o->svel[i], vj = o->svel[j], wi = o->sw[i], wj = o->sw[j];
        V3 vc = vi + cross(wi, rc - pi) - (vj + cross(wj, rc - pj));
        V3 F = mk(0, 0, 0), T = mk(0, 0, 0);
        Springs sp = o->ssSpring[k];
        int cij = combined_index(o->nmat, o->smat[i], o->smat[j]);
        run_contact_model(o, cij, F, T, sp, vc, wi - wj, n, delta, m, R, dt);
        o->ssForce[k] = F;
        o->ssTorque[k] = T;
        o->ssSpring[k] = sp;
    }
}

/* calSolidBondedForceTorque (integrate.cu:717-804) */
static void solid_bonds(orc_ctx* o, double dt) {
    int nb = (int)o->bI.size();
    for (int b = 0; b < nb; b++) {   /* serial: bonds add into contacts */
        if (o->bOn[b] == 0) {
            o->bFn[b] = 0; o->bTt[b] = 0;
            o->bFs[b] = mk(0, 0, 0); o->bTb[b] = mk(0, 0, 0); o->bN[b] = mk(0, 0, 0);
            continue;
        }
        int i = o->bI[b], j = o->bJ[b];
        int kc = -1;
        for (int k = o->ss.seg_begin(i); k < o->ss.seg_end(i); k++)
            if (o->ss.pointing[k] == j) { kc = k; break; }
        if (kc < 0) { o->bOn[b] = 0; continue; }
        double ri = o->srad[i], rj = o->srad[j];
        V3 pi = o->spos[i], pj = o->spos[j];
        V3 n = unit(pi - pj);
        V3 n0 = o->bN[b];
        double delta = ri + rj - len(pi - pj);
        V3 rc = pj + (rj - 0.5 * delta) * n;
        V3 vi = o->svel[i], vj = o->svel[j], wi = o->sw[i], wj = o->sw[j];
        V3 vc = vi + cross(wi, rc - pi) - (vj + cross(wj, rc - pj));
        int cij = combined_index(o->nmat, o->smat[i], o->smat[j]);
        double p[6];
        for (int f = 0; f < 6; f++) p[f] = o->bonded[f * o->ncomb + cij];
        double Fn = o->bFn[b], Tt = o->bTt[b];
        V3 Fs = o->bFs[b], Tb = o->bTb[b];
        o->bOn[b] = parallel_bond(Fn, Tt, Fs, Tb, n0, n, vc, wi, wj, ri, rj, dt, p);
        o->bFn[b] = Fn; o->bFs[b] = Fs; o->bTt[b] = Tt; o->bTb[b] = Tb;
        o->bN[b] = n; o->bPt[b] = rc;
        o->ssForce[kc] += Fn * n + Fs;
        o->ssTorque[kc] += Tt * n + Tb;
    }
}

/* calSolidForceTorque (integrate.cu:806-858) */
static void solid_force_sum(orc_ctx* o) {
    const V3 g = o->g;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->ns; i++) {
        double ri = o->srad[i];
        V3 pi = o->spos[i];
        V3 F = mk(0, 0, 0), T = mk(0, 0, 0);
        for (int k = o->ss.seg_begin(i); k < o->ss.seg_end(i); k++) {
            int j = o->ss.pointing[k];
            double rj = o->srad[j];
            V3 pj = o->spos[j];
            V3 n = unit(pi - pj);
            double delta = ri + rj - len(pi - pj);
            V3 rc = pj + (rj - 0.5 * delta) * n;
            F += o->ssForce[k];
            T += o->ssTorque[k] + cross(rc - pi, o->ssForce[k]);
        }
        if (o->ss.rstart[i] != -1)
            for (int k = o->ss.rstart[i]; k < o->ss.rend[i]; k++) {
                int k1 = o->ss.tindex[k];
                int j = o->ss.pointed[k1];
                double rj = o->srad[j];
                V3 pj = o->spos[j];
                V3 n = unit(pi - pj);
                double delta = ri + rj - len(pi - pj);
                V3 rc = pj + (rj - 0.5 * delta) * n;
                F -= o->ssForce[k1];
                T -= o->ssTorque[k1];
                T -= cross(rc - pi, o->ssForce[k1]);
            }
        if (!o->fs.rstart.empty() && o->fs.rstart[i] != -1)
            for (int k = o->fs.rstart[i]; k < o->fs.rend[i]; k++) F -= o->fsForce[o->fs.tindex[k]];
        double gon = (o->sinvm[i] > 0. && o->sclump[i] < 0) ? 1.0 : 0.0;
        o->sacc[i] = F * o->sinvm[i] + g * gon;
        o->storq[i] = T;
    }
}

/* calInfiniteWallForceTorque (integrate.cu:514-610) */
static void wall_forces(orc_ctx* o, double dt) {
    int nw = o->nw;
    bool cont = (o->flags & DSPH_FLAG_WALL_CONTINUE) != 0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->ns; i++) {
        if (o->sdummy[i] == 1 || o->sinvm[i] < 1.e-10) continue;
        double ri = o->srad[i];
        V3 pi = o->spos[i], vi = o->svel[i], wi = o->sw[i];
        for (int k = 0; k < nw; k++) {
            Springs& h = o->wallHist[(size_t)i * nw + k];
            V3 n = o->wn[k];
            double delta = ri - dot(pi - o->wpt[k], n);
            if (delta < 0.) {
                h = Springs{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
                if (cont) continue;
                break;
            }
            double m = 1. / o->sinvm[i];
            V3 rc = pi - (ri - delta) * n;
            V3 vc = vi + cross(wi, rc - pi);
            V3 F = mk(0, 0, 0), T = mk(0, 0, 0);
            int cij = combined_index(o->nmat, o->smat[i], o->wmat[k]);
            run_contact_model(o, cij, F, T, h, vc, wi, n, delta, m, ri, dt);
            o->sacc[i] += F * o->sinvm[i];
            o->storq[i] += T + cross(rc - pi, F);
        }
    }
}

/* calCylinderWallForceTorque (integrate.cu:612-715): sphere against the
 * wall of radius R around an axis through p along a.  With d the centre's
 * distance from the axis, an inside contact overlaps by d + r - R and an
 * outside one (d >= R) by r + R - d; the contact point is R - delta/2 from
 * the axis.  The reference takes the outward normal for both; an inside
 * contact here takes the inward one, so the wall holds the sphere in (the
 * reference never launches this kernel: its cylinder walls are not copied
 * to the device).  The loop stops at the first wall out of reach, as the
 * infinite-wall loop does, unless DSPH_FLAG_WALL_CONTINUE. */
static void cylinder_forces(orc_ctx* o, double dt) {
    int ny = o->ny;
    bool cont = (o->flags & DSPH_FLAG_WALL_CONTINUE) != 0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->ns; i++) {
        if (o->sdummy[i] == 1 || o->sinvm[i] < 1.e-10) continue;
        double ri = o->srad[i];
        V3 pi = o->spos[i], vi = o->svel[i], wi = o->sw[i];
        for (int k = 0; k < ny; k++) {
            Springs& h = o->cylHist[(size_t)i * ny + k];
            double R = o->yrad[k];
            V3 a = o->yax[k], pk = o->ypt[k];
            V3 c0 = pk + a * dot(pi - pk, a);
            V3 out = unit(pi - c0);
            double d = len(pi - c0);
            double delta = d + ri - R;
            bool outside = false;
            if (delta >= ri) {
                delta = ri + R - d;
                outside = true;
            }
            if (delta < 0.) {
                h = Springs{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
                if (cont) continue;
                break;
            }
            V3 n = outside ? out : -out;
            V3 rc = c0 + out * (R - 0.5 * delta);
            V3 wk = o->yw[k];
            V3 vc = vi + cross(wi, rc - pi) - cross(wk, rc - pk);
            V3 wrel = wi - wk;
            double m = 1. / o->sinvm[i];
            V3 F = mk(0, 0, 0), T = mk(0, 0, 0);
            int cij = combined_index(o->nmat, o->smat[i], o->ymat[k]);
            run_contact_model(o, cij, F, T, h, vc, wrel, n, delta, m, ri, dt);
            o->sacc[i] += F * o->sinvm[i];
            o->storq[i] += T + cross(rc - pi, F);
        }
    }
}

/* calClumpForceTorque (integrate.cu:860-881) */
static void clump_force_sum(orc_ctx* o) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; c++) {
        V3 F = mk(0, 0, 0), T = mk(0, 0, 0);
        for (int i = o->cstart[c]; i < o->cend[c]; i++) {
            V3 Fi = mk(0, 0, 0);
            if (o->sinvm[i] > 1.e-10) Fi = o->sacc[i] / o->sinvm[i];
            F += Fi;
            T += o->storq[i] + cross(o->spos[i] - o->cpos[c], Fi);
            o->sacc[i] = mk(0, 0, 0);
            o->storq[i] = mk(0, 0, 0);
        }
        o->cacc[c] = F * o->cinvm[c] + o->g * (o->cinvm[c] > 0. ? 1.0 : 0.0);
        o->ctorq[c] = T;
    }
}

int orc_contact_forces(orc_ctx* o, double dt) {
    solid_contacts(o, dt);
    solid_bonds(o, dt);
    solid_force_sum(o);
    if (o->nw > 0) wall_forces(o, dt);
    if (o->ny > 0) cylinder_forces(o, dt);
    clump_force_sum(o);
    return 0;
}

/* solidVelocityAngularVelocityIntegrate (integrate.cu:883-897) */
static void solid_kick(orc_ctx* o, double dt) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->ns; i++) {
        if (o->sclump[i] >= 0) continue;
        o->svel[i] += o->sacc[i] * dt;
        double r = o->srad[i], im = o->sinvm[i];
        if (im < 1.e-10 || r < 1.e-10) continue;
        double I = 0.4 * r * r / im;
        o->sw[i] += o->storq[i] / I * dt;
    }
}
/* solidPositionIntegrate (integrate.cu:899-906) */
static void solid_drift(orc_ctx* o, double dt) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->ns; i++) o->spos[i] += dt * o->svel[i];
}
/* clumpVelocityAngularVelocityIntegrate (integrate.cu:908-929) */
static void clump_kick(orc_ctx* o, double dt) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; c++) {
        V3 ac = o->cacc[c];
        V3 w0 = o->cw[c];
        o->cvel[c] += ac * dt;
        if (o->cinvm[c] > 0.) o->cw[c] += mul(rotate_inv_inertia(o->cq[c], o->cinvI[c]), o->ctorq[c]) * dt;
        V3 w = o->cw[c];
        for (int i = o->cstart[c]; i < o->cend[c]; i++) {
            V3 r = o->spos[i] - o->cpos[c];
            o->sacc[i] = ac + cross(ac, r) + cross(w, cross(w, r));
            o->svel[i] = o->cvel[c] + cross(w, r);
            o->sw[i] = w;
            o->snorm[i] = rodrigues(o->snorm[i], 0.5 * (w0 + w) * dt);
        }
    }
}
/* clumpPositionIntegrate (integrate.cu:931-938); orientation update only
 * with DSPH_FLAG_CLUMP_QUATERNION */
static void clump_drift(orc_ctx* o, double dt) {
    bool quat = (o->flags & DSPH_FLAG_CLUMP_QUATERNION) != 0;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; c++) {
        o->cpos[c] += dt * o->cvel[c];
        if (quat) o->cq[c] = quat_rotate(o->cq[c], o->cw[c], dt);
    }
}

/* solidIntegrateBeforeContact (integrate.cu:964-980) */
int orc_solid_before_contact(orc_ctx* o, double dt) {
    solid_kick(o, 0.5 * dt);
    clump_kick(o, 0.5 * dt);
    clump_drift(o, 0.5 * dt);
    solid_drift(o, 0.5 * dt);
    return 0;
}
/* solidIntegrateAfterContact (integrate.cu:982-998) */
int orc_solid_after_contact(orc_ctx* o, double dt) {
    solid_drift(o, 0.5 * dt);
    clump_drift(o, 0.5 * dt);
    clump_kick(o, 0.5 * dt);
    solid_kick(o, 0.5 * dt);
    return 0;
}

/* dataBase::addExternalForce (dataBase.cpp:179-195): a force on one solid
 * between the contact phase and solidIntegrateAfterContact (the user hook
 * of solverBase::update), as an acceleration increment F / m */
int orc_add_external_force(orc_ctx* o, int32_t i, const double f[3]) {
    if (i < 0 || i >= o->ns) return -1;
    o->sacc[i] += mk(f[0], f[1], f[2]) * o->sinvm[i];
    return 0;
}

/* dataBase::addGlobalDamping (dataBase.cpp:197-224): the solid's force and
 * torque each reduced by C_d times their magnitude along its velocity and
 * angular velocity (skipped for fixed solids and slower than 1e-10) */
int orc_add_global_damping(orc_ctx* o, int32_t i, double cd) {
    if (i < 0 || i >= o->ns) return -1;
    const double im = o->sinvm[i];
    if (im < 1.e-10) return 0;
    V3 F = o->sacc[i] / im, T = o->storq[i];
    const V3 v = o->svel[i], w = o->sw[i];
    if (len(v) > 1.e-10) F += -cd * len(F) * unit(v);
    if (len(w) > 1.e-10) T += -cd * len(T) * unit(w);
    o->sacc[i] = F * im;
    o->storq[i] = T;
    return 0;
}

/* solverBase::update (solverBase.cpp:47-59) without hook and output */
int orc_step(orc_ctx* o, int32_t i_step, double dt, int32_t gap) {
    orc_neighbor_search(o, i_step, gap);
    orc_solid_before_contact(o, dt);
    orc_fluid_integrate(o, dt, i_step, gap);
    orc_contact_forces(o, dt);
    orc_solid_after_contact(o, dt);
    return 0;
}

/* ---------------- read-back -------------------------------------------- */
int orc_download_fluid(orc_ctx* o, double* pos, double* vel, double* acc, double* drho,
                       double* sumw0) {
    for (int i = 0; i < o->nf; i++) {
        if (pos) st3(pos, i, o->fpos[i]);
        if (vel) st3(vel, i, o->fvel[i]);
        if (acc) st3(acc, i, o->facc[i]);
        if (drho) drho[i] = o->drho[i];
        if (sumw0) sumw0[i] = o->sumw0[i];
    }
    return 0;
}
int orc_download_solid(orc_ctx* o, double* pos, double* vel, double* acc, double* angvel,
                       double* torque, double* normal) {
    for (int i = 0; i < o->ns; i++) {
        if (pos) st3(pos, i, o->spos[i]);
        if (vel) st3(vel, i, o->svel[i]);
        if (acc) st3(acc, i, o->sacc[i]);
        if (angvel) st3(angvel, i, o->sw[i]);
        if (torque) st3(torque, i, o->storq[i]);
        if (normal) st3(normal, i, o->snorm[i]);
    }
    return 0;
}
int orc_download_clump(orc_ctx* o, double* centroid, double* vel, double* acc, double* angvel,
                       double* torque, double* orientation) {
    for (int c = 0; c < o->nc; c++) {
        if (centroid) st3(centroid, c, o->cpos[c]);
        if (vel) st3(vel, c, o->cvel[c]);
        if (acc) st3(acc, c, o->cacc[c]);
        if (angvel) st3(angvel, c, o->cw[c]);
        if (torque) st3(torque, c, o->ctorq[c]);
        if (orientation) {
            orientation[4 * c] = o->cq[c].q0; orientation[4 * c + 1] = o->cq[c].q1;
            orientation[4 * c + 2] = o->cq[c].q2; orientation[4 * c + 3] = o->cq[c].q3;
        }
    }
    return 0;
}
int orc_download_fluid_solid_force(orc_ctx* o, double* force, int64_t cap, int64_t* n_out) {
    int64_t n = o->fs.num();
    if (n_out) *n_out = n;
    if (force) {
        if (cap < n) return DSPH_E_INVALID;
        for (int64_t k = 0; k < n; k++) st3(force, (int)k, o->fsForce[k]);
    }
    return 0;
}
int orc_download_contacts(orc_ctx* o, double* force, double* torque, double* sliding,
                          double* rolling, double* torsion, int64_t cap, int64_t* n_out) {
    int64_t n = o->ss.num();
    if (n_out) *n_out = n;
    if (cap < n && (force || torque || sliding || rolling || torsion)) return DSPH_E_INVALID;
    for (int64_t k = 0; k < n; k++) {
        if (force) st3(force, (int)k, o->ssForce[k]);
        if (torque) st3(torque, (int)k, o->ssTorque[k]);
        if (sliding) st3(sliding, (int)k, o->ssSpring[k].s);
        if (rolling) st3(rolling, (int)k, o->ssSpring[k].r);
        if (torsion) st3(torsion, (int)k, o->ssSpring[k].t);
    }
    return 0;
}
int orc_download_bonds(orc_ctx* o, double* normal_force, double* torsion, double* shear,
                       double* bending, int32_t* is_bonded, int64_t cap, int64_t* n_out) {
    int64_t n = (int64_t)o->bI.size();
    if (n_out) *n_out = n;
    if (cap < n && (normal_force || torsion || shear || bending || is_bonded)) return DSPH_E_INVALID;
    for (int64_t b = 0; b < n; b++) {
        if (normal_force) normal_force[b] = o->bFn[b];
        if (torsion) torsion[b] = o->bTt[b];
        if (shear) st3(shear, (int)b, o->bFs[b]);
        if (bending) st3(bending, (int)b, o->bTb[b]);
        if (is_bonded) is_bonded[b] = o->bOn[b];
    }
    return 0;
}
static const PairList* list_of(orc_ctx* o, int kind) {
    if (kind == DSPH_LIST_FLUID_FLUID) return &o->ff;
    if (kind == DSPH_LIST_FLUID_SOLID) return &o->fs;
    if (kind == DSPH_LIST_SOLID_SOLID) return &o->ss;
    return nullptr;
}
int orc_download_list(orc_ctx* o, int kind, int32_t* pointed, int32_t* pointing, int64_t cap,
                      int64_t* n_out, int32_t* count) {
    if (kind == DSPH_LIST_BOND) {
        int64_t n = (int64_t)o->bI.size();
        if (n_out) *n_out = n;
        if ((pointed || pointing) && cap < n) return DSPH_E_INVALID;
        for (int64_t b = 0; b < n; b++) {
            if (pointed) pointed[b] = o->bI[b];
            if (pointing) pointing[b] = o->bJ[b];
        }
        return 0;
    }
    const PairList* L = list_of(o, kind);
    if (!L) return DSPH_E_INVALID;
    int64_t n = L->num();
    if (n_out) *n_out = n;
    if ((pointed || pointing) && cap < n) return DSPH_E_INVALID;
    for (int64_t k = 0; k < n; k++) {
        if (pointed) pointed[k] = L->pointed[k];
        if (pointing) pointing[k] = L->pointing[k];
    }
    if (count)
        for (size_t i = 0; i < L->count.size(); i++) count[i] = L->count[i];
    return 0;
}
int orc_download_transpose(orc_ctx* o, int kind, int32_t* index, int64_t cap,
                           int32_t* range_start, int32_t* range_end) {
    const PairList* L = list_of(o, kind);
    if (!L) return DSPH_E_INVALID;
    if (index) {
        if (cap < L->num()) return DSPH_E_INVALID;
        for (int k = 0; k < L->num(); k++) index[k] = L->tindex[k];
    }
    for (size_t i = 0; i < L->rstart.size(); i++) {
        if (range_start) range_start[i] = L->rstart[i];
        if (range_end) range_end[i] = L->rend[i];
    }
    return 0;
}
int orc_grid_info(orc_ctx* o, int32_t* gs, double* cs, int32_t* n_cells) {
    if (gs) { gs[0] = o->gx; gs[1] = o->gy; gs[2] = o->gz; }
    if (cs) { cs[0] = o->cell.x; cs[1] = o->cell.y; cs[2] = o->cell.z; }
    if (n_cells) *n_cells = o->ncell;
    return 0;
}
int orc_download_grid(orc_ctx* o, int species, int32_t* hash, int32_t* perm, int32_t* cs,
                      int32_t* ce) {
    const std::vector<int>& h = species == 0 ? o->fhash : o->shash;
    const std::vector<int>& p = species == 0 ? o->fperm : o->sperm;
    const std::vector<int>& s = species == 0 ? o->fstart : o->sstart;
    const std::vector<int>& e = species == 0 ? o->fend : o->send;
    if (hash) std::copy(h.begin(), h.end(), hash);
    if (perm) std::copy(p.begin(), p.end(), perm);
    if (cs) std::copy(s.begin(), s.end(), cs);
    if (ce) std::copy(e.begin(), e.end(), ce);
    return 0;
}

/* ---------------- single-formula entry points -------------------------- */
double orc_wendland(double r, double h) { return wendland(r, h); }
void orc_grad_wendland(const double rij[3], double h, double out[3]) {
    V3 g = grad_wendland(mk(rij[0], rij[1], rij[2]), h);
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}
void orc_contact(int model, const double v_rel[3], const double w_rel[3], const double n[3],
                 double overlap, double m_eff, double r_eff, double dt, const double* params,
                 double springs[9], double out[6]) {
    Springs sp{ld3(springs, 0), ld3(springs, 1), ld3(springs, 2)};
    V3 F = mk(0, 0, 0), T = mk(0, 0, 0);
    if (model == 0)
        hertz_contact(F, T, sp, ld3(v_rel, 0), ld3(w_rel, 0), ld3(n, 0), overlap, m_eff, r_eff, dt, params);
    else
        linear_contact(F, T, sp, ld3(v_rel, 0), ld3(w_rel, 0), ld3(n, 0), overlap, m_eff, r_eff, dt, params);
    st3(springs, 0, sp.s); st3(springs, 1, sp.r); st3(springs, 2, sp.t);
    st3(out, 0, F); st3(out, 1, T);
}
int orc_parallel_bond(double st[8], const double n_prev[3], const double n[3], const double v_rel[3],
                      const double w_a[3], const double w_b[3], double r_a, double r_b, double dt,
                      const double params[6]) {
    double Fn = st[0], Tt = st[1];
    V3 Fs = mk(st[2], st[3], st[4]), Tb = mk(st[5], st[6], st[7]);
    int on = parallel_bond(Fn, Tt, Fs, Tb, ld3(n_prev, 0), ld3(n, 0), ld3(v_rel, 0), ld3(w_a, 0),
                           ld3(w_b, 0), r_a, r_b, dt, params);
    st[0] = Fn; st[1] = Tt; st[2] = Fs.x; st[3] = Fs.y; st[4] = Fs.z;
    st[5] = Tb.x; st[6] = Tb.y; st[7] = Tb.z;
    return on;
}
void orc_rotate_inverse_inertia(const double q[4], const double I[6], double out[6]) {
    Sym m = rotate_inv_inertia(Quat{q[0], q[1], q[2], q[3]}, Sym{I[0], I[1], I[2], I[3], I[4], I[5]});
    out[0] = m.xx; out[1] = m.yy; out[2] = m.zz; out[3] = m.xy; out[4] = m.xz; out[5] = m.yz;
}
void orc_quaternion_rotate(const double q[4], const double w[3], double dt, double out[4]) {
    Quat r = quat_rotate(Quat{q[0], q[1], q[2], q[3]}, ld3(w, 0), dt);
    out[0] = r.q0; out[1] = r.q1; out[2] = r.q2; out[3] = r.q3;
}
void orc_rotate_vector(const double v[3], const double a[3], double out[3]) {
    st3(out, 0, rodrigues(ld3(v, 0), ld3(a, 0)));
}
int orc_combined_index(int n, int a, int b) { return combined_index(n, a, b); }

} // extern "C"
