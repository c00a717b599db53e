// k_dem.hip — DEM contact forces, bonds, walls, clumps and solid
// integration, in FP64 (the contact history and the rigid-body state stay
// FP64 in both precision modes).
//
// Follows core/integrate.cu: calSolidContactForceTorque (:423-512),
// calInfiniteWallForceTorque (:514-610), calSolidBondedForceTorque
// (:717-804), calSolidForceTorque (:806-858), calClumpForceTorque
// (:860-881), solid/clump integrators (:883-938) and the phase drivers
// (:940-998); the runtime mutations of core/dataBase.cpp:179-224.
//
// Per-solid sums are gathers over the forward (CSR) and reverse (transposed
// CSR) contact lists, as in the reference — no floating-point atomics, so a
// run is bit-reproducible.
#include "dsph_internal.h"

namespace dsph {

namespace {

constexpr int kBS = 256;

__device__ __forceinline__ d3 v3(double4 v) { return d3{v.x, v.y, v.z}; }
__device__ __forceinline__ double4 d4(d3 v, double w = 0.0) { return make_double4(v.x, v.y, v.z, w); }
__device__ __forceinline__ Springs ld_springs(const double* p) {
    return Springs{d3{p[0], p[1], p[2]}, d3{p[3], p[4], p[5]}, d3{p[6], p[7], p[8]}};
}
__device__ __forceinline__ void st_springs(double* p, const Springs& s) {
    p[0] = s.s.x; p[1] = s.s.y; p[2] = s.s.z;
    p[3] = s.r.x; p[4] = s.r.y; p[5] = s.r.z;
    p[6] = s.t.x; p[7] = s.t.y; p[8] = s.t.z;
}

__device__ __forceinline__ int64_t ss_total(const DemArgs& a) {
    int64_t t = a.L.ssoff[a.ns];
    return t > a.L.ss_cap ? a.L.ss_cap : t;
}

// contact point r_c = r_j + (r_j - delta/2) n_ij, n_ij = normalize(r_i - r_j)
__device__ __forceinline__ d3 contact_point(d3 pi, d3 pj, double ri, double rj, d3& n, double& delta) {
    d3 dd = pi - pj;
    n = unit(dd);
    delta = ri + rj - vlen(dd);
    return pj + (rj - 0.5 * delta) * n;
}

__global__ void k_ss_contact(DemArgs a, double dt) {
    const int64_t tot = ss_total(a);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < tot;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int i = a.L.ss_i[k], j = a.L.ss_j[k];
        const double dti = step_of(a.sdt, a.sscene, i, dt);
        const double4 pri = a.so.pr[i], prj = a.so.pr[j];
        const d3 pi = v3(pri), pj = v3(prj);
        d3 n;
        double delta;
        const d3 rc = contact_point(pi, pj, pri.w, prj.w, n, delta);
        const double R = pri.w * prj.w / (pri.w + prj.w);
        const double4 vi4 = a.so.vel[i], vj4 = a.so.vel[j];
        const double m = 1. / (vi4.w + vj4.w);
        const d3 wi = v3(a.so.w[i]), wj = v3(a.so.w[j]);
        const d3 vc = v3(vi4) + cross(wi, rc - pi) - (v3(vj4) + cross(wj, rc - pj));
        d3 F{0, 0, 0}, T{0, 0, 0};
        Springs sp = ld_springs(a.L.ssS + 9 * k);
        const int c = combined_index(a.tb.nmat, a.so.mat[i], a.so.mat[j]);
        contact_model(a.tb, c, F, T, sp, vc, wi - wj, n, delta, m, R, dti);
        a.L.ssF[k] = d4(F);
        a.L.ssT[k] = d4(T);
        st_springs(a.L.ssS + 9 * k, sp);
    }
}

__global__ void k_bonds(DemArgs a, double dt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.b.n) return;
    if (a.b.on[b] == 0) {
        a.b.Fn[b] = 0;
        a.b.Tt[b] = 0;
        a.b.Fs[b] = make_double4(0, 0, 0, 0);
        a.b.Tb[b] = make_double4(0, 0, 0, 0);
        a.b.nrm[b] = make_double4(0, 0, 0, 0);
        return;
    }
    const int i = a.b.i[b], j = a.b.j[b];
    int kc = -1;
    const int64_t e1 = (int64_t)a.L.ssoff[i] + a.L.sscnt[i];
    const int64_t e = e1 < a.L.ss_cap ? e1 : a.L.ss_cap;
    for (int64_t k = a.L.ssoff[i]; k < e; k++)
        if (a.L.ss_j[k] == j) {
            kc = (int)k;
            break;
        }
    if (kc < 0) {
        a.b.on[b] = 0;
        return;
    }
    const double4 pri = a.so.pr[i], prj = a.so.pr[j];
    const d3 pi = v3(pri), pj = v3(prj);
    d3 n;
    double delta;
    const d3 rc = contact_point(pi, pj, pri.w, prj.w, n, delta);
    const d3 wi = v3(a.so.w[i]), wj = v3(a.so.w[j]);
    const d3 vc = v3(a.so.vel[i]) + cross(wi, rc - pi) - (v3(a.so.vel[j]) + cross(wj, rc - pj));
    const int c = combined_index(a.tb.nmat, a.so.mat[i], a.so.mat[j]);
    double Fn = a.b.Fn[b], Tt = a.b.Tt[b];
    d3 Fs = v3(a.b.Fs[b]), Tb = v3(a.b.Tb[b]);
    a.b.on[b] = parallel_bond(Fn, Tt, Fs, Tb, v3(a.b.nrm[b]), n, vc, wi, wj, pri.w, prj.w,
                              step_of(a.sdt, a.sscene, i, dt), a.tb.bonded, a.tb.ncomb, c);
    a.b.Fn[b] = Fn;
    a.b.Tt[b] = Tt;
    a.b.Fs[b] = d4(Fs);
    a.b.Tb[b] = d4(Tb);
    a.b.nrm[b] = d4(n);
    a.b.pt[b] = d4(rc);
    const double4 F0 = a.L.ssF[kc], T0 = a.L.ssT[kc];
    a.L.ssF[kc] = d4(v3(F0) + (Fn * n + Fs));
    a.L.ssT[kc] = d4(v3(T0) + (Tt * n + Tb));
}

// the per-solid DEM kernels run over the live solids only: wall dummies
// never move, take no contacts and keep zero acceleration and torque
// infinite walls on live solid i (after its contact sums): adds the wall
// forces and torques to acc and tq
__device__ void wall_forces(const DemArgs& a, int i, double dt, d3& acc, d3& tq) {
    const double4 v4 = a.so.vel[i];
    if (a.so.dummy[i] == 1 || v4.w < 1.e-10) return;
    const bool cont = (a.flags & DSPH_FLAG_WALL_CONTINUE) != 0;
    const double4 pr = a.so.pr[i];
    const d3 pi = v3(pr), vi = v3(v4), wi = v3(a.so.w[i]);
    // the walls of the solid's own scene, in their upload order
    const int sc = a.nscene == 1 ? 0 : a.sscene[i];
    const int k0 = a.nscene == 1 ? 0 : a.woff[sc], k1 = a.nscene == 1 ? a.nw : a.woff[sc + 1];
    for (int k = k0; k < k1; k++) {
        double* h = a.whist + 9 * ((int64_t)i * a.nwmax + (k - k0));
        const d3 n = v3(a.wn[k]);
        const double delta = pr.w - dot(pi - v3(a.wpt[k]), n);
        if (delta < 0.) {
            for (int m = 0; m < 9; m++) h[m] = 0.0;
            if (cont) continue;
            break;
        }
        const double m = 1. / v4.w;
        const d3 rc = pi - (pr.w - delta) * n;
        const d3 vc = vi + cross(wi, rc - pi);
        d3 F{0, 0, 0}, T{0, 0, 0};
        Springs sp = ld_springs(h);
        const int c = combined_index(a.tb.nmat, a.so.mat[i], a.wmat[k]);
        contact_model(a.tb, c, F, T, sp, vc, wi, n, delta, m, pr.w, dt);
        st_springs(h, sp);
        acc += F * v4.w;
        tq += T + cross(rc - pi, F);
    }
}

// cylinder walls on live solid i (calCylinderWallForceTorque,
// integrate.cu:612-715), after the infinite walls.  With d the centre's
// distance from the axis, an inside contact overlaps by d + r - R and an
// outside one (d >= R) by r + R - d; the contact point is R - delta/2 from
// the axis.  The reference takes the outward normal for both (and never
// launches the kernel: its cylinder walls are not copied to the device);
// an inside contact here takes the inward one, so the wall holds the
// sphere in.  The loop stops at the first wall out of reach, as the
// infinite-wall loop does, unless DSPH_FLAG_WALL_CONTINUE.
__device__ void cylinder_forces(const DemArgs& a, int i, double dt, d3& acc, d3& tq) {
    const double4 v4 = a.so.vel[i];
    if (a.so.dummy[i] == 1 || v4.w < 1.e-10) return;
    const bool cont = (a.flags & DSPH_FLAG_WALL_CONTINUE) != 0;
    const double4 pr = a.so.pr[i];
    const d3 pi = v3(pr), vi = v3(v4), wi = v3(a.so.w[i]);
    const double ri = pr.w;
    const int sc = a.nscene == 1 ? 0 : a.sscene[i];
    const int k0 = a.nscene == 1 ? 0 : a.yoff[sc], k1 = a.nscene == 1 ? a.ny : a.yoff[sc + 1];
    for (int k = k0; k < k1; k++) {
        double* h = a.yhist + 9 * ((int64_t)i * a.nymax + (k - k0));
        const double4 p4 = a.ypt[k];
        const double R = p4.w;
        const d3 pk = v3(p4), ax = v3(a.yax[k]);
        const d3 c0 = pk + ax * dot(pi - pk, ax);
        const d3 out = unit(pi - c0);
        const double d = vlen(pi - c0);
        double delta = d + ri - R;
        bool outside = false;
        if (delta >= ri) {
            delta = ri + R - d;
            outside = true;
        }
        if (delta < 0.) {
            for (int m = 0; m < 9; m++) h[m] = 0.0;
            if (cont) continue;
            break;
        }
        const d3 n = outside ? out : -out;
        const d3 rc = c0 + out * (R - 0.5 * delta);
        const d3 wk = v3(a.yw[k]);
        const d3 vc = vi + cross(wi, rc - pi) - cross(wk, rc - pk);
        const d3 wrel = wi - wk;
        d3 F{0, 0, 0}, T{0, 0, 0};
        Springs sp = ld_springs(h);
        const int c = combined_index(a.tb.nmat, a.so.mat[i], a.ymat[k]);
        contact_model(a.tb, c, F, T, sp, vc, wrel, n, delta, 1. / v4.w, ri, dt);
        st_springs(h, sp);
        acc += F * v4.w;
        tq += T + cross(rc - pi, F);
    }
}

__global__ void k_solid_sum(DemArgs a, double dt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.so.nlive) return;
    const int i = a.so.live[t];
    const double4 pri = a.so.pr[i];
    const d3 pi = v3(pri);
    d3 F{0, 0, 0}, T{0, 0, 0};
    const bool ok = a.L.ssoff[a.ns] <= a.L.ss_cap;
    if (ok) {
        // segments are [off[i], off[i] + cnt[i]): offsets are kept only
        // for solids that can have pairs
        for (int k = a.L.ssoff[i], e = k + a.L.sscnt[i]; k < e; k++) {
            const double4 prj = a.so.pr[a.L.ss_j[k]];
            d3 n;
            double delta;
            const d3 rc = contact_point(pi, v3(prj), pri.w, prj.w, n, delta);
            const d3 f = v3(a.L.ssF[k]);
            F += f;
            T += v3(a.L.ssT[k]) + cross(rc - pi, f);
        }
        for (int k = a.L.ssr_off[i], e = k + a.L.ssr_cnt[i]; k < e; k++) {
            const int k1 = a.L.ssr_idx[k];
            const double4 prj = a.so.pr[a.L.ss_i[k1]];
            d3 n;
            double delta;
            const d3 rc = contact_point(pi, v3(prj), pri.w, prj.w, n, delta);
            const d3 f = v3(a.L.ssF[k1]);
            F -= f;
            T -= v3(a.L.ssT[k1]);
            T -= cross(rc - pi, f);
        }
    }
    const double4 v = a.so.vel[i];
    // fluid forces only matter where 1/m > 0 (F * 0 = 0 otherwise)
    if (v.w > 0.0 && a.L.fsoff && a.L.fsr_off)
        for (int k = a.L.fsr_off[i], e = k + a.L.fsr_cnt[i]; k < e; k++) F -= v3(a.L.fsF[a.L.fsr_idx[k]]);
    const double gon = (v.w > 0.0 && a.so.clump[i] < 0) ? 1.0 : 0.0;
    d3 acc = F * v.w + d3{a.gx, a.gy, a.gz} * gon;
    const double dti = step_of(a.sdt, a.sscene, i, dt);
    if (a.nw > 0) wall_forces(a, i, dti, acc, T);
    if (a.ny > 0) cylinder_forces(a, i, dti, acc, T);
    a.so.acc[i] = d4(acc);
    a.so.tq[i] = d4(T);
}

__global__ void k_clump_sum(DemArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    const d3 pc = v3(a.c.pos[c]);
    d3 F{0, 0, 0}, T{0, 0, 0};
    for (int i = a.c.start[c]; i < a.c.end[c]; i++) {
        const double im = a.so.vel[i].w;
        d3 Fi{0, 0, 0};
        if (im > 1.e-10) Fi = v3(a.so.acc[i]) / im;
        F += Fi;
        T += v3(a.so.tq[i]) + cross(v3(a.so.pr[i]) - pc, Fi);
        a.so.acc[i] = make_double4(0, 0, 0, 0);
        a.so.tq[i] = make_double4(0, 0, 0, 0);
    }
    const double im = a.c.vel[c].w;
    a.c.acc[c] = d4(F * im + d3{a.gx, a.gy, a.gz} * (im > 0. ? 1.0 : 0.0));
    a.c.tq[c] = d4(T);
}

__device__ __forceinline__ void solid_kick_one(const DemArgs& a, int i, double dt) {
    if (a.so.clump[i] >= 0) return;
    dt = step_of(a.sdt, a.sscene, i, dt);
    double4 v = a.so.vel[i];
    const double4 ac = a.so.acc[i];
    v.x += ac.x * dt;
    v.y += ac.y * dt;
    v.z += ac.z * dt;
    a.so.vel[i] = v;
    const double r = a.so.pr[i].w;
    if (v.w < 1.e-10 || r < 1.e-10) return;
    const double I = 0.4 * r * r / v.w;
    const double4 tq = a.so.tq[i];
    double4 w = a.so.w[i];
    w.x += tq.x / I * dt;
    w.y += tq.y / I * dt;
    w.z += tq.z / I * dt;
    a.so.w[i] = w;
}
__global__ void k_solid_kick(DemArgs a, double dt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.so.nlive) return;
    solid_kick_one(a, a.so.live[t], dt);
}

__device__ __forceinline__ void solid_drift_one(const DemArgs& a, int i, double dt) {
    dt = step_of(a.sdt, a.sscene, i, dt);
    double4 p = a.so.pr[i];
    const double4 v = a.so.vel[i];
    p.x += dt * v.x;
    p.y += dt * v.y;
    p.z += dt * v.z;
    a.so.pr[i] = p;
}
__global__ void k_solid_drift(DemArgs a, double dt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.so.nlive) return;
    solid_drift_one(a, a.so.live[t], dt);
}
// without clumps a phase's kick and drift touch only the solid's own state:
// one kernel, kick first (before contact) or drift first (after contact)
__global__ void k_solid_kick_drift(DemArgs a, double dt, int kick_first) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.so.nlive) return;
    const int i = a.so.live[t];
    if (!kick_first) solid_drift_one(a, i, dt);
    solid_kick_one(a, i, dt);
    if (kick_first) solid_drift_one(a, i, dt);
}

// the scene of a clump is its pebbles' (a clump with none has no scene step)
__device__ __forceinline__ double clump_step(const DemArgs& a, int c, double dt) {
    return a.sdt && a.c.start[c] < a.c.end[c] ? step_of(a.sdt, a.sscene, a.c.start[c], dt) : dt;
}

__global__ void k_clump_kick(DemArgs a, double dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    dt = clump_step(a, c, dt);
    const d3 ac = v3(a.c.acc[c]);
    const d3 w0 = v3(a.c.w[c]);
    double4 cv = a.c.vel[c];
    const d3 vc = v3(cv) + ac * dt;
    a.c.vel[c] = d4(vc, cv.w);
    d3 w = w0;
    if (cv.w > 0.) {
        const double4 q4 = a.c.q[c];
        const double* I = a.c.invI + 6 * (int64_t)c;
        const Sym Iw = rotate_inv_inertia(Quat{q4.x, q4.y, q4.z, q4.w},
                                          Sym{I[0], I[1], I[2], I[3], I[4], I[5]});
        w = w0 + mul(Iw, v3(a.c.tq[c])) * dt;
        a.c.w[c] = d4(w);
    }
    const d3 pc = v3(a.c.pos[c]);
    const d3 th = 0.5 * (w0 + w) * dt;
    for (int i = a.c.start[c]; i < a.c.end[c]; i++) {
        const d3 r = v3(a.so.pr[i]) - pc;
        a.so.acc[i] = d4(ac + cross(ac, r) + cross(w, cross(w, r)));
        a.so.vel[i] = d4(vc + cross(w, r), a.so.vel[i].w);
        a.so.w[i] = d4(w);
        a.so.nrm[i] = d4(rodrigues(v3(a.so.nrm[i]), th));
    }
}

__global__ void k_clump_drift(DemArgs a, double dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    dt = clump_step(a, c, dt);
    double4 p = a.c.pos[c];
    const double4 v = a.c.vel[c];
    p.x += dt * v.x;
    p.y += dt * v.y;
    p.z += dt * v.z;
    a.c.pos[c] = p;
    if (a.flags & DSPH_FLAG_CLUMP_QUATERNION) {
        const double4 q4 = a.c.q[c];
        const Quat q = quat_rotate(Quat{q4.x, q4.y, q4.z, q4.w}, v3(a.c.w[c]), dt);
        a.c.q[c] = make_double4(q.q0, q.q1, q.q2, q.q3);
    }
}

// dataBase::addExternalForce (dataBase.cpp:179-195)
__global__ void k_ext_force(DemArgs a, int i, double fx, double fy, double fz) {
    const double im = a.so.vel[i].w;
    double4 ac = a.so.acc[i];
    ac.x += fx * im;
    ac.y += fy * im;
    ac.z += fz * im;
    a.so.acc[i] = ac;
}

// dataBase::addGlobalDamping (dataBase.cpp:197-224)
__global__ void k_damping(DemArgs a, int i, double cd) {
    const double im = a.so.vel[i].w;
    if (im < 1.e-10) return;
    d3 F = v3(a.so.acc[i]) / im;
    d3 T = v3(a.so.tq[i]);
    const d3 v = v3(a.so.vel[i]), w = v3(a.so.w[i]);
    if (vlen(v) > 1.e-10) F += -cd * vlen(F) * unit(v);
    if (vlen(w) > 1.e-10) T += -cd * vlen(T) * unit(w);
    a.so.acc[i] = d4(F * im);
    a.so.tq[i] = d4(T);
}

// dataBase::addBondData, initial branch (dataBase.cpp:270-308): a contact
// of the initial list becomes a bond when both spheres touch physically
// (r_i + r_j + 1e-10 > |x_i - x_j|), belong to one cluster and their
// material pair has a bond modulus.  FLAG pass: one flag per contact;
// FILL pass: the bonds at their scanned offsets, so they keep the contact
// list's order.  FP64 without contraction, as the host loop computes it.
__device__ __forceinline__ bool bond_candidate(const DemArgs& a, int64_t k, int& i, int& j) {
#pragma clang fp contract(off)
    i = a.L.ss_i[k];
    j = a.L.ss_j[k];
    const double4 pi = a.so.pr[i], pj = a.so.pr[j];
    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (pi.w + pj.w + 1.e-10 <= d) return false;
    const int ci = a.so.clus[i], cj = a.so.clus[j];
    const int cij = combined_index(a.tb.nmat, a.so.mat[i], a.so.mat[j]);
    const double Eb = a.tb.bonded[1 * a.tb.ncomb + cij];
    return Eb > 1.e-10 && ci >= 0 && ci == cj;
}

__global__ void k_bond_flags(DemArgs a, const int* __restrict__ total, int* __restrict__ flag) {
    const int64_t n = *total;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        int i, j;
        flag[k] = bond_candidate(a, k, i, j) ? 1 : 0;
    }
}

__global__ void k_bond_fill(DemArgs a, const int* __restrict__ total, const int* __restrict__ off, Bonds b) {
#pragma clang fp contract(off)
    const int64_t n = *total;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        int i, j;
        if (!bond_candidate(a, k, i, j)) continue;
        const int m = off[k];
        const double4 pi = a.so.pr[i], pj = a.so.pr[j];
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const double nx = dx / d, ny = dy / d, nz = dz / d;
        const double delta = pi.w + pj.w - d;
        const double arm = pj.w - 0.5 * delta;
        b.i[m] = i;
        b.j[m] = j;
        b.on[m] = 1;
        b.nrm[m] = make_double4(nx, ny, nz, 0.0);
        b.pt[m] = make_double4(pj.x + arm * nx, pj.y + arm * ny, pj.z + arm * nz, 0.0);
        b.Fs[m] = make_double4(0.0, 0.0, 0.0, 0.0);
        b.Tb[m] = make_double4(0.0, 0.0, 0.0, 0.0);
        b.Fn[m] = 0.0;
        b.Tt[m] = 0.0;
    }
}

int grid_stride_blocks(int64_t cap) {
    int b = blocks_for(cap);
    return b > 2048 ? 2048 : b;
}

}  // namespace

void dem_contacts(hipStream_t s, const DemArgs& a, double dt) {
    if (a.ns <= 0 || a.L.ss_cap <= 0) return;
    hipLaunchKernelGGL(k_ss_contact, dim3(grid_stride_blocks(a.L.ss_cap)), dim3(kBS), 0, s, a, dt);
}
void dem_bonds(hipStream_t s, const DemArgs& a, double dt) {
    if (a.b.n > 0) hipLaunchKernelGGL(k_bonds, dim3(blocks_for(a.b.n)), dim3(kBS), 0, s, a, dt);
}
void dem_solid_sum(hipStream_t s, const DemArgs& a, double dt) {
    if (a.so.nlive > 0) hipLaunchKernelGGL(k_solid_sum, dim3(blocks_for(a.so.nlive)), dim3(kBS), 0, s, a, dt);
}
void dem_solid_kick_drift(hipStream_t s, const DemArgs& a, double dt, bool kick_first) {
    if (a.so.nlive > 0)
        hipLaunchKernelGGL(k_solid_kick_drift, dim3(blocks_for(a.so.nlive)), dim3(kBS), 0, s, a, dt,
                           (int)kick_first);
}
void dem_clump_sum(hipStream_t s, const DemArgs& a) {
    if (a.nc > 0) hipLaunchKernelGGL(k_clump_sum, dim3(blocks_for(a.nc)), dim3(kBS), 0, s, a);
}
void dem_solid_kick(hipStream_t s, const DemArgs& a, double dt) {
    if (a.so.nlive > 0) hipLaunchKernelGGL(k_solid_kick, dim3(blocks_for(a.so.nlive)), dim3(kBS), 0, s, a, dt);
}
void dem_solid_drift(hipStream_t s, const DemArgs& a, double dt) {
    if (a.so.nlive > 0) hipLaunchKernelGGL(k_solid_drift, dim3(blocks_for(a.so.nlive)), dim3(kBS), 0, s, a, dt);
}
void dem_clump_kick(hipStream_t s, const DemArgs& a, double dt) {
    if (a.nc > 0) hipLaunchKernelGGL(k_clump_kick, dim3(blocks_for(a.nc)), dim3(kBS), 0, s, a, dt);
}
void dem_clump_drift(hipStream_t s, const DemArgs& a, double dt) {
    if (a.nc > 0) hipLaunchKernelGGL(k_clump_drift, dim3(blocks_for(a.nc)), dim3(kBS), 0, s, a, dt);
}
void dem_bond_flags(hipStream_t s, const DemArgs& a, int* flag) {
    if (a.ns > 0 && a.L.ss_cap > 0)
        hipLaunchKernelGGL(k_bond_flags, dim3(grid_stride_blocks(a.L.ss_cap)), dim3(kBS), 0, s, a,
                           (const int*)(a.L.ssoff + a.ns), flag);
}
void dem_bond_fill(hipStream_t s, const DemArgs& a, const int* off, const Bonds& b) {
    if (a.ns > 0 && a.L.ss_cap > 0 && b.n > 0)
        hipLaunchKernelGGL(k_bond_fill, dim3(grid_stride_blocks(a.L.ss_cap)), dim3(kBS), 0, s, a,
                           (const int*)(a.L.ssoff + a.ns), off, b);
}

void dem_external_force(hipStream_t s, const DemArgs& a, int idx, double fx, double fy, double fz) {
    hipLaunchKernelGGL(k_ext_force, dim3(1), dim3(1), 0, s, a, idx, fx, fy, fz);
}
void dem_global_damping(hipStream_t s, const DemArgs& a, int idx, double cd) {
    hipLaunchKernelGGL(k_damping, dim3(1), dim3(1), 0, s, a, idx, cd);
}

}  // namespace dsph
