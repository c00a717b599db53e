// k_fluid.hip — WCSPH pair loops and fluid time integration.
//
// Follows core/integrate.cu: densityReinitialization (:3-53),
// solveMassConservationEquationDensityIntegrate (:55-161),
// solveMomentumConservationEquation (:163-280), calFluid2SolidForce
// (:282-330), fluidVelocityIntegrate / fluidPositionIntegrate (:367-383).
//
// One thread per fluid particle in cell-sorted order walks its full
// neighbour list (ELL, both directions) and its fluid–solid segment; there
// is no per-pair scratch and no atomics.  Two precisions from one template:
//   MIX = false  FP64 throughout (the reference's arithmetic);
//   MIX = true   pair geometry r_i - r_j subtracted in FP64 then rounded,
//                everything else per pair in FP32 with Kahan-compensated
//                neighbour sums; the FP64 masters (position, velocity, dRho)
//                are updated in FP64 and the FP32 shadows refreshed.
//
// The Riemann algebra is rearranged (same quantities, fewer operations):
//   grad W_ij = F(r,h) r_ij with F = -5 sigma t^3 / h^2 = -105/(16 pi) t^3/h^5,
//   t = 1 - r/(2h)  (d/dq[(1-q/2)^4(1+2q)] = -5q(1-q/2)^3);
//   v_i - v*_ij = [rho_j (v_i - v_j) + (P_i - P_j)/c e_ij] / (rho_i + rho_j);
//   the viscous term's (v_i - v_j)/|r| * (e . grad W) = F (v_i - v_j);
//   2 V_i / m_i = 2 / rho_i.
// Pairs closer than 1e-10 contribute nothing (grad W = 0 there upstream).
// The continuity update is Jacobi: neighbours' dRho is read from the buffer
// `cur` and the result written to `1 - cur`.
#include "dsph_internal.h"

namespace dsph {

namespace {

constexpr int kBS = 256;
constexpr double kGradC = -105.0 / (16.0 * kPi);  // -5 * 21/(16 pi)
constexpr double kWC = 21.0 / (16.0 * kPi);

struct KahanF {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// per-particle values of the owner, in the pair-loop precision
template <class R>
struct Owner {
    double px, py, pz;  // FP64 position for the pair difference
    R vx, vy, vz, rho0, drho, rho, c, P, nu, gamma, h, m;
};

template <bool MIX>
__device__ __forceinline__ auto load_fluid(const FluidSet& f, int cur, int i) {
    using R = typename std::conditional<MIX, float, double>::type;
    Owner<R> o;
    const double4 p = f.ph[i];
    o.px = p.x; o.py = p.y; o.pz = p.z;
    if constexpr (MIX) {
        const float4 v = f.vr32[i];
        const float4 q = f.p32[i];  // c, nu, m, h
        o.vx = v.x; o.vy = v.y; o.vz = v.z; o.rho0 = v.w;
        o.drho = f.drho32[cur][i];
        o.c = q.x; o.nu = q.y; o.m = q.z; o.h = q.w;
    } else {
        const double4 v = f.vr[i];
        const double2 cn = f.cn[i];
        o.vx = v.x; o.vy = v.y; o.vz = v.z; o.rho0 = v.w;
        o.drho = f.drho[cur][i];
        o.c = cn.x; o.nu = cn.y; o.h = p.w;
        const double h13 = p.w / 1.3;
        o.m = h13 * h13 * h13 * v.w;
    }
    o.rho = o.rho0 + o.drho;
    o.P = o.drho * o.c * o.c;
    o.gamma = o.nu * o.rho;
    return o;
}

template <class R> __device__ __forceinline__ R rcp_(R x) { return R(1) / x; }
template <> __device__ __forceinline__ float rcp_<float>(float x) { return frcp(x); }

// solid-side values for a fluid–solid pair
template <class R>
struct SolidP {
    R rx, ry, rz;        // r_i - r_a
    R vx, vy, vz;        // v_a
    R ax, ay, az;        // a_a
    R nx, ny, nz;        // n_a
    R rad;
};
template <class R>
__device__ __forceinline__ SolidP<R> load_solid(const Solids& so, int a, double px, double py,
                                                double pz) {
    SolidP<R> s;
    const double4 pr = so.pr[a], v = so.vel[a], ac = so.acc[a], n = so.nrm[a];
    s.rx = (R)(px - pr.x); s.ry = (R)(py - pr.y); s.rz = (R)(pz - pr.z);
    s.vx = (R)v.x; s.vy = (R)v.y; s.vz = (R)v.z;
    s.ax = (R)ac.x; s.ay = (R)ac.y; s.az = (R)ac.z;
    s.nx = (R)n.x; s.ny = (R)n.y; s.nz = (R)n.z;
    s.rad = (R)pr.w;
    return s;
}

// F(r,h) of grad W = F r_ij; 0 outside (1e-10, 2h)
template <class R>
__device__ __forceinline__ R grad_factor(R r, R h, R& ih) {
    ih = rcp_(h);
    const R t = R(1) - R(0.5) * r * ih;
    const R ih2 = ih * ih;
    const R f = R(kGradC) * t * t * t * (ih2 * ih2 * ih);
    return (r >= R(1e-10) && r < R(2) * h) ? f : R(0);
}

// wall / DEM pressure and density mirror (integrate.cu:139-158, 251-277)
template <class R>
struct WallState { R Pa, rhoa; };
template <class R>
__device__ __forceinline__ WallState<R> wall_state(const Owner<R>& o, const SolidP<R>& s, R gx, R gy,
                                                   R gz) {
    const R gn = (gx - s.ax) * s.nx + (gy - s.ay) * s.ny + (gz - s.az) * s.nz;
    const R rn = s.rx * s.nx + s.ry * s.ny + s.rz * s.nz;
    WallState<R> w;
    w.Pa = o.P + o.rho * (gn > R(0) ? gn : R(0)) * rn;
    w.rhoa = w.Pa / (o.c * o.c) + o.rho0;
    return w;
}

template <class R>
struct Acc1 {
    R s = 0;
    __device__ __forceinline__ void add(R x) { s += x; }
    __device__ __forceinline__ R get() const { return s; }
};
template <>
struct Acc1<float> {
    KahanF k;
    __device__ __forceinline__ void add(float x) { k.add(x); }
    __device__ __forceinline__ float get() const { return k.s; }
};

// ------------------------------------------------------------- continuity
template <bool MIX>
__global__ void __launch_bounds__(kBS) k_continuity(FluidArgs a, double dt) {
    using R = typename std::conditional<MIX, float, double>::type;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nf) return;
    const FluidSet& f = a.f;
    const Owner<R> o = load_fluid<MIX>(f, a.cur, s);
    Acc1<R> acc;
    const int n = min(a.L.cnt[s], a.L.K);
    for (int k = 0; k < n; k++) {
        const int j = a.L.nbr[(int64_t)k * a.nf + s];
        const Owner<R> q = load_fluid<MIX>(f, a.cur, j);
        const R rx = (R)(o.px - q.px), ry = (R)(o.py - q.py), rz = (R)(o.pz - q.pz);
        const R r2 = rx * rx + ry * ry + rz * rz;
        R r;
        if constexpr (MIX) r = r2 > 0.f ? r2 * frsq(r2) : 0.f;
        else r = sqrt(r2);
        R ih;
        const R F = grad_factor<R>(r, R(0.5) * (o.h + q.h), ih);
        if (F == R(0)) continue;
        const R c = R(0.5) * (o.c + q.c);
        const R Vj = q.m * rcp_(q.rho);
        const R dvr = (o.vx - q.vx) * rx + (o.vy - q.vy) * ry + (o.vz - q.vz) * rz;
        const R term = R(2) * o.rho * Vj * F * (q.rho * dvr + (o.P - q.P) * r * rcp_(c)) *
                       rcp_(o.rho + q.rho);
        acc.add(term);
    }
    const int b = a.L.fsoff[a.nf] > a.L.fs_cap ? 0 : 1;  // skip if the list overflowed
    if (b) {
        const int e = a.L.fsoff[s + 1];
        for (int k = a.L.fsoff[s]; k < e; k++) {
            const int ai = a.L.fs_a[k];
            const SolidP<R> sp = load_solid<R>(a.so, ai, o.px, o.py, o.pz);
            const R r2 = sp.rx * sp.rx + sp.ry * sp.ry + sp.rz * sp.rz;
            R r;
            if constexpr (MIX) r = r2 > 0.f ? r2 * frsq(r2) : 0.f;
            else r = sqrt(r2);
            const R ha = R(2.6) * sp.rad;
            R ih;
            const R F = grad_factor<R>(r, R(0.5) * (o.h + ha), ih);
            if (F == R(0)) continue;
            const R d2 = R(2) * sp.rad;
            const R Va = d2 * d2 * d2;
            const WallState<R> w = wall_state<R>(o, sp, (R)a.gx, (R)a.gy, (R)a.gz);
            const R dvr = (o.vx - sp.vx) * sp.rx + (o.vy - sp.vy) * sp.ry + (o.vz - sp.vz) * sp.rz;
            const R term = R(2) * o.rho * Va * F *
                           (R(2) * w.rhoa * dvr + (o.P - w.Pa) * r * rcp_(o.c)) *
                           rcp_(o.rho + w.rhoa);
            acc.add(term);
        }
    }
    const double nd = f.drho[a.cur][s] + dt * (double)acc.get();
    f.drho[1 - a.cur][s] = nd;
    if constexpr (MIX) f.drho32[1 - a.cur][s] = (float)nd;
}

// --------------------------------------------------------------- momentum
template <class R>
struct Acc3 {
    R x = 0, y = 0, z = 0;
    __device__ __forceinline__ void add(R a, R b, R c) { x += a; y += b; z += c; }
    __device__ __forceinline__ R gx() const { return x; }
    __device__ __forceinline__ R gy() const { return y; }
    __device__ __forceinline__ R gz() const { return z; }
};
template <>
struct Acc3<float> {
    KahanF x, y, z;
    __device__ __forceinline__ void add(float a, float b, float c) { x.add(a); y.add(b); z.add(c); }
    __device__ __forceinline__ float gx() const { return x.s; }
    __device__ __forceinline__ float gy() const { return y.s; }
    __device__ __forceinline__ float gz() const { return z.s; }
};

// fluid–solid momentum term per unit fluid mass (integrate.cu:251-277)
template <class R>
__device__ __forceinline__ bool fs_term(const Owner<R>& o, const SolidP<R>& sp, R gx, R gy, R gz,
                                        R& tx, R& ty, R& tz) {
    const R r2 = sp.rx * sp.rx + sp.ry * sp.ry + sp.rz * sp.rz;
    R r;
    if constexpr (std::is_same<R, float>::value) r = r2 > 0.f ? r2 * frsq(r2) : 0.f;
    else r = sqrt(r2);
    const R ha = R(2.6) * sp.rad;
    R ih;
    const R F = grad_factor<R>(r, R(0.5) * (o.h + ha), ih);
    if (F == R(0)) return false;
    const R d2 = R(2) * sp.rad;
    const R Va = d2 * d2 * d2;
    const WallState<R> w = wall_state<R>(o, sp, gx, gy, gz);
    // U_L = v_i . (-n_a), U_R = (2 v_a - v_i) . (-n_a)
    const R UL = -(o.vx * sp.nx + o.vy * sp.ny + o.vz * sp.nz);
    const R UR = -((R(2) * sp.vx - o.vx) * sp.nx + (R(2) * sp.vy - o.vy) * sp.ny +
                   (R(2) * sp.vz - o.vz) * sp.nz);
    const R dU = UL - UR;
    R beta = R(3) * (dU > R(0) ? dU : R(0));
    beta = beta < o.c ? beta : o.c;
    const R Ps = (o.rho * w.Pa + w.rhoa * o.P + o.rho * w.rhoa * beta * dU) * rcp_(o.rho + w.rhoa);
    const R pre = R(2) * rcp_(o.rho) * Va * F;
    const R g2 = R(2) * o.gamma;
    tx = pre * (g2 * (o.vx - sp.vx) - Ps * sp.rx);
    ty = pre * (g2 * (o.vy - sp.vy) - Ps * sp.ry);
    tz = pre * (g2 * (o.vz - sp.vz) - Ps * sp.rz);
    return true;
}

template <bool MIX>
__global__ void __launch_bounds__(kBS) k_momentum(FluidArgs a, int write_fs) {
    using R = typename std::conditional<MIX, float, double>::type;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nf) return;
    const FluidSet& f = a.f;
    const Owner<R> o = load_fluid<MIX>(f, a.cur, s);
    const R two_over_rho = R(2) * rcp_(o.rho);
    Acc3<R> acc;
    const int n = min(a.L.cnt[s], a.L.K);
    for (int k = 0; k < n; k++) {
        const int j = a.L.nbr[(int64_t)k * a.nf + s];
        const Owner<R> q = load_fluid<MIX>(f, a.cur, j);
        const R rx = (R)(o.px - q.px), ry = (R)(o.py - q.py), rz = (R)(o.pz - q.pz);
        const R r2 = rx * rx + ry * ry + rz * rz;
        R r, ir;
        if constexpr (MIX) {
            ir = r2 > 0.f ? frsq(r2) : 0.f;
            r = r2 * ir;
        } else {
            r = sqrt(r2);
            ir = r > 0 ? 1.0 / r : 0.0;
        }
        R ih;
        const R F = grad_factor<R>(r, R(0.5) * (o.h + q.h), ih);
        if (F == R(0)) continue;
        const R c = R(0.5) * (o.c + q.c);
        const R Vj = q.m * rcp_(q.rho);
        const R dvx = o.vx - q.vx, dvy = o.vy - q.vy, dvz = o.vz - q.vz;
        // U_L - U_R = -(v_i - v_j) . e_ij
        const R dU = -(dvx * rx + dvy * ry + dvz * rz) * ir;
        R beta = R(3) * (dU > R(0) ? dU : R(0));
        beta = beta < c ? beta : c;
        const R inv_sum = rcp_(o.rho + q.rho);
        const R Ps = (o.rho * q.P + q.rho * o.P + o.rho * q.rho * beta * dU) * inv_sum;
        const R gbar = R(2) * o.gamma * q.gamma * rcp_(o.gamma + q.gamma);
        const R pre = two_over_rho * Vj * F;
        acc.add(pre * (gbar * dvx - Ps * rx), pre * (gbar * dvy - Ps * ry),
                pre * (gbar * dvz - Ps * rz));
    }
    const bool fs_ok = a.L.fsoff[a.nf] <= a.L.fs_cap;
    if (fs_ok) {
        const int e = a.L.fsoff[s + 1];
        for (int k = a.L.fsoff[s]; k < e; k++) {
            const SolidP<R> sp = load_solid<R>(a.so, a.L.fs_a[k], o.px, o.py, o.pz);
            R tx, ty, tz;
            if (fs_term<R>(o, sp, (R)a.gx, (R)a.gy, (R)a.gz, tx, ty, tz)) {
                acc.add(tx, ty, tz);
                if (write_fs) a.L.fsF[k] = make_double4(o.m * tx, o.m * ty, o.m * tz, 0.0);
            } else if (write_fs) {
                a.L.fsF[k] = make_double4(0, 0, 0, 0);
            }
        }
    }
    f.acc[s] = make_double4((double)acc.gx() + a.gx, (double)acc.gy() + a.gy,
                            (double)acc.gz() + a.gz, 0.0);
}

// calFluid2SolidForce (integrate.cu:282-330), one thread per fluid particle
template <bool MIX>
__global__ void __launch_bounds__(kBS) k_fs_force(FluidArgs a) {
    using R = typename std::conditional<MIX, float, double>::type;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nf) return;
    if (a.L.fsoff[a.nf] > a.L.fs_cap) return;
    const int b0 = a.L.fsoff[s], e = a.L.fsoff[s + 1];
    if (b0 == e) return;
    const Owner<R> o = load_fluid<MIX>(a.f, a.cur, s);
    for (int k = b0; k < e; k++) {
        const SolidP<R> sp = load_solid<R>(a.so, a.L.fs_a[k], o.px, o.py, o.pz);
        R tx = 0, ty = 0, tz = 0;
        fs_term<R>(o, sp, (R)a.gx, (R)a.gy, (R)a.gz, tx, ty, tz);
        a.L.fsF[k] = make_double4(o.m * tx, o.m * ty, o.m * tz, 0.0);
    }
}

// -------------------------------------------------------- density re-init
template <bool MIX>
__global__ void __launch_bounds__(kBS) k_density_reinit(FluidArgs a, int i_step) {
    using R = typename std::conditional<MIX, float, double>::type;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nf) return;
    const FluidSet& f = a.f;
    const double4 p = f.ph[s];
    const R h = (R)p.w;
    const R ih = rcp_(h);
    const R sigma = R(kWC) * ih * ih * ih;
    const double h13 = p.w / 1.3;
    const R V0 = (R)(h13 * h13 * h13);
    Acc1<R> sw;
    sw.add(sigma);  // W(0, h)
    const int n = min(a.L.cnt[s], a.L.K);
    for (int k = 0; k < n; k++) {
        const int j = a.L.nbr[(int64_t)k * a.nf + s];
        const double4 q = f.ph[j];
        const R rx = (R)(p.x - q.x), ry = (R)(p.y - q.y), rz = (R)(p.z - q.z);
        const R r2 = rx * rx + ry * ry + rz * rz;
        R r;
        if constexpr (MIX) r = r2 > 0.f ? r2 * frsq(r2) : 0.f;
        else r = sqrt(r2);
        const R qq = r * ih;
        if (qq >= R(2)) continue;
        const R t = R(1) - R(0.5) * qq;
        const R t2 = t * t;
        sw.add(sigma * t2 * t2 * (R(1) + R(2) * qq));
    }
    if (a.L.fsoff[a.nf] <= a.L.fs_cap) {
        const int e = a.L.fsoff[s + 1];
        const R invV0 = rcp_(V0);
        for (int k = a.L.fsoff[s]; k < e; k++) {
            const double4 q = a.so.pr[a.L.fs_a[k]];
            const R rx = (R)(p.x - q.x), ry = (R)(p.y - q.y), rz = (R)(p.z - q.z);
            const R r2 = rx * rx + ry * ry + rz * rz;
            R r;
            if constexpr (MIX) r = r2 > 0.f ? r2 * frsq(r2) : 0.f;
            else r = sqrt(r2);
            const R qq = r * ih;
            if (qq >= R(2)) continue;
            const R t = R(1) - R(0.5) * qq;
            const R t2 = t * t;
            const R d = R(2) * (R)q.w;
            sw.add(sigma * t2 * t2 * (R(1) + R(2) * qq) * (d * d * d) * invV0);
        }
    }
    const double sumw = (double)sw.get();
    if (i_step == 1) {
        f.sumw0[s] = sumw;
    } else {
        const double rho0 = f.vr[s].w;
        const double rs = rho0 + f.drho[a.cur][s];
        const double r1 = rho0 * sumw / f.sumw0[s];
        const double nd = (rs > r1 ? rs : r1) - rho0;
        f.drho[a.cur][s] = nd;
        if constexpr (MIX) f.drho32[a.cur][s] = (float)nd;
    }
}

// ------------------------------------------------------------ integration
__global__ void k_drift(FluidSet f, int n, double dt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double4 p = f.ph[s];
    const double4 v = f.vr[s];
    p.x += dt * v.x;
    p.y += dt * v.y;
    p.z += dt * v.z;
    f.ph[s] = p;
}

// v += dt a (fluidVelocityIntegrate) then x += dt/2 v (fluidPositionIntegrate)
__global__ void k_kick_drift(FluidSet f, int n, double dt, int mixed) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double4 v = f.vr[s];
    const double4 ac = f.acc[s];
    v.x += dt * ac.x;
    v.y += dt * ac.y;
    v.z += dt * ac.z;
    f.vr[s] = v;
    double4 p = f.ph[s];
    const double hd = 0.5 * dt;
    p.x += hd * v.x;
    p.y += hd * v.y;
    p.z += hd * v.z;
    f.ph[s] = p;
    if (mixed) f.vr32[s] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
}

}  // namespace

void density_reinit(hipStream_t s, const FluidArgs& a, int i_step, bool mixed) {
    if (a.nf <= 0) return;
    if (mixed)
        hipLaunchKernelGGL(k_density_reinit<true>, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a, i_step);
    else
        hipLaunchKernelGGL(k_density_reinit<false>, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a, i_step);
}

void continuity(hipStream_t s, const FluidArgs& a, double dt, bool mixed) {
    if (a.nf <= 0) return;
    if (mixed)
        hipLaunchKernelGGL(k_continuity<true>, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a, dt);
    else
        hipLaunchKernelGGL(k_continuity<false>, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a, dt);
}

void fluid_drift(hipStream_t s, const FluidArgs& a, double dt) {
    if (a.nf > 0) hipLaunchKernelGGL(k_drift, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a.f, a.nf, dt);
}

void momentum(hipStream_t s, const FluidArgs& a, bool mixed, bool write_fs_force) {
    if (a.nf <= 0) return;
    if (mixed)
        hipLaunchKernelGGL(k_momentum<true>, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a, (int)write_fs_force);
    else
        hipLaunchKernelGGL(k_momentum<false>, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a, (int)write_fs_force);
}

void fluid_solid_force(hipStream_t s, const FluidArgs& a, bool mixed) {
    if (a.nf <= 0) return;
    if (mixed)
        hipLaunchKernelGGL(k_fs_force<true>, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a);
    else
        hipLaunchKernelGGL(k_fs_force<false>, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a);
}

void fluid_kick_drift(hipStream_t s, const FluidArgs& a, double dt, bool mixed) {
    if (a.nf > 0)
        hipLaunchKernelGGL(k_kick_drift, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a.f, a.nf, dt, (int)mixed);
}

}  // namespace dsph
