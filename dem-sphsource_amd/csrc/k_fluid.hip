// k_fluid.hip — WCSPH pair loops and fluid time integration.
//
// Follows core/integrate.cu: densityReinitialization (:3-53),
// solveMassConservationEquationDensityIntegrate (:55-161),
// solveMomentumConservationEquation (:163-280), calFluid2SolidForce
// (:282-330), fluidVelocityIntegrate / fluidPositionIntegrate (:367-383).
//
// One thread per fluid particle in slot order walks its neighbour row, then
// adds its fluid–solid terms, which a kernel with one thread per pair stored
// just before (k_fs_terms); there are no atomics.  A row is
// the particle's Verlet superset (all particles within 2h + skin at the last
// rebuild, both directions): candidates outside the kernel support give
// grad W = 0 and are skipped, so the contributing pairs and their order are
// exactly those of the filtered list.  Rows the search walked exactly this
// step (spare-cell cases) are read from the exact list instead (cnt >= 0).
// Two precisions from one template:
//   MIX = false  FP64 throughout (the reference's arithmetic);
//   MIX = true   pair loops in FP32, neighbour sums in FP64;
//                the FP64 masters (position, velocity, dRho) are updated in
//                FP64 and the FP32 records refreshed.
//
// Neighbour records (ping-ponged with dRho so the Jacobi continuity never
// overwrites what neighbours read), two quads per slot, interleaved:
//   FP64: (x, y, z, table index) + (v, dRho), two f64x4 (64 B);
//   MIX:  (x, y, z as 32-bit fixed point, table index) + (v, dRho) as f32x4
//         (32 B).
// The per-particle constants (h, c, m, rho0, nu) live in a small property
// table indexed by the table index (one entry when the scene's fluid is
// uniform, and then not read at all); rho = rho0 + dRho and P = c^2 dRho are
// derived in registers.  Every candidate's records are loaded and its term
// computed, then kept only inside the kernel support (no data-dependent
// branch around the loads).  The fixed-point map spans the domain plus one
// domain length on each side; r_i - r_j is the exact integer difference
// (|d| < 2^24 units converts to FP32 exactly) times the unit, so pair
// geometry never suffers the cancellation of absolute FP32 coordinates.
//
// The Riemann algebra is rearranged (same quantities, fewer operations):
//   grad W_ij = F(r,h) r_ij with F = -105/(16 pi) t^3 / h^5, t = 1 - r/(2h)
//   (d/dq[(1-q/2)^4 (1+2q)] = -5q (1-q/2)^3);
//   v_i - v*_ij = [rho_j (v_i - v_j) + (P_i - P_j)/c e_ij] / (rho_i + rho_j);
//   the viscous term's (v_i - v_j)/|r| (e . grad W) = F (v_i - v_j);
//   2 V_i / m_i = 2 / rho_i.
// Pairs closer than 1e-10 contribute nothing (grad W = 0 there upstream).
#include <algorithm>
#include <type_traits>
#include <utility>

#include "dsph_internal.h"

namespace dsph {

namespace {

constexpr int kBS = 256;     // workgroup of the streaming kernels
constexpr int kMaxBS = 512;  // largest tunable workgroup of the pair loops
constexpr double kGradC = -105.0 / (16.0 * kPi);  // -5 * 21/(16 pi)
constexpr double kWC = 21.0 / (16.0 * kPi);

template <bool MIX> struct Prec { using R = double; using R4 = double4; };
template <> struct Prec<true> { using R = float; using R4 = float4; };

template <class R> __device__ __forceinline__ R rcp_(R x) { return R(1) / x; }
template <> __device__ __forceinline__ float rcp_<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <class R> __device__ __forceinline__ R rsq_(R x) { return R(1) / sqrt(x); }
template <> __device__ __forceinline__ float rsq_<float>(float x) { return __builtin_amdgcn_rsqf(x); }

// Neighbour sums.  FP64 path: plain FP64 sums.  Mixed path: FP32 terms
// summed in FP64 (a convert and an FP64 add per term, no compensation
// needed), so a sum does not depend on how rows are batched or shared.
// add_if(on, x) adds x when on and +0 otherwise, without a branch: the pair
// loops compute every candidate's term and keep it only inside the kernel
// support, so the compiler cannot sink a candidate's record loads behind its
// support test; +0 leaves an FP64 sum unchanged.
template <class R>
struct Acc1 {
    double s = 0.0;
    __device__ __forceinline__ void add(R x) { s += (double)x; }
    __device__ __forceinline__ void add_if(bool on, R x) { s += (double)(on ? x : R(0)); }
    __device__ __forceinline__ double get() const { return s; }
};
template <class R>
struct Acc3 {
    double x = 0.0, y = 0.0, z = 0.0;
    __device__ __forceinline__ void add(R a, R b, R c) {
        x += (double)a;
        y += (double)b;
        z += (double)c;
    }
    __device__ __forceinline__ void add_if(bool on, R a, R b, R c) {
        x += (double)(on ? a : R(0));
        y += (double)(on ? b : R(0));
        z += (double)(on ? c : R(0));
    }
};

// one particle as the pair loops see it; the position is FP64 (FP64 path)
// or fixed point (mixed path).  Its record holds the position, the index of
// its entry in the fluid property table, the velocity and dRho; the table
// gives h, c, m, rho0 (and nu) and the rest follows: rho = rho0 + dRho,
// P = c^2 dRho, V = m / rho, gamma = nu rho.
template <bool MIX>
struct Nb {
    using R = typename Prec<MIX>::R;
    using PT = typename std::conditional<MIX, unsigned, double>::type;
    PT px, py, pz;
    int id;
    R vx, vy, vz, drho;
    R h, c, m, rho0, nu;
    R rho, P;
};

// the record quads of slot j, interleaved: the position quad, then the
// state quad (16 bytes each in the mixed path, f64x4 in the FP64 path), so
// one cache line serves both gathers of a candidate
template <bool MIX>
__device__ __forceinline__ typename Prec<MIX>::R4* pos_w(const FluidSet& f, int cur, int j) {
    return (typename Prec<MIX>::R4*)f.rec[cur] + 2 * (int64_t)j;
}
template <bool MIX>
__device__ __forceinline__ typename Prec<MIX>::R4* state_w(const FluidSet& f, int cur, int j) {
    return (typename Prec<MIX>::R4*)f.rec[cur] + 2 * (int64_t)j + 1;
}

// the position quad (position and table index)
template <bool MIX>
__device__ __forceinline__ void load_pos(const FluidSet& f, int cur, int j, Nb<MIX>& o) {
    if constexpr (MIX) {
        const uint4 a = *(const uint4*)pos_w<MIX>(f, cur, j);
        o.px = a.x; o.py = a.y; o.pz = a.z; o.id = (int)a.w;
    } else {
        const double4 a = *pos_w<MIX>(f, cur, j);
        o.px = a.x; o.py = a.y; o.pz = a.z; o.id = (int)a.w;
    }
}
// the state quad (velocity, dRho)
template <bool MIX>
__device__ __forceinline__ void load_state(const FluidSet& f, int cur, int j, Nb<MIX>& o) {
    const auto b = *state_w<MIX>(f, cur, j);
    o.vx = b.x; o.vy = b.y; o.vz = b.z; o.drho = b.w;
}

// Neighbour records are gathered through a buffer resource over the record
// array: the per-lane address is a 32-bit byte offset (no 64-bit address
// arithmetic per candidate) and an offset past the array reads zeros
// instead of faulting, so candidate indices need no clamping.
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw buffer access
template <bool MIX> constexpr unsigned kRecBytes = MIX ? 32u : 64u;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
using Rsrc = __amdgpu_buffer_rsrc_t;
template <bool MIX>
__device__ __forceinline__ Rsrc rec_rsrc(const FluidSet& f, int cur, int nf) {
    return __builtin_amdgcn_make_buffer_rsrc(f.rec[cur], (short)0, (int)((unsigned)nf * kRecBytes<MIX>),
                                             kRsrcWord3);
}
template <bool MIX, bool STATE>
__device__ __forceinline__ void load_rec(Rsrc rs, int j, Nb<MIX>& o) {
    const unsigned off = (unsigned)j * kRecBytes<MIX>;
    if constexpr (MIX) {
        const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
        o.px = a.x; o.py = a.y; o.pz = a.z; o.id = (int)a.w;
        if constexpr (STATE) {
            const f32x4 b = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16, 0, 0));
            o.vx = b.x; o.vy = b.y; o.vz = b.z; o.drho = b.w;
        }
    } else {
        const f64x2 a0 = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        const f64x2 a1 = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16, 0, 0));
        o.px = a0.x; o.py = a0.y; o.pz = a1.x; o.id = (int)a1.y;
        if constexpr (STATE) {
            const f64x2 b0 = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 32, 0, 0));
            const f64x2 b1 = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 48, 0, 0));
            o.vx = b0.x; o.vy = b0.y; o.vz = b1.x; o.drho = b1.y;
        }
    }
}

// UNI: every fluid particle has the same entry, held by the launch (no load)
template <bool MIX, bool NU, bool UNI>
__device__ __forceinline__ void load_props(const FluidArgs& a, Nb<MIX>& o) {
    using R = typename Prec<MIX>::R;
    if constexpr (UNI) {
        o.h = (R)a.u.h; o.c = (R)a.u.c; o.m = (R)a.u.m; o.rho0 = (R)a.u.rho0;
        if constexpr (NU) o.nu = (R)a.u.nu;
    } else {
        const auto t = ((const typename Prec<MIX>::R4*)a.f.tab)[o.id];
        o.h = t.x; o.c = t.y; o.m = t.z; o.rho0 = t.w;
        if constexpr (NU) o.nu = ((const R*)a.f.tabnu)[o.id];
    }
}
template <bool MIX>
__device__ __forceinline__ void derive(Nb<MIX>& o) {
    o.rho = o.rho0 + o.drho;
    o.P = o.drho * o.c * o.c;
}
template <bool MIX, bool UNI = false>
__device__ __forceinline__ Nb<MIX> load_nb(const FluidArgs& a, int cur, int j) {
    Nb<MIX> o;
    load_pos<MIX>(a.f, cur, j, o);
    load_state<MIX>(a.f, cur, j, o);
    load_props<MIX, true, UNI>(a, o);
    derive<MIX>(o);
    return o;
}

// r_i - r_j in position units: the exact integer difference of the fixed
// point coordinates (mixed; one unit = FluidSet::qr metres on every axis) or
// the FP64 difference in metres
template <bool MIX>
__device__ __forceinline__ void rel(const Nb<MIX>& o, const Nb<MIX>& q, typename Prec<MIX>::R& rx,
                                    typename Prec<MIX>::R& ry, typename Prec<MIX>::R& rz) {
    if constexpr (MIX) {
        rx = (float)(int)(o.px - q.px);
        ry = (float)(int)(o.py - q.py);
        rz = (float)(int)(o.pz - q.pz);
    } else {
        rx = o.px - q.px;
        ry = o.py - q.py;
        rz = o.pz - q.pz;
    }
}
// metres per position unit
template <bool MIX>
__device__ __forceinline__ typename Prec<MIX>::R pos_unit(const FluidSet& f) {
    if constexpr (MIX) return f.qr[0];
    else return 1.0;
}
template <class R> __device__ __forceinline__ R sqrt_(R x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrt_<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// 32-bit fixed point of one coordinate (clamped to the mapped span)
__device__ __forceinline__ unsigned fixq(double x, double origin, double inv_unit) {
    double t = rint((x - origin) * inv_unit);
    t = t < 0.0 ? 0.0 : (t > 4294967295.0 ? 4294967295.0 : t);
    return (unsigned)t;
}

// the position quad of slot s from its FP64 position
template <bool MIX>
__device__ __forceinline__ void store_pos(const FluidSet& f, int buf, int s, const double4& p) {
    auto* r = pos_w<MIX>(f, buf, s);
    const int id = f.tid[s];
    if constexpr (MIX)
        *(uint4*)r = make_uint4(fixq(p.x, f.qo[0], f.qi[0]), fixq(p.y, f.qo[1], f.qi[1]),
                                fixq(p.z, f.qo[2], f.qi[2]), (unsigned)id);
    else
        r[0] = make_double4(p.x, p.y, p.z, (double)id);
}

// the pair record of particle s from its FP64 masters (p = position)
template <bool MIX>
__device__ __forceinline__ void write_records(const FluidSet& f, int buf, int s, const double4& p,
                                              double vx, double vy, double vz, double drho) {
    using R = typename Prec<MIX>::R;
    store_pos<MIX>(f, buf, s, p);
    *state_w<MIX>(f, buf, s) = {(R)vx, (R)vy, (R)vz, (R)drho};
}

// the records of slot s from its records o of the other buffer with a new
// dRho: the position quad as it is, the velocity as it is (o holds the
// record's own FP32 / FP64 values)
template <bool MIX>
__device__ __forceinline__ void renew_records(const FluidSet& f, int buf, int s, const Nb<MIX>& o, double drho) {
    using R = typename Prec<MIX>::R;
    if constexpr (MIX)
        *(uint4*)pos_w<MIX>(f, buf, s) = make_uint4(o.px, o.py, o.pz, (unsigned)o.id);
    else
        *pos_w<MIX>(f, buf, s) = make_double4(o.px, o.py, o.pz, (double)o.id);
    *state_w<MIX>(f, buf, s) = {o.vx, o.vy, o.vz, (R)drho};
}

// the neighbour row of slot s: the exact walked row or the Verlet superset;
// both are ELL column-major with stride nf
struct Row {
    const int* p;  // the row's storage slot
    int n;
    int lg;        // interleave (row_at); walked rows have none
};
__device__ __forceinline__ Row nb_row(const Lists& L, int s) {
    const int c = L.cnt[s];
    if (c >= 0) return {L.nbr + s, c < L.K ? c : L.K, 0};
    const int g = L.grow[s];  // superset rows are shared by a group of slots
    const int r = row_slot(g);
    const int m = L.sup_cnt[r];
    return {L.sup + r, m < L.supK ? m : L.supK, row_lg(g)};
}

// What a pair loop reads of each candidate: its position only (POS), its
// whole state (STATE), or that and its viscosity (STATE_NU).
enum NbMode { POS = 0, STATE = 1, STATE_NU = 2 };

// Visits the row of slot s in batches of U candidates.  The batch's indices
// (column-major, so coalesced over the wavefront) are fetched one batch
// ahead and its U records are gathered before any of them is used: a
// wavefront keeps U record gathers in flight instead of walking one
// dependent index -> record chain per candidate.  Every load is issued
// unconditionally so the loads of a batch overlap: the index reads run up
// to 2U - 1 columns past the row end, into the kRowPad columns every row
// array carries, and the record gathers are buffer loads, safe for any
// index.  fn(q, ok) sees every loaded candidate in row order, ok = false
// past the row end, and must not branch on it (keep its contribution with
// add_if).  The particle itself (a shared row holds its group) is at
// distance 0, outside the support tests of continuity and momentum; in POS
// mode ok is also false for it.
// The table entries of a batch are read once its position quads are in; a
// table shared by many particles is served to the whole wavefront from one
// cache line.
constexpr int kBatch = 4;
static_assert(2 * kBatch <= kRowPad, "row arrays need 2 kBatch pad columns");
// The first batch's indices pass through an identity quad permutation that
// the optimiser cannot see through.  Without it the loads before the loop
// and the loads inside it (same kind, addresses one batch apart) are merged
// into one load at the loop head, which every iteration then waits on
// before its gathers: the index read is no longer a batch ahead.
__device__ __forceinline__ int apart(int x) { return __builtin_amdgcn_mov_dpp(x, 0xE4, 0xF, 0xF, false); }
// column c of a row at its storage slot p
__device__ __forceinline__ int row_col(const int* p, int lg, int c, int64_t st) {
    return __builtin_nontemporal_load(p + (int64_t)(c >> lg) * st + (c & ((1 << lg) - 1)));
}
template <bool MIX, int U, NbMode MODE, bool UNI, bool IL, class Fn>
__device__ __forceinline__ void for_row_each(const FluidArgs& a, Rsrc rs, int s, const Row& row, Fn&& fn) {
    const int n = row.n;
    const int64_t st = a.nf;
    const int* ip = row.p;
    int jn[U];
#pragma unroll
    for (int u = 0; u < U; u++)
        jn[u] = apart(IL ? row_col(row.p, row.lg, u, st) : __builtin_nontemporal_load(ip + u * st));
    for (int k = 0; k < n; k += U) {
        ip += U * st;
        Nb<MIX> q[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            load_rec<MIX, MODE != POS>(rs, jn[u], q[u]);
            // POS (the Shepard sum) counts coincident particles, so it
            // tells the particle itself apart by index
            ok[u] = (k + u < n) & (MODE != POS || jn[u] != s);
        }
        // read once per pass: streamed past the caches the records live in
#pragma unroll
        for (int u = 0; u < U; u++)
            jn[u] = IL ? row_col(row.p, row.lg, k + U + u, st) : __builtin_nontemporal_load(ip + u * st);
        if constexpr (MODE != POS) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                load_props<MIX, MODE == STATE_NU, UNI>(a, q[u]);
                derive<MIX>(q[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) fn(q[u], ok[u]);
    }
}

// Lanes of a wavefront whose slots share one superset row (a group of 2 or
// 4 consecutive slots, aligned in the wavefront) split each batch of four
// candidates: with G lanes sharing, lane l of the group gathers candidates
// k + l (+ G), and every candidate's record is then broadcast to the
// group's lanes by a quad permutation (DPP), so a record crosses the
// texture path once per group instead of once per lane.  Each lane still
// visits candidates k, k + 1, k + 2, k + 3 in row order with the same
// records, so the sums are bitwise those of for_row_each.
typedef int i32x2 __attribute__((ext_vector_type(2)));
template <int C>
__device__ __forceinline__ int bc32(int x) {
    return __builtin_amdgcn_mov_dpp(x, C, 0xF, 0xF, false);
}
template <int C, class T>
__device__ __forceinline__ T bc(T x) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, bc32<C>(__builtin_bit_cast(int, x)));
    } else {
        i32x2 v = __builtin_bit_cast(i32x2, x);
        v.x = bc32<C>(v.x);
        v.y = bc32<C>(v.y);
        return __builtin_bit_cast(T, v);
    }
}
// quad permutation that gives every lane of its group of G the value of the
// group's lane m
template <int G, int M>
constexpr int kQuadFrom = G == 4 ? M * 85 : 160 + 85 * M;
template <int C, bool MIX, bool STATE>
__device__ __forceinline__ void bc_rec(const Nb<MIX>& q, Nb<MIX>& r) {
    r.px = bc<C>(q.px); r.py = bc<C>(q.py); r.pz = bc<C>(q.pz); r.id = bc<C>(q.id);
    if constexpr (STATE) {
        r.vx = bc<C>(q.vx); r.vy = bc<C>(q.vy); r.vz = bc<C>(q.vz); r.drho = bc<C>(q.drho);
    }
}
// candidate M of a batch (row order): gathered by lane M % G of the group
// as its (M / G)-th load; broadcast to the group, its props read
template <int M, int G, bool MIX, NbMode MODE, bool UNI>
__device__ __forceinline__ void split_cand(const FluidArgs& a, const Nb<MIX>* q, const int* jq, int s, int k,
                                           int n, Nb<MIX>& c, bool& ok) {
    constexpr int C = kQuadFrom<G, M % G>;
    bc_rec<C, MIX, MODE != POS>(q[M / G], c);
    ok = k + M < n;
    if constexpr (MODE == POS) ok = ok & (bc<C>(jq[M / G]) != s);
    if constexpr (MODE != POS) {
        load_props<MIX, MODE == STATE_NU, UNI>(a, c);
        derive<MIX>(c);
    }
}
template <int G, int H, bool MIX, NbMode MODE, bool UNI, class Fn, int... M>
__device__ __forceinline__ void split_quarter(const FluidArgs& a, const Nb<MIX>* q, const int* jq, int s, int k,
                                              int n, Fn&& fn, std::integer_sequence<int, M...>) {
    Nb<MIX> c[4];
    bool ok[4];
    (split_cand<4 * H + M, G, MIX, MODE, UNI>(a, q, jq, s, k, n, c[M], ok[M]), ...);
    (fn(c[M], ok[M]), ...);
}
template <int G, bool MIX, NbMode MODE, bool UNI, class Fn, int... H>
__device__ __forceinline__ void split_batch(const FluidArgs& a, const Nb<MIX>* q, const int* jq, int s, int k,
                                            int n, Fn&& fn, std::integer_sequence<int, H...>) {
    (split_quarter<G, H, MIX, MODE, UNI>(a, q, jq, s, k, n, fn, std::make_integer_sequence<int, 4>{}), ...);
}
// B candidates per iteration (a multiple of 4): each lane gathers B / G of
// them, and the next iteration's indices are read ahead, up to 2B - 1
// columns past the batch start
// IL: the group's row is interleaved over its G slots (row_at), so lane l
// reads the dense storage column of slot r + l
template <bool MIX, int G, int B, NbMode MODE, bool UNI, bool IL, class Fn>
__device__ __forceinline__ void for_row_split(const FluidArgs& a, Rsrc rs, int s, const Row& row, Fn&& fn) {
    static_assert(B % 4 == 0 && 2 * B - 1 <= kRowPad, "row arrays need 2B - 1 pad columns");
    constexpr int L = B / G;  // gathers per lane per iteration
    constexpr bool ST = MODE != POS;
    constexpr int CS = IL ? 1 : G;  // storage columns between a lane's candidates
    const int n = row.n;
    const int64_t st = a.nf;
    const int l = threadIdx.x & (G - 1);
    const int* ip = IL ? row.p + l : row.p + (int64_t)l * st;
    int jn[L];
#pragma unroll
    for (int u = 0; u < L; u++) jn[u] = apart(__builtin_nontemporal_load(ip + u * CS * st));
    for (int k = 0; k < n; k += B) {
        ip += (B / G) * CS * st;
        Nb<MIX> q[L];
        int jq[L];
#pragma unroll
        for (int u = 0; u < L; u++) {
            load_rec<MIX, ST>(rs, jn[u], q[u]);
            jq[u] = jn[u];
        }
#pragma unroll
        for (int u = 0; u < L; u++) jn[u] = __builtin_nontemporal_load(ip + u * CS * st);
        split_batch<G, MIX, MODE, UNI>(a, q, jq, s, k, n, fn, std::make_integer_sequence<int, B / 4>{});
    }
}

constexpr int kSplitBatch = 8;
// The row walk of slot s.  Wavefronts whose slots all share superset rows
// in aligned groups of 4 (or 2) take the split walk; any other wavefront
// (walked rows, unshared rows, a partial wavefront) walks every row itself.
template <bool MIX, int U, NbMode MODE, bool UNI, class Fn>
__device__ __forceinline__ void for_row(const FluidArgs& a, Rsrc rs, int s, Fn&& fn) {
    static_assert(U == 4, "the split walk takes batches of four");
    const Row row = nb_row(a.L, s);
    // walked rows are never shared: a key no other lane has
    const int key = a.L.cnt[s] >= 0 ? -1 - s : row_slot(a.L.grow[s]);
    const bool full = __builtin_amdgcn_read_exec() == ~0ull;
    const bool q4 = full && __ballot(key != bc<kQuadFrom<4, 0>>(key)) == 0;
    const bool q2 = full && __ballot(key != bc<kQuadFrom<2, 0>>(key)) == 0;
    if (q4 && __ballot(row.lg != 2) == 0) {
        for_row_split<MIX, 4, kSplitBatch, MODE, UNI, true>(a, rs, s, row, fn);
    } else if (q2 && __ballot(row.lg != 1) == 0) {
        for_row_split<MIX, 2, kSplitBatch, MODE, UNI, true>(a, rs, s, row, fn);
    } else if (__ballot(row.lg != 0) != 0) {
        for_row_each<MIX, U, MODE, UNI, true>(a, rs, s, row, fn);
    } else if (q4) {
        for_row_split<MIX, 4, kSplitBatch, MODE, UNI, false>(a, rs, s, row, fn);
    } else if (q2) {
        for_row_split<MIX, 2, kSplitBatch, MODE, UNI, false>(a, rs, s, row, fn);
    } else {
        for_row_each<MIX, U, MODE, UNI, false>(a, rs, s, row, fn);
    }
}

// --------------------------------------------------------------- tiles
// A tile's workgroup stages the records of its candidates' union in LDS
// once (the position quads, and for continuity and momentum the state
// quads), then every slot walks its local row there: indices into the union
// (u16, ELL column-major), in the order of its global row, in batches of U
// with the next batch's indices read ahead.  An LDS read replaces the
// gather of the global walk.  Position and state quads sit in two arrays of
// 16-byte entries, so consecutive entries fall in consecutive bank groups (a
// 32-byte stride would leave half of them unused and make the lanes' reads
// collide); union position k sits at entry tile_swz(k).
// Entry kTileCap is a sentinel that local rows carry past their end: a
// record at fixed-point (0, 0, 0) at rest with dRho = 0.  Every particle of
// a tiled row lies in the domain, which the fixed-point map places at least
// 2^32 / 3 units from 0 on each axis, so the sentinel is far outside every
// support: its t clamps to 0 and its term is an exact zero (finite
// factors), and the walk needs no row-end test.
using TileRec = uint4;
struct TileLds {
    TileRec pos[kTileCap + 1];
    TileRec st[kTileCap + 1];
};
template <bool STATE>
__device__ __forceinline__ int stage_tile(const FluidArgs& a, int b, TileLds& L) {
    const int m = a.T.ucnt[b];
    const TileRec* rec = (const TileRec*)a.f.rec[a.cur];
    const int* uni = a.T.uni + (int64_t)b * kTileCap;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        const int j = uni[k], e = tile_swz(k);  // the local rows hold entries e
        L.pos[e] = rec[2 * (int64_t)j];
        if constexpr (STATE) L.st[e] = rec[2 * (int64_t)j + 1];
    }
    if (threadIdx.x == 0) {
        L.pos[kTileCap] = make_uint4(0u, 0u, 0u, 0u);
        L.st[kTileCap] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();
    return m;
}
template <bool STATE>
__device__ __forceinline__ void load_tile_rec(const TileLds& L, int j, Nb<true>& o) {
    const uint4 p = L.pos[j];
    o.px = p.x; o.py = p.y; o.pz = p.z; o.id = (int)p.w;
    if constexpr (STATE) {
        const float4 v = __builtin_bit_cast(float4, L.st[j]);
        o.vx = v.x; o.vy = v.y; o.vz = v.z; o.drho = v.w;
    }
}
// quads of a local row: four entries in one 8-byte load, kept two batches
// ahead of use (the rows stream from HBM).  The loop takes two batches per
// iteration, each on a quad register of its own that is reloaded right
// after its batch: a single register pair shifted down every batch would
// wait on the newest load at the top of each batch, one batch ahead only.
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void quad_entries(u32x2 v, int* j) {
    j[0] = (int)(v.x & 0xffffu);
    j[1] = (int)(v.x >> 16);
    j[2] = (int)(v.y & 0xffffu);
    j[3] = (int)(v.y >> 16);
}
__device__ __forceinline__ u32x2 apart2(u32x2 v) {
    v.x = (unsigned)apart((int)v.x);
    v.y = (unsigned)apart((int)v.y);
    return v;
}
static_assert(kRowPad >= 15, "tile rows need three quads read ahead past the last");
template <int U, NbMode MODE, bool UNI, class Fn>
__device__ __forceinline__ void for_row_tile(const FluidArgs& a, const TileLds& lrec, int s, Fn&& fn) {
    static_assert(U == 4, "a batch is one quad of the local row");
    const int n = a.T.rcnt[s];
    const int64_t st = a.nf;
    const u32x2* ip = (const u32x2*)a.T.row + s;
    u32x2 qa = apart2(__builtin_nontemporal_load(ip)), qb = apart2(__builtin_nontemporal_load(ip + st));
    const auto batch = [&](u32x2 qd, int k) {
        int jn[U];
        quad_entries(qd, jn);
        Nb<true> q[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            load_tile_rec<MODE != POS>(lrec, jn[u], q[u]);
            ok[u] = true;  // entries past the row end are the sentinel
        }
        if constexpr (MODE != POS) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                load_props<true, MODE == STATE_NU, UNI>(a, q[u]);
                derive<true>(q[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) fn(q[u], ok[u]);
    };
    // quads up to three past the batch pair read ahead: the row carries zero
    // entries to the end of the third quad after its last
    for (int k = 0; k < n; k += 2 * U) {
        batch(qa, k);
        qa = __builtin_nontemporal_load(ip + 2 * st);
        if (k + U < n) batch(qb, k + U);
        qb = __builtin_nontemporal_load(ip + 3 * st);
        ip += 2 * st;
    }
}

// How a slot's candidates are visited: its global row (superset or exact),
// records gathered from HBM; or, in a tiled launch, its local row in LDS,
// with the global row, walked by the slot alone, for a row walked exactly
// this step or a tile whose union did not fit.  run<MODE>(fn) as for_row.
template <bool MIX, bool UNI>
struct GlobalWalk {
    const FluidArgs& a;
    Rsrc rs;
    int s;
    template <NbMode MODE, class Fn>
    __device__ __forceinline__ void run(Fn&& fn) const {
        for_row<MIX, kBatch, MODE, UNI>(a, rs, s, fn);
    }
};
// U: candidates per batch
template <bool UNI, int U = kBatch>
struct TileWalk {
    const FluidArgs& a;
    Rsrc rs;
    int s;
    const TileLds& lrec;
    bool tiled;
    template <NbMode MODE, class Fn>
    __device__ __forceinline__ void run(Fn&& fn) const {
        if (tiled) {
            for_row_tile<U, MODE, UNI>(a, lrec, s, fn);
        } else {
            const Row row = nb_row(a.L, s);
            for_row_each<true, U, MODE, UNI, true>(a, rs, s, row, fn);
        }
    }
};
// the tile of this workgroup and the slot of this thread (-1: none)
__device__ __forceinline__ int tile_slot(const FluidArgs& a, int b) {
    const int2 tr = a.T.range[b];
    const int s = tr.x + (int)threadIdx.x;
    return s < tr.y ? s : -1;
}

template <class R>
__device__ __forceinline__ R dist(R r2) {
    return r2 > R(0) ? r2 * rsq_(r2) : R(0);
}

// F(r,h) of grad W = F r_ij; 0 outside [1e-10, 2h); ih = 1/h
template <class R>
__device__ __forceinline__ R grad_factor_ih(R r, R h, R ih) {
    const R t = R(1) - R(0.5) * r * ih;
    const R ih2 = ih * ih;
    const R f = R(kGradC) * t * t * t * (ih2 * ih2 * ih);
    return (r >= R(1e-10) && r < R(2) * h) ? f : R(0);
}
template <class R>
__device__ __forceinline__ R grad_factor(R r, R h) {
    return grad_factor_ih<R>(r, h, rcp_(h));
}

// The pair loops' kernel factor t = 1 - r/(2h) of a candidate, set to 0
// where the candidate does not count (keep false: past the row end or
// closer than 1e-10 m) and clamped at 0 outside the support.  Every pair
// term carries t^3, so a candidate that does not count adds an exact zero
// (+0 or -0, which leave an FP64 sum that starts at +0 unchanged) and the
// sums need no select.
template <class R>
__device__ __forceinline__ R support_t(bool keep, R t) {
    return keep ? fmax(t, R(0)) : R(0);
}

// solid-side values of a fluid–solid pair (r_i - r_a subtracted in FP64)
template <class R>
struct SolidP {
    R rx, ry, rz, vx, vy, vz, ax, ay, az, nx, ny, nz, rad;
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

// wall / DEM mirror pressure and density (integrate.cu:139-158, 251-277):
// P_a = P_i + rho_i max(0, (g - a_a).n_a) (r_ia.n_a), rho_a = P_a/c_i^2 + rho0_i
template <class R>
__device__ __forceinline__ void wall_state(R P, R rho, R c, R rho0, const SolidP<R>& s, R gx,
                                           R gy, R gz, R& Pa, R& rhoa) {
    const R gn = (gx - s.ax) * s.nx + (gy - s.ay) * s.ny + (gz - s.az) * s.nz;
    const R rn = s.rx * s.nx + s.ry * s.ny + s.rz * s.nz;
    Pa = P + rho * (gn > R(0) ? gn : R(0)) * rn;
    rhoa = Pa * rcp_(c * c) + rho0;
}

// momentum contribution of a fluid–solid pair per unit fluid mass
template <bool MIX>
__device__ __forceinline__ bool fs_term(const Nb<MIX>& o, typename Prec<MIX>::R rho0,
                                        const SolidP<typename Prec<MIX>::R>& sp,
                                        typename Prec<MIX>::R gx, typename Prec<MIX>::R gy,
                                        typename Prec<MIX>::R gz, typename Prec<MIX>::R& tx,
                                        typename Prec<MIX>::R& ty, typename Prec<MIX>::R& tz) {
    using R = typename Prec<MIX>::R;
    const R r = dist<R>(sp.rx * sp.rx + sp.ry * sp.ry + sp.rz * sp.rz);
    const R F = grad_factor<R>(r, R(0.5) * (o.h + R(2.6) * sp.rad));
    const R d2 = R(2) * sp.rad;
    const R Va = d2 * d2 * d2;
    R Pa, rhoa;
    wall_state<R>(o.P, o.rho, o.c, rho0, sp, gx, gy, gz, Pa, rhoa);
    // U_L = v_i . (-n_a), U_R = (2 v_a - v_i) . (-n_a)
    const R UL = -(o.vx * sp.nx + o.vy * sp.ny + o.vz * sp.nz);
    const R UR = -((R(2) * sp.vx - o.vx) * sp.nx + (R(2) * sp.vy - o.vy) * sp.ny +
                   (R(2) * sp.vz - o.vz) * sp.nz);
    const R dU = UL - UR;
    R beta = R(3) * (dU > R(0) ? dU : R(0));
    beta = beta < o.c ? beta : o.c;
    const R Ps = (o.rho * Pa + rhoa * o.P + o.rho * rhoa * beta * dU) * rcp_(o.rho + rhoa);
    const R pre = R(2) * rcp_(o.rho) * Va * F;
    const R g2 = R(2) * o.nu * o.rho;
    // outside the support (F = 0) the pair's force is zero
    const bool on = F != R(0);
    tx = on ? pre * (g2 * (o.vx - sp.vx) - Ps * sp.rx) : R(0);
    ty = on ? pre * (g2 * (o.vy - sp.vy) - Ps * sp.ry) : R(0);
    tz = on ? pre * (g2 * (o.vz - sp.vz) - Ps * sp.rz) : R(0);
    return on;
}

__device__ __forceinline__ bool fs_valid(const Lists& L, int nf) { return L.fsoff[nf] <= L.fs_cap; }

// Fluid–solid pairs are evaluated one thread per pair (k_fs_terms) before
// each pass, not by the pair loops: pairs crowd at the walls (a wavefront
// beside a wall holds tens of pairs per lane, one in the bulk none), so a
// slot walking its own pairs would keep most lanes of the wall wavefronts
// idle through a chain of dependent gathers.  The pass then adds its slot's
// stored terms [fsoff[s], fsoff[s+1]) in list order, as the slot would have
// added them itself; pairs outside the support store zeros.  Terms are read
// in batches of four (reads past the segment end re-read its last term and
// are not used); fn(term, ok) must not branch on ok.
template <class T, class Fn>
__device__ __forceinline__ void for_fs_terms(const FluidArgs& a, int s, Fn&& fn) {
    const int b = a.L.fsoff[s], e = a.L.fsoff[s + 1];
    const T* __restrict__ t = (const T*)a.L.fsT;
    for (int k = b; k < e; k += 4) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = t[min(k + u, e - 1)];
#pragma unroll
        for (int u = 0; u < 4; u++) fn(v[u], k + u < e);
    }
}

// --------------------------------------------------------------- records
template <bool MIX>
__global__ void __launch_bounds__(kBS) k_records(FluidSet f, int n, int buf) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double4 v = f.vr[s], p = f.ph[s];
    write_records<MIX>(f, buf, s, p, v.x, v.y, v.z, f.drho[buf][s]);
}

// ------------------------------------------------------------- continuity
// The pair sums are the reference's algebra with the reciprocals merged:
//   continuity  2 rho_i V_j F (rho_j dvr + dP r / c) / (rho_i + rho_j)
//             = 2 rho_i m_j F (rho_j c dvr + dP r) / (c rho_j (rho_i + rho_j)),
//   momentum    (2 / rho_i) V_j F [gbar dv - Ps r] with Ps = N / (rho_i + rho_j)
//             = common [(rho_i + rho_j) gbar dv - N r],
//               common = (2 / rho_i) m_j F / (rho_j (rho_i + rho_j)),
// (V_j = m_j / rho_j), so a pair costs one reciprocal besides 1/h and the
// distance.  Pair geometry is in position units (r = u r_u with u metres per
// unit): t = 1 - r/(2h) = 1 + (-u / (2h)) r_u, and u goes into per-particle
// constants with the constant of F.  UNI: all fluid shares one (h, c, m,
// rho0, nu), so 1/h, c and m are constants, dP / c = c (dRho_i - dRho_j) and
// (rho_i + rho_j) gbar = 2 nu rho_i rho_j.  A candidate counts when
// r >= 1e-10 m (r_u^2 >= r2min) and t > 0 (r < 2h).
// drift != 0: the fluid drift that follows the first continuity pass of a
// step (fluidPositionIntegrate, integrate.cu:367-375), x += drift v, done
// here for the slot's own position and written into its new records
template <bool MIX, bool UNI, class Walk>
__device__ __forceinline__ void continuity_slot(const FluidArgs& a, int s, double dt, double drift,
                                                const Walk& walk) {
    using R = typename Prec<MIX>::R;
    if (a.sdt) {  // factors of the scene's own step
        const double k = a.sdt[a.f.scene[s]];
        dt *= k;
        drift *= k;
    }
    const FluidSet& f = a.f;
    const Nb<MIX> o = load_nb<MIX, UNI>(a, a.cur, s);
    const R u = pos_unit<MIX>(f);
    const R r2min = R(1e-20) / (u * u);
    // UNI: term = K t^3 [rho_j dvr_u + (c dRho_i - c dRho_j) r_u] / (rho_j (rho_i + rho_j))
    const R ih = UNI ? R(1) / o.h : R(1);
    const R ih2 = ih * ih;
    const R nh = R(-0.5) * u * ih;
    const R K = R(2) * o.rho * o.m * R(kGradC) * (ih2 * ih2 * ih) * u;
    const R cd_i = o.c * o.drho;
    Acc1<R> acc;
    walk.template run<STATE>(
        [&](const Nb<MIX>& q, bool ok) {
            R rx, ry, rz;
            rel<MIX>(o, q, rx, ry, rz);
            const R r2 = rx * rx + ry * ry + rz * rz;
            const R r = sqrt_(r2);
            const R dvr = (o.vx - q.vx) * rx + (o.vy - q.vy) * ry + (o.vz - q.vz) * rz;
            // mixed: a fixed-point unit is over 1e-10 m, so only coincident
            // records are closer, and their term has dvr = r = 0
            const bool keep = MIX ? ok : (ok & (r2 >= r2min));
            R t, term;
            if constexpr (UNI) {
                t = support_t(keep, R(1) + nh * r);
                const R num = q.rho * dvr + (cd_i - o.c * q.drho) * r;
                term = K * (t * t * t) * num * rcp_(q.rho * (o.rho + q.rho));
            } else {
                const R hb = R(0.5) * (o.h + q.h);
                const R ihb = rcp_(hb);
                t = support_t(keep, R(1) - R(0.5) * u * ihb * r);
                const R ihb2 = ihb * ihb;
                const R F = R(kGradC) * (t * t * t) * (ihb2 * ihb2 * ihb);
                const R c = R(0.5) * (o.c + q.c);
                const R num = q.rho * c * dvr + (o.P - q.P) * r;
                term = R(2) * o.rho * q.m * u * F * num * rcp_(c * q.rho * (o.rho + q.rho));
            }
            acc.add(term);
        });
    if (fs_valid(a.L, a.nf)) for_fs_terms<R>(a, s, [&](R t, bool ok) { acc.add_if(ok, t); });
    const double nd = f.drho[a.cur][s] + dt * acc.get();
    f.drho[1 - a.cur][s] = nd;
    if (drift != 0.0) {
        const double4 vm = f.vr[s];
        double4 p = f.ph[s];
        p.x += drift * vm.x;
        p.y += drift * vm.y;
        p.z += drift * vm.z;
        f.ph[s] = p;
        write_records<MIX>(f, 1 - a.cur, s, p, vm.x, vm.y, vm.z, nd);
    } else {
        // position and velocity are those the slot's records in buffer cur
        // were written from: the new records differ in dRho only
        renew_records<MIX>(f, 1 - a.cur, s, o, nd);
    }
}
template <bool MIX, bool UNI>
__global__ void __launch_bounds__(kMaxBS) k_continuity(FluidArgs a, double dt, double drift) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= a.nf) return;
    continuity_slot<MIX, UNI>(a, s, dt, drift, GlobalWalk<MIX, UNI>{a, rec_rsrc<MIX>(a.f, a.cur, a.nf), s});
}
template <bool UNI>
__global__ void __launch_bounds__(kTileSlots) k_continuity_tile(FluidArgs a, double dt, double drift) {
    __shared__ TileLds lrec;
    const int b = xcd_block();
    const int m = stage_tile<true>(a, b, lrec);
    const int s = tile_slot(a, b);
    if (s < 0) return;
    const bool tiled = m >= 0 && a.L.cnt[s] < 0;
    continuity_slot<true, UNI>(a, s, dt, drift, TileWalk<UNI>{a, rec_rsrc<true>(a.f, a.cur, a.nf), s, lrec, tiled});
}

// --------------------------------------------------------------- momentum
// kick != 0: the kick and drift that follow momentum (fluidVelocityIntegrate
// and fluidPositionIntegrate, integrate.cu:355-375), v += kick a, then
// x += kick/2 v, for the slot's own state; the new records go to the other
// buffer with the slot's dRho (other slots still read this one), and the
// caller flips the current buffer
template <bool MIX, bool UNI, class Walk>
__device__ __forceinline__ void momentum_slot(const FluidArgs& a, int s, double kick,
                                              const Walk& walk) {
    using R = typename Prec<MIX>::R;
    if (a.sdt) kick *= a.sdt[a.f.scene[s]];  // a factor of the scene's own step
    const FluidSet& f = a.f;
    const Nb<MIX> o = load_nb<MIX, UNI>(a, a.cur, s);
    const double4 ps = f.ph[s];
    const R u = pos_unit<MIX>(f);
    const R r2min = R(1e-20) / (u * u);
    const R two_over_rho = R(2) * rcp_(o.rho);
    const R gam_i = o.nu * o.rho;
    // UNI: common = Km t^3 / (rho_j (rho_i + rho_j)), and the pair's force is
    // common (G rho_j dv - uN r_u) with uN = u N =
    //   rho_j (u P_i + u rho_i beta dU) + (u rho_i c^2) dRho_j
    const R ih = UNI ? R(1) / o.h : R(1);
    const R ih2 = ih * ih;
    const R nh = R(-0.5) * u * ih;
    const R Km = two_over_rho * o.m * R(kGradC) * (ih2 * ih2 * ih);
    const R G = R(2) * o.nu * o.rho;
    const R uP = u * o.P;
    const R uB = u * o.rho * o.c * o.c;
    const R uE = u * o.rho;
    // artificial viscosity (SPHKernelDEMContactModel.h:74-77), opt-in: the
    // particle's alpha times m_j h_ij c_ij (v_ij . r_ij) / (rho_ij r^2 +
    // 0.01 h_ij^2) grad W_ij, oriented to remove kinetic energy
    const bool art = (a.flags & DSPH_FLAG_ART_VISCOSITY) != 0;
    const R alpha = art ? (R)f.alpha[s] : R(0);
    Acc3<R> acc;
    // the artificial-viscosity switch is taken once, outside the loop
    auto rows = [&](auto art_tag) {
        constexpr bool ART = decltype(art_tag)::value;
        walk.template run<STATE_NU>(
            [&](const Nb<MIX>& q, bool ok) {
                R rx, ry, rz;
                rel<MIX>(o, q, rx, ry, rz);
                const R r2 = rx * rx + ry * ry + rz * rz;
                const bool keep = ok & (r2 >= r2min);
                // finite for coincident records too (their t is 0)
                const R ir = rsq_(fmax(r2, r2min));
                const R r = r2 * ir;
                const R dvx = o.vx - q.vx, dvy = o.vy - q.vy, dvz = o.vz - q.vz;
                const R dvr = dvx * rx + dvy * ry + dvz * rz;
                // U_L - U_R = -(v_i - v_j) . e_ij (no unit)
                const R dU = -dvr * ir;
                R t, cv, cn, hb, c;
                if constexpr (UNI) {
                    hb = o.h;
                    c = o.c;
                    t = support_t(keep, R(1) + nh * r);
                    const R beta = fmin(R(3) * fmax(dU, R(0)), c);
                    const R common = Km * (t * t * t) * rcp_(q.rho * (o.rho + q.rho));
                    cv = common * G * q.rho;
                    cn = common * (q.rho * (uP + uE * beta * dU) + uB * q.drho);
                } else {
                    hb = R(0.5) * (o.h + q.h);
                    c = R(0.5) * (o.c + q.c);
                    const R ihb = rcp_(hb);
                    t = support_t(keep, R(1) - R(0.5) * u * ihb * r);
                    const R ihb2 = ihb * ihb;
                    const R F = R(kGradC) * (t * t * t) * (ihb2 * ihb2 * ihb);
                    const R beta = fmin(R(3) * fmax(dU, R(0)), c);
                    const R rsum = o.rho + q.rho;
                    const R common = two_over_rho * q.m * F * rcp_(q.rho * rsum);
                    // (rho_i + rho_j) gbar, gbar = 2 gamma_i gamma_j / (gamma_i + gamma_j)
                    const R gam_j = q.nu * q.rho;
                    cv = common * rsum * R(2) * gam_i * gam_j * rcp_(gam_i + gam_j);
                    // N = rho_i P_j + rho_j P_i + rho_i rho_j beta dU
                    cn = common * u * (o.rho * q.P + q.rho * o.P + o.rho * q.rho * beta * dU);
                }
                if constexpr (ART) {
                    // in metres: dvr u, r^2 u^2, grad W = F u r_u
                    const R ihb = rcp_(hb);
                    const R ihb2 = ihb * ihb;
                    const R F = R(kGradC) * (t * t * t) * (ihb2 * ihb2 * ihb);
                    const R rb = R(0.5) * (o.rho + q.rho);
                    cn -= u * q.m * alpha * hb * c * (dvr * u) * rcp_(rb * (r2 * u * u) + R(0.01) * hb * hb) *
                          F;
                }
                acc.add(cv * dvx - cn * rx, cv * dvy - cn * ry, cv * dvz - cn * rz);
            });
    };
    if (art) rows(std::true_type{});
    else rows(std::false_type{});
    if (fs_valid(a.L, a.nf))
        for_fs_terms<typename Prec<MIX>::R4>(
            a, s, [&](const typename Prec<MIX>::R4& t, bool ok) { acc.add_if(ok, t.x, t.y, t.z); });
    const double4 am = make_double4(acc.x + a.gx, acc.y + a.gy, acc.z + a.gz, 0.0);
    f.acc[s] = am;
    // the other buffer gets the slot's records whatever the kick (the caller
    // flips the current buffer)
    double4 v = f.vr[s];
    double4 p = ps;
    if (kick != 0.0) {
        v.x += kick * am.x;
        v.y += kick * am.y;
        v.z += kick * am.z;
        f.vr[s] = v;
        const double hd = 0.5 * kick;
        p.x += hd * v.x;
        p.y += hd * v.y;
        p.z += hd * v.z;
        f.ph[s] = p;
    }
    const double d = f.drho[a.cur][s];
    f.drho[1 - a.cur][s] = d;
    write_records<MIX>(f, 1 - a.cur, s, p, v.x, v.y, v.z, d);
}
template <bool MIX, bool UNI>
__global__ void __launch_bounds__(kMaxBS) k_momentum(FluidArgs a, double kick) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= a.nf) return;
    momentum_slot<MIX, UNI>(a, s, kick, GlobalWalk<MIX, UNI>{a, rec_rsrc<MIX>(a.f, a.cur, a.nf), s});
}
template <bool UNI>
__global__ void __launch_bounds__(kTileSlots) k_momentum_tile(FluidArgs a, double kick) {
    __shared__ TileLds lrec;
    const int b = xcd_block();
    const int m = stage_tile<true>(a, b, lrec);
    const int s = tile_slot(a, b);
    if (s < 0) return;
    const bool tiled = m >= 0 && a.L.cnt[s] < 0;
    momentum_slot<true, UNI>(a, s, kick,
                             TileWalk<UNI>{a, rec_rsrc<true>(a.f, a.cur, a.nf), s, lrec, tiled});
}

// --------------------------------------------------- fluid–solid pair terms
// One thread per pair k of the fluid–solid list (grid-stride over the
// list's device-side length): the pair's term of the coming pass from its
// fluid slot fs_own[k] and solid fs_a[k], stored in fsT[k]:
//   FS_CONT     the continuity term (integrate.cu:103-161), one R;
//   FS_MOM      the momentum term per unit fluid mass (integrate.cu:214-280,
//               calFluid2SolidForce :282-330) as (x, y, z, 0), and with
//               write_fs the pair's force on the solid, m_i times it, in fsF;
//   FS_SHEPARD  the solid's part of the Shepard sum (integrate.cu:3-53).
// The slot's state is that of the pass the terms are for (the same record
// buffer cur, the FP64 master position before the pass moves it).
enum FsMode { FS_CONT = 0, FS_MOM = 1, FS_SHEPARD = 2 };
constexpr int kFsTermBlocks = 2048;  // grid cap: 8 workgroups per CU

// the solid side of pair (s, ai) for the fluid slot whose records are o: in
// the mixed path a static solid comes from its record (r_i - r_a the exact
// fixed-point difference, zero velocity and acceleration), every other
// solid from its FP64 state (r_i - r_a subtracted in FP64)
template <bool MIX>
__device__ __forceinline__ SolidP<typename Prec<MIX>::R> fs_solid(const FluidArgs& a, int s, int ai,
                                                                   const Nb<MIX>& o) {
    using R = typename Prec<MIX>::R;
    if constexpr (MIX) {
        const uint4 q0 = a.so.mrec[2 * (int64_t)ai], q1 = a.so.mrec[2 * (int64_t)ai + 1];
        if (q1.w == 0u) {
            const R u = pos_unit<true>(a.f);
            SolidP<R> sp;
            sp.rx = u * (float)(int)(o.px - q0.x);
            sp.ry = u * (float)(int)(o.py - q0.y);
            sp.rz = u * (float)(int)(o.pz - q0.z);
            sp.vx = sp.vy = sp.vz = R(0);
            sp.ax = sp.ay = sp.az = R(0);
            sp.nx = __uint_as_float(q1.x);
            sp.ny = __uint_as_float(q1.y);
            sp.nz = __uint_as_float(q1.z);
            sp.rad = __uint_as_float(q0.w);
            return sp;
        }
    }
    const double4 ps = a.f.ph[s];
    return load_solid<R>(a.so, ai, ps.x, ps.y, ps.z);
}

template <bool MIX, bool UNI, int MODE>
__global__ void __launch_bounds__(kBS) k_fs_terms(FluidArgs a, int write_fs) {
    using R = typename Prec<MIX>::R;
    using R4 = typename Prec<MIX>::R4;
    if (!fs_valid(a.L, a.nf)) return;
    const FluidSet& f = a.f;
    const int64_t tot = a.L.fsoff[a.nf];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < tot; k += stride) {
        const int s = a.L.fs_own[k];
        const int ai = a.L.fs_a[k];
        if constexpr (MODE == FS_SHEPARD) {
            Nb<MIX> o;
            load_pos<MIX>(f, a.cur, s, o);
            const double hs = UNI ? a.u.h : f.ph[s].w;
            const R h = (R)hs;
            const R ih = rcp_(h);
            const R sigma = R(kWC) * ih * ih * ih;
            const double h13 = hs / 1.3;
            const R invV0 = rcp_((R)(h13 * h13 * h13));
            const SolidP<R> sp = fs_solid<MIX>(a, s, ai, o);
            const R qq = dist<R>(sp.rx * sp.rx + sp.ry * sp.ry + sp.rz * sp.rz) * ih;
            const R t = R(1) - R(0.5) * qq;
            const R t2 = t * t;
            const R d = R(2) * sp.rad;
            ((R*)a.L.fsT)[k] = qq < R(2) ? sigma * t2 * t2 * (R(1) + R(2) * qq) * (d * d * d) * invV0 : R(0);
        } else {
            const Nb<MIX> o = load_nb<MIX, UNI>(a, a.cur, s);
            const SolidP<R> sp = fs_solid<MIX>(a, s, ai, o);
            if constexpr (MODE == FS_CONT) {
                const R r = dist<R>(sp.rx * sp.rx + sp.ry * sp.ry + sp.rz * sp.rz);
                const R F = grad_factor<R>(r, R(0.5) * (o.h + R(2.6) * sp.rad));
                const R d2 = R(2) * sp.rad;
                const R Va = d2 * d2 * d2;
                R Pa, rhoa;
                wall_state<R>(o.P, o.rho, o.c, o.rho0, sp, (R)a.gx, (R)a.gy, (R)a.gz, Pa, rhoa);
                const R dvr = (o.vx - sp.vx) * sp.rx + (o.vy - sp.vy) * sp.ry + (o.vz - sp.vz) * sp.rz;
                const R t = R(2) * o.rho * Va * F * (R(2) * rhoa * dvr + (o.P - Pa) * r * rcp_(o.c)) *
                            rcp_(o.rho + rhoa);
                ((R*)a.L.fsT)[k] = F != R(0) ? t : R(0);
            } else {
                R tx, ty, tz;
                fs_term<MIX>(o, o.rho0, sp, (R)a.gx, (R)a.gy, (R)a.gz, tx, ty, tz);
                ((R4*)a.L.fsT)[k] = R4{tx, ty, tz, R(0)};
                if (write_fs) {
                    const double h13 = (UNI ? a.u.h : f.ph[s].w) / 1.3;
                    const double mi = h13 * h13 * h13 * (UNI ? a.u.rho0 : f.vr[s].w);
                    a.L.fsF[k] = make_double4(mi * tx, mi * ty, mi * tz, 0.0);
                }
            }
        }
    }
}

// Solids::mrec from the solids' state (mixed path).  A solid outside the
// fixed-point map, or with a velocity or acceleration, is flagged; the
// solids the DEM phases move are flagged by the second kernel.
__global__ void __launch_bounds__(kBS) k_solid_mrec(Solids so, int ns, FluidSet f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const double4 p = so.pr[i], v = so.vel[i], ac = so.acc[i], n = so.nrm[i];
    bool in_map = true;
    const double c[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; d++) {
        const double t = (c[d] - f.qo[d]) * f.qi[d];
        in_map = in_map && t >= 0.0 && t <= 4294967295.0;
    }
    const bool moving = v.x != 0.0 || v.y != 0.0 || v.z != 0.0 || ac.x != 0.0 || ac.y != 0.0 || ac.z != 0.0;
    so.mrec[2 * (int64_t)i] = make_uint4(fixq(p.x, f.qo[0], f.qi[0]), fixq(p.y, f.qo[1], f.qi[1]),
                                         fixq(p.z, f.qo[2], f.qi[2]), __float_as_uint((float)p.w));
    so.mrec[2 * (int64_t)i + 1] = make_uint4(__float_as_uint((float)n.x), __float_as_uint((float)n.y),
                                             __float_as_uint((float)n.z), (!in_map || moving) ? 1u : 0u);
}
__global__ void __launch_bounds__(kBS) k_solid_mrec_live(Solids so) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < so.nlive) so.mrec[2 * (int64_t)so.live[t] + 1].w = 1u;
}

// -------------------------------------------------------- density re-init
template <bool MIX, class Walk>
__device__ __forceinline__ void reinit_slot(const FluidArgs& a, int s, int i_step, const Walk& walk) {
    using R = typename Prec<MIX>::R;
    const FluidSet& f = a.f;
    const double4 p = f.ph[s];
    const R h = (R)p.w;
    const R ih = rcp_(h);
    const R sigma = R(kWC) * ih * ih * ih;
    Acc1<R> sw;
    sw.add(sigma);  // W(0, h)
    Nb<MIX> o;
    load_pos<MIX>(f, a.cur, s, o);
    const R ihu = pos_unit<MIX>(f) * ih;  // 1/h per position unit
    walk.template run<POS>(
        [&](const Nb<MIX>& q, bool ok) {
            R rx, ry, rz;
            rel<MIX>(o, q, rx, ry, rz);
            const R qq = sqrt_(rx * rx + ry * ry + rz * rz) * ihu;
            const R t = support_t(ok, R(1) - R(0.5) * qq);
            const R t2 = t * t;
            sw.add(sigma * t2 * t2 * (R(1) + R(2) * qq));
        });
    if (fs_valid(a.L, a.nf)) for_fs_terms<R>(a, s, [&](R t, bool ok) { sw.add_if(ok, t); });
    const double sumw = (double)sw.get();
    if (i_step == 1) {
        f.sumw0[s] = sumw;
    } else {
        const double4 vm = f.vr[s];
        const double rho0 = vm.w;
        const double rs = rho0 + f.drho[a.cur][s];
        const double r1 = rho0 * sumw / f.sumw0[s];
        const double nd = (rs > r1 ? rs : r1) - rho0;
        f.drho[a.cur][s] = nd;
        write_records<MIX>(f, a.cur, s, p, vm.x, vm.y, vm.z, nd);
    }
}
template <bool MIX>
__global__ void __launch_bounds__(kMaxBS) k_density_reinit(FluidArgs a, int i_step) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= a.nf) return;
    reinit_slot<MIX>(a, s, i_step, GlobalWalk<MIX, false>{a, rec_rsrc<MIX>(a.f, a.cur, a.nf), s});
}
__global__ void __launch_bounds__(kTileSlots) k_density_reinit_tile(FluidArgs a, int i_step) {
    __shared__ TileLds lrec;
    const int b = xcd_block();
    const int m = stage_tile<false>(a, b, lrec);
    const int s = tile_slot(a, b);
    if (s < 0) return;
    const bool tiled = m >= 0 && a.L.cnt[s] < 0;
    reinit_slot<true>(a, s, i_step, TileWalk<false>{a, rec_rsrc<true>(a.f, a.cur, a.nf), s, lrec, tiled});
}

}  // namespace

#define DSPH_LAUNCH_BS(kern, n, bs, s, ...)                                                 \
    do {                                                                                    \
        if (mixed)                                                                          \
            hipLaunchKernelGGL(kern<true>, dim3(blocks_for(n, bs)), dim3(bs), 0, s, __VA_ARGS__);  \
        else                                                                                \
            hipLaunchKernelGGL(kern<false>, dim3(blocks_for(n, bs)), dim3(bs), 0, s, __VA_ARGS__); \
    } while (0)
#define DSPH_LAUNCH(kern, n, s, ...) DSPH_LAUNCH_BS(kern, n, kBS, s, __VA_ARGS__)

void fluid_records(hipStream_t s, const FluidArgs& a, int buf, bool mixed) {
    if (a.nf > 0) DSPH_LAUNCH(k_records, a.nf, s, a.f, a.nf, buf);
}

// the fluid–solid pair terms of the coming pass (none without solids)
template <int MODE>
void fs_terms(hipStream_t s, const FluidArgs& a, bool mixed, int write_fs) {
    if (a.nf <= 0 || a.ns <= 0 || a.L.fs_cap <= 0) return;
    const dim3 g((unsigned)std::min<int64_t>((a.L.fs_cap + kBS - 1) / kBS, kFsTermBlocks)), b(kBS);
    if (mixed && a.uni)
        hipLaunchKernelGGL((k_fs_terms<true, true, MODE>), g, b, 0, s, a, write_fs);
    else if (mixed)
        hipLaunchKernelGGL((k_fs_terms<true, false, MODE>), g, b, 0, s, a, write_fs);
    else if (a.uni)
        hipLaunchKernelGGL((k_fs_terms<false, true, MODE>), g, b, 0, s, a, write_fs);
    else
        hipLaunchKernelGGL((k_fs_terms<false, false, MODE>), g, b, 0, s, a, write_fs);
}

void density_reinit(hipStream_t s, const FluidArgs& a, int i_step, bool mixed) {
    if (a.nf <= 0) return;
    fs_terms<FS_SHEPARD>(s, a, mixed, 0);
    if (mixed && a.T.n > 0)
        hipLaunchKernelGGL(k_density_reinit_tile, dim3(a.T.n), dim3(kTileSlots), 0, s, a, i_step);
    else
        DSPH_LAUNCH_BS(k_density_reinit, a.nf, a.bs, s, a, i_step);
}

#define DSPH_LAUNCH_UNI(kern, n, bs, s, ...)                                                     \
    do {                                                                                         \
        const dim3 g(blocks_for(n, bs)), b(bs);                                                  \
        if (mixed && a.uni)                                                                      \
            hipLaunchKernelGGL((kern<true, true>), g, b, 0, s, __VA_ARGS__);                      \
        else if (mixed)                                                                          \
            hipLaunchKernelGGL((kern<true, false>), g, b, 0, s, __VA_ARGS__);                     \
        else if (a.uni)                                                                          \
            hipLaunchKernelGGL((kern<false, true>), g, b, 0, s, __VA_ARGS__);                     \
        else                                                                                     \
            hipLaunchKernelGGL((kern<false, false>), g, b, 0, s, __VA_ARGS__);                    \
    } while (0)

// the tiled launch: one workgroup of kTileSlots threads per tile
#define DSPH_LAUNCH_TILE(kern, s, ...)                                                                \
    do {                                                                                          \
        if (a.uni)                                                                                \
            hipLaunchKernelGGL((kern<true>), dim3(a.T.n), dim3(kTileSlots), 0, s, __VA_ARGS__);   \
        else                                                                                      \
            hipLaunchKernelGGL((kern<false>), dim3(a.T.n), dim3(kTileSlots), 0, s, __VA_ARGS__);  \
    } while (0)

void continuity(hipStream_t s, const FluidArgs& a, double dt, bool mixed, double drift) {
    if (a.nf <= 0) return;
    fs_terms<FS_CONT>(s, a, mixed, 0);
    if (mixed && a.T.n > 0)
        DSPH_LAUNCH_TILE(k_continuity_tile, s, a, dt, drift);
    else
        DSPH_LAUNCH_UNI(k_continuity, a.nf, a.bs, s, a, dt, drift);
}


void momentum(hipStream_t s, const FluidArgs& a, bool mixed, bool write_fs_force, double kick) {
    if (a.nf <= 0) return;
    fs_terms<FS_MOM>(s, a, mixed, (int)write_fs_force);
    if (mixed && a.T.n > 0)
        DSPH_LAUNCH_TILE(k_momentum_tile, s, a, kick);
    else
        DSPH_LAUNCH_UNI(k_momentum, a.nf, a.bs, s, a, kick);
}

void solid_records(hipStream_t s, const Solids& so, int ns, const FluidSet& f) {
    if (ns <= 0 || !so.mrec) return;
    hipLaunchKernelGGL(k_solid_mrec, dim3(blocks_for(ns)), dim3(kBS), 0, s, so, ns, f);
    if (so.nlive > 0) hipLaunchKernelGGL(k_solid_mrec_live, dim3(blocks_for(so.nlive)), dim3(kBS), 0, s, so);
}

// calFluid2SolidForce (integrate.cu:282-330) on steps without a fluid update:
// the pair forces alone
void fluid_solid_force(hipStream_t s, const FluidArgs& a, bool mixed) { fs_terms<FS_MOM>(s, a, mixed, 1); }


}  // namespace dsph
