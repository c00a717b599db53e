// k_search.hip — hashed-grid neighbour search.
//
// Follows core/neighborSearch.cu (hash :19-36, fluid lists :115-220, solid
// contact lists with spring history :222-312, driver :314-381) and the
// solid-normal estimate of core/dataBase.cpp:348-413.
//
// Bit-exactness: the hash and every membership test are evaluated in FP64
// with FMA contraction off, IEEE division and correctly rounded sqrt, so the
// cell of a particle and the membership of a pair are the reference's.  The
// distance test skips the sqrt when d^2 is outside [R^2(1-1e-14),
// R^2(1+1e-14)] (sqrt_rn is monotone, so the answer is the same).
//
// MI355X design.  The reference grid's cells are sized by the largest
// effective radius of any species (2.2 r of the DEM spheres), so a fluid
// particle would test ~27 x 85 candidates for ~80 neighbours at 1M
// particles.  Fluid lists therefore come from a Verlet superset: pairs within
// 2 max(h_i,h_j) + skin (fluid) and 2 h_i + skin (solid), built by a
// cell-block search only when the largest displacements since the last build
// could have brought a new pair inside the cut-off
// (d_fluid + max(d_fluid, d_solid) >= skin, decided on the device).  Every
// step an exact FP64 filter of the superset yields the reference's lists.
// For in-domain particles the reference's stencil membership is exactly
// geometric, so the sets are identical; particles whose stencil reaches the
// spare (out-of-domain) cell, and every particle while any is out of the
// domain, take the reference's own per-lane stencil walk instead.
//
// The cell-block build puts one wavefront on <= 64 particles of one cell: the
// 27-cell stencil walk is wave-uniform, candidate records are loaded once
// per wave, and a stencil cell is skipped when no lane can reach it (ballot).
// After a build the fluid particles are reordered into cell-sorted slots
// (gather + copy-back, so device pointers never change and a captured HIP
// graph stays valid).
#include <type_traits>

#include "dsph_internal.h"

namespace dsph {

namespace {

constexpr int kBS = 256;

__device__ __forceinline__ bool in_domain(const GridP& g, double x, double y, double z) {
    return g.minx <= x && x < g.maxx && g.miny <= y && y < g.maxy && g.minz <= z && z < g.maxz;
}

// calculateGridPosition (neighborSearch.h:44-49)
__device__ __forceinline__ void grid_pos(const GridP& g, double x, double y, double z, int& ix,
                                         int& iy, int& iz) {
    ix = (int)((x - g.minx) / g.csx);
    iy = (int)((y - g.miny) / g.csy);
    iz = (int)((z - g.minz) / g.csz);
}

// calculatePointHash (neighborSearch.cu:19-36).  A point within rounding of
// the upper bound can truncate to grid index G on an axis; its hash is then
// sent to the spare cell instead of past the end of the cell arrays.
__device__ __forceinline__ int point_hash(const GridP& g, double x, double y, double z) {
    if (in_domain(g, x, y, z)) {
        int ix, iy, iz;
        grid_pos(g, x, y, z, ix, iy, iz);
        const int h = iz * g.gy * g.gx + iy * g.gx + ix;
        return (h >= 0 && h < g.ncell - 1) ? h : g.ncell - 1;
    }
    return g.ncell - 1;
}

// sqrt_rn(|p-q|^2) <= R with the reference's evaluation order:
// ((dx*dx + dy*dy) + dz*dz), no FMA, R - dis >= 0
__device__ __forceinline__ bool within(double px, double py, double pz, double qx, double qy,
                                       double qz, double R) {
#pragma clang fp contract(off)
    const double dx = px - qx, dy = py - qy, dz = pz - qz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double R2 = R * R;
    if (d2 < R2 * 0.99999999999999) return true;
    if (d2 > R2 * 1.00000000000001) return false;
    return R - sqrt(d2) >= 0.0;
}

// inclusive scan of one int per thread across the workgroup (lds: one int
// per wavefront)
__device__ __forceinline__ int block_scan_int(int v, int* lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    if (lane == 63) lds[w] = v;
    __syncthreads();
    int add = 0;
    for (int k = 0; k < w; k++) add += lds[k];
    __syncthreads();
    return v + add;
}

__device__ __forceinline__ void wave_max_atomic(unsigned long long* addr, unsigned long long v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        unsigned long long u = __shfl_xor(v, d, 64);
        v = v > u ? v : u;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(addr, v);
}

// the start of a search: the per-search status words, and on fluid steps
// the near-out-of-domain cell flags (near_out), cleared in one launch
__global__ void k_status_reset(unsigned long long* st, int* near_out, int ncell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // overflow flags are sticky until read; counters are per search
    if (i == ST_OUT_OF_DOMAIN || i == ST_FF_MAX || i == ST_FS_TOTAL || i == ST_SS_TOTAL || i == ST_SUP_MAX ||
        i == ST_NONFINITE)
        st[i] = 0;
    if (near_out)
        for (int c = i; c < ncell; c += gridDim.x * blockDim.x) near_out[c] = 0;
}

// Morton code of a point's sub-cell, b bits per axis: (x - min) / cs is the
// same quotient the cell index truncates, so the sub-cell is its fraction
__device__ __forceinline__ unsigned sub_code(const GridP& g, double x, double y, double z, int b) {
    const double n = (double)(1 << b);
    const double tx = (x - g.minx) / g.csx, ty = (y - g.miny) / g.csy, tz = (z - g.minz) / g.csz;
    const int mx = (1 << b) - 1;
    const int qx = min(max((int)((tx - floor(tx)) * n), 0), mx);
    const int qy = min(max((int)((ty - floor(ty)) * n), 0), mx);
    const int qz = min(max((int)((tz - floor(tz)) * n), 0), mx);
    unsigned code = 0;
    for (int k = 0; k < b; k++)
        code |= (((qx >> k) & 1u) << (3 * k)) | (((qy >> k) & 1u) << (3 * k + 1)) |
                (((qz >> k) & 1u) << (3 * k + 2));
    return code;
}

__device__ __forceinline__ unsigned long long dbits(double d) {
    return (unsigned long long)__double_as_longlong(d);
}

// squared displacement of p since the build position b, in units of the
// scene's skin^2 (a NaN position forces a rebuild)
__device__ __forceinline__ double disp2(const GridP& g, const double4& p, const double4& b) {
    const double dx = p.x - b.x, dy = p.y - b.y, dz = p.z - b.z;
    const double d2 = (dx * dx + dy * dy + dz * dz) * g.iskin2;
    return d2 == d2 ? d2 : 1e300;
}

// largest value of a workgroup (positive doubles order like their bit
// patterns), valid in thread 0
__device__ __forceinline__ unsigned long long block_max(unsigned long long m) {
    __shared__ unsigned long long wmax[kBS / 64];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const unsigned long long u = __shfl_xor(m, d, 64);
        m = u > m ? u : m;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = wmax[w] > m ? wmax[w] : m;
    return m;
}

// Cells within one cell of an out-of-domain particle's position (clamped to
// the grid): every fluid particle within 2 max(h) <= one cell of it lies in
// them, and only those can pair with it under the reference's asymmetric
// spare-cell rule, so only their rows are walked on the reference stencil.
__device__ __forceinline__ void mark_near(const GridP& g, const double4& p, int* __restrict__ near_out) {
    const double tx = (p.x - g.minx) / g.csx, ty = (p.y - g.miny) / g.csy, tz = (p.z - g.minz) / g.csz;
    if (!(tx == tx && ty == ty && tz == tz)) return;  // a NaN position pairs with nothing
    const int cx = (int)fmin(fmax(floor(tx), 0.0), (double)(g.gx - 1));
    const int cy = (int)fmin(fmax(floor(ty), 0.0), (double)(g.gy - 1));
    const int cz = (int)fmin(fmax(floor(tz), 0.0), (double)(g.gz - 1));
    for (int z = max(cz - 1, 0); z <= min(cz + 1, g.gz - 1); z++)
        for (int y = max(cy - 1, 0); y <= min(cy + 1, g.gy - 1); y++)
            for (int x = max(cx - 1, 0); x <= min(cx + 1, g.gx - 1); x++)
                near_out[g.cell0 + (z * g.gy + y) * g.gx + x] = 1;
}

// near_out (fluid steps): marks the cells near out-of-domain particles; a
// particle whose build position (xb) was out of the domain and that is back
// in it is in no superset row and requests a rebuild
__global__ void __launch_bounds__(kBS) k_hash_fluid(Grids G, const double4* __restrict__ ph,
                                                    const int* __restrict__ orig, int n, int sub_bits,
                                                    const unsigned* __restrict__ crank,
                                                    const unsigned char* __restrict__ cspare,
                                                    unsigned* __restrict__ key, unsigned* __restrict__ rank,
                                                    unsigned* __restrict__ hslot,
                                                    const double4* __restrict__ xb,
                                                    unsigned long long* __restrict__ bmax,
                                                    int* __restrict__ near_out,
                                                    unsigned long long* ctl, unsigned long long* st) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    bool out = false, near = false, back = false, bad = false;
    double d2 = 0.0;
    if (s < n) {
        const GridP& g = grid_at(G, fluid_scene(G, s));
        double4 p = ph[s];
        // a NaN/inf position is outside every domain: the spare cell, whose
        // walks skip it (no distance test holds), and it is counted
        bad = !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z));
        const int hl = point_hash(g, p.x, p.y, p.z);
        const int h = g.cell0 + hl;
        const int o = orig[s];
        key[o] = (unsigned)h;
        out = hl == g.ncell - 1;
        const unsigned sub = out ? 0u : sub_code(g, p.x, p.y, p.z, sub_bits / 3);
        rank[o] = (crank[h] << sub_bits) | sub;
        hslot[s] = (unsigned)h;
        near = cspare[h] != 0;
        if (xb) {
            const double4 b = xb[s];
            d2 = disp2(g, p, b);
            back = !out && point_hash(g, b.x, b.y, b.z) == g.ncell - 1;
        }
        if (out && near_out) mark_near(g, p, near_out);
    }
    const unsigned long long m = __ballot(out);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&st[ST_OUT_OF_DOMAIN], (unsigned long long)__popcll(m));
    const unsigned long long mb = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && mb) atomicAdd(&st[ST_NONFINITE], (unsigned long long)__popcll(mb));
    if ((threadIdx.x & 63) == 0 && __ballot(near) != 0) atomicOr(&ctl[CTL_NEAR_SPARE], 1ull);
    if ((threadIdx.x & 63) == 0 && __ballot(back) != 0) atomicOr(&ctl[CTL_FORCE], 1ull);
    // one maximum per workgroup, reduced by the decision (no contended
    // atomic on one word)
    if (xb) {
        const unsigned long long m2 = block_max(dbits(d2));
        if (threadIdx.x == 0) bmax[blockIdx.x] = m2;
    }
}

// calculatePointHash of caller points (interleaved xyz doubles) on one grid
__global__ void __launch_bounds__(kBS) k_hash_points(GridP g, const double* __restrict__ xyz, int n,
                                                     unsigned* __restrict__ key) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double* p = xyz + 3 * (int64_t)s;
    key[s] = (unsigned)point_hash(g, p[0], p[1], p[2]);
}

__global__ void k_hash_solid(Grids G, const double4* __restrict__ pr, int n,
                             unsigned* __restrict__ key, unsigned long long* st) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long out = 0;
    bool bad = false;
    if (s < n) {
        const GridP& g = grid_at(G, solid_scene(G, s));
        double4 p = pr[s];
        bad = !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z));
        const int hl = point_hash(g, p.x, p.y, p.z);
        key[s] = (unsigned)(g.cell0 + hl);
        out = hl == g.ncell - 1;
    }
    unsigned long long m = __ballot(out != 0);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&st[ST_OUT_OF_DOMAIN], (unsigned long long)__popcll(m));
    const unsigned long long mb = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && mb) atomicAdd(&st[ST_NONFINITE], (unsigned long long)__popcll(mb));
}

// --------------------------------------------------- rebuild decision
// rebuild when two particles may have closed the skin: d_f + max(d_f, d_s)
// >= skin, with displacements in skin units (each scene its own skin)
__device__ void decide_rebuild(unsigned long long* ctl, unsigned long long* st, int always, int always_sort) {
    const double df = sqrt(__longlong_as_double((long long)ctl[CTL_DISP_F]));
    const double ds = sqrt(__longlong_as_double((long long)ctl[CTL_DISP_S]));
    const bool move = df + (df > ds ? df : ds) >= 1.0 - 1e-9;
    // out-of-domain particles are in no superset row (a build finds
    // particles by position): the rows near them are walked exactly
    // instead, and one that comes back in requests a rebuild (CTL_FORCE)
    const unsigned long long r = (ctl[CTL_FORCE] != 0 || always || move) ? 1ull : 0ull;
    ctl[CTL_REBUILD] = r;
    // the grids are sorted for a build and for exact stencil walks (rows of
    // cells whose stencil reaches the spare cell, and rows near particles
    // in it: a fluid particle out of the domain raises CTL_NEAR_SPARE too)
    ctl[CTL_SORT] = (r || ctl[CTL_NEAR_SPARE] != 0 || always_sort) ? 1ull : 0ull;
    ctl[CTL_NEAR_SPARE] = 0;
    ctl[CTL_FORCE] = 0;
    ctl[CTL_DISP_F] = 0;
    ctl[CTL_DISP_S] = 0;
    st[ST_REBUILDS] += r;
}

// the live solids' hashes (and, with decide, their displacements and then
// the rebuild decision in the last workgroup to finish)
__global__ void __launch_bounds__(kBS) k_hash_live(Grids G, const double4* __restrict__ pr,
                                                   const int* __restrict__ live, int n,
                                                   unsigned* __restrict__ key, unsigned long long static_out,
                                                   int decide, const double4* __restrict__ xb,
                                                   const unsigned long long* __restrict__ fbmax, int nfb,
                                                   int always, int always_sort, unsigned long long* ctl,
                                                   unsigned long long* st) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool out = false, back = false, bad = false;
    double d2 = 0.0;
    if (t < n) {
        const int i = live[t];
        const GridP& g = grid_at(G, solid_scene(G, i));
        const double4 p = pr[i];
        bad = !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z));
        const int hl = point_hash(g, p.x, p.y, p.z);
        key[i] = (unsigned)(g.cell0 + hl);
        out = hl == g.ncell - 1;
        if (decide) {
            // a solid out of the domain at the build is in no fluid row's
            // solid superset: back in, it requests a rebuild
            const double4 b = xb[i];
            d2 = disp2(g, p, b);
            back = !out && point_hash(g, b.x, b.y, b.z) == g.ncell - 1;
        }
    }
    const unsigned long long m = __ballot(out);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&st[ST_OUT_OF_DOMAIN], (unsigned long long)__popcll(m));
    const unsigned long long mb = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && mb) atomicAdd(&st[ST_NONFINITE], (unsigned long long)__popcll(mb));
    if (t == 0 && static_out) atomicAdd(&st[ST_OUT_OF_DOMAIN], static_out);
    if ((threadIdx.x & 63) == 0 && __ballot(back) != 0) atomicOr(&ctl[CTL_FORCE], 1ull);
    if (!decide) return;
    const unsigned long long ms = block_max(dbits(d2));
    if (threadIdx.x == 0 && ms) atomicMax(&ctl[CTL_DISP_S], ms);
    // every thread's counts and maxima are visible device-wide before this
    // workgroup takes its ticket; the last ticket sees all of them
    __threadfence();
    __syncthreads();
    __shared__ bool last;
    if (threadIdx.x == 0) last = atomicAdd(&ctl[CTL_TICKET], 1ull) == (unsigned long long)gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    // the fluid's largest displacement from the hash kernel's workgroup maxima
    unsigned long long mf = 0;
    for (int b = threadIdx.x; b < nfb; b += blockDim.x) mf = fbmax[b] > mf ? fbmax[b] : mf;
    mf = block_max(mf);
    if (threadIdx.x == 0) {
        ctl[CTL_TICKET] = 0;
        if (mf > ctl[CTL_DISP_F]) ctl[CTL_DISP_F] = mf;
        decide_rebuild(ctl, st, always, always_sort);
    }
}

// --------------------------------------------- sorted-order reorder
template <class R4>
__global__ void k_reorder(FluidSet src, FluidSet dst, const int* __restrict__ perm,
                          int* __restrict__ slot_of, int n, int cur,
                          const unsigned long long* __restrict__ ctl) {
    if (ctl[CTL_REBUILD] == 0) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int o = perm[t];
    int old = slot_of[o];
    dst.ph[t] = src.ph[old];
    dst.vr[t] = src.vr[old];
    dst.drho[cur][t] = src.drho[cur][old];
    dst.cn[t] = src.cn[old];
    dst.sumw0[t] = src.sumw0[old];
    dst.alpha[t] = src.alpha[old];
    dst.scene[t] = src.scene[old];
    dst.acc[t] = src.acc[old];
    dst.orig[t] = o;
    for (int u = 0; u < 2; u++) ((R4*)dst.rec[cur])[2 * t + u] = ((const R4*)src.rec[cur])[2 * old + u];
    dst.tid[t] = src.tid[old];
    slot_of[o] = t;
}

// After a sort (CTL_SORT): the slot of every sorted position with its
// hash.  After a rebuild's reorder (CTL_REBUILD), the reordered set is
// copied back in the same pass.
template <class R4>
__global__ void k_after_sort(FluidSet dst, FluidSet src, int n, int cur, int reorder,
                             const unsigned* __restrict__ sorted, int sub_bits, const int* __restrict__ perm,
                             const int* __restrict__ slot_of, const unsigned* __restrict__ cinv,
                             int* __restrict__ member, unsigned* __restrict__ hslot,
                             const unsigned long long* __restrict__ ctl) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl[CTL_SORT] != 0) {
        if (t < n) {
            const int s = slot_of[perm[t]];
            member[t] = s;
            hslot[s] = cinv[sorted[t] >> sub_bits];
        }
    }
    if (reorder && ctl[CTL_REBUILD] != 0 && t < n) {
        dst.ph[t] = src.ph[t];
        dst.vr[t] = src.vr[t];
        dst.drho[cur][t] = src.drho[cur][t];
        dst.cn[t] = src.cn[t];
        dst.sumw0[t] = src.sumw0[t];
        dst.alpha[t] = src.alpha[t];
        dst.scene[t] = src.scene[t];
        dst.acc[t] = src.acc[t];
        dst.orig[t] = src.orig[t];
        for (int u = 0; u < 2; u++) ((R4*)dst.rec[cur])[2 * t + u] = ((const R4*)src.rec[cur])[2 * t + u];
        dst.tid[t] = src.tid[t];
    }
}

// ------------------------------------------------------- cell blocks
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
// distance from p to the slab [lo, lo + w] along one axis
__device__ __forceinline__ float slab_dist(float p, float lo, float w) {
    return fmaxf(fmaxf(lo - p, p - (lo + w)), 0.0f);
}

__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
// lanes below this one in a wave mask
__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// A wavefront's queue of staged candidates in LDS: up to 2 x 64 entries of
// (x, y, z relative to the box corner, reach^2) and the candidate's index.
// Candidates are appended in walk order, and the membership scan takes them
// 64 at a time, so a scan runs on full batches however few candidates each
// cell contributes.
struct LdsQueue {
    float4* q;
    int* j;
    int n;  // wave-uniform
};
constexpr int kQueue = 128;

// Verlet superset build (setFluidInteractionsKernel's loop, radius + skin).
// One wavefront per chunk of <= 64 consecutive sorted particles of one scene
// (Morton cell order, so a chunk covers a few neighbouring cells), one lane
// per particle.  The wavefront walks the cells of the chunk's bounding box
// widened by the superset stencil half-width, skipping every z slab, y row
// and cell that no lane can reach.  The members of a cell in reach are
// loaded 64 at a time (one coalesced load per lane), and only those inside
// the lanes' bounding box widened by the largest reach enter the
// wavefront's LDS queue: on the reference grid of inflated cells (cell = 2
// max effective radius, DEM spheres included) that box holds about a third
// of the cells' members.  The queue is then read back by every lane at the
// same address, so the membership scan is a broadcast LDS read and a few
// FP32 operations per candidate, with no global-memory latency in its loop.
// Culling by the box removes only candidates no lane can reach, and keeps
// the walk order, so the rows are those of a scan over every member.
// Geometry is FP32 relative to a corner of the box (a few cells away: ~1e-8
// m rounding); membership is d^2 <= R^2 (1 + 2e-4) with R = 2 max(h_i, h_j)
// + skin (2 h_i + skin for solids), evaluated as max((2 h_i + skin)^2,
// (2 h_j + skin)^2).  A superset only has to contain every pair that can
// come within range before the next rebuild, so the margin errs on the
// inclusive side.  Spare-cell particles and spare-cell candidates are left
// out: they are walked exactly every step.
// G1: rows of single slots (group 1, the tiled path's setting): a lane's
// own test decides and the lane writes its own row, with its store offset
// advanced per member (no ballot, no row_at); the rows are the same.
template <bool G1>
__global__ void __launch_bounds__(kBS) k_build(SearchArgs a, const int2* __restrict__ chunks,
                                               int nchunk) {
    if (a.ctl[CTL_REBUILD] == 0) return;
    // the positions of this build: the reference of the next displacement
    // tests (fluid slots, and the solids that can move)
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < a.nf; t += gridDim.x * blockDim.x)
        a.fxb[t] = a.ph[t];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < a.nslive; t += gridDim.x * blockDim.x) {
        const int i = a.slive[t];
        a.sxb[i] = a.spr[i];
    }
    // per wavefront: a fluid queue and a solid queue
    __shared__ float4 lq[2 * kQueue * (kBS / 64)];
    __shared__ int lj[2 * kQueue * (kBS / 64)];
    const int lane = threadIdx.x & 63;
    const int wq = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * 2 * kQueue;
    const int wv = xcd_block() * (kBS / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwv = gridDim.x * (kBS / 64);
    const int supK = a.L.supK, fsupK = a.L.fsupK;
    const float grow = 1.0f + 2e-4f;
    bool ovf = false;
    int cmax = 0;
    for (int b = wv; b < nchunk; b += nwv) {
        const int2 ch = chunks[b];
        const int t0 = __builtin_amdgcn_readfirstlane(ch.x);
        const int t1 = __builtin_amdgcn_readfirstlane(ch.y);
        bool act = t0 + lane < t1;
        const int s = a.fmember[act ? t0 + lane : t0];
        // a chunk never spans two scenes
        const int sc = __builtin_amdgcn_readfirstlane(fluid_scene(a.G, a.fmember[t0]));
        const GridP& g = grid_at(a.G, sc);
        const int spare = g.ncell - 1;
        const int lc = (int)a.hslot[s] - g.cell0;  // the cell in its scene's own indices
        if (act && lc == spare) {  // walked exactly while out of the domain
            a.L.sup_cnt[s] = 0;
            a.L.fsup_cnt[s] = 0;
            a.L.grow[s] = s;
            act = false;
        }
        // Groups of G consecutive lanes share one row, the union of their
        // supersets, kept by the group's first active lane (the writer): the
        // lanes of a group then read the same candidates in the same order,
        // so their record loads coalesce.  A lane's own row is a subset.
        const int G = G1 ? 1 : a.group;
        const int gl = lane & ~(G - 1);
        const unsigned long long gmask = (G == 64 ? ~0ull : ((1ull << G) - 1)) << gl;
        const unsigned long long actm = __ballot(act) & gmask;
        const int wl = actm ? __ffsll((long long)actm) - 1 : lane;
        const bool writer = act && lane == wl;
        const int ws = __shfl(s, wl, 64);
        // a complete group on consecutive slots interleaves its row over
        // their storage (row_at)
        const bool dense = G > 1 && actm == gmask &&
                           (__ballot(act && s == ws + (lane - gl)) & gmask) == gmask;
        const int gcode = dense ? ws | ((31 - __builtin_clz(G)) << kRowSlotBits) : ws;
        const int gxy = g.gx * g.gy;
        const int cz = lc / gxy, rem = lc - cz * gxy;
        const int cy = rem / g.gx, cx = rem - cy * g.gx;
        const int x0 = wave_min(act ? cx : 0x7fffffff);
        if (x0 == 0x7fffffff) continue;  // no lane in a real cell
        const int w = g.sw;
        const int X0 = max(x0 - w, 0), X1 = min(wave_max(act ? cx : -1) + w, g.gx - 1);
        const int Y0 = max(wave_min(act ? cy : 0x7fffffff) - w, 0);
        const int Y1 = min(wave_max(act ? cy : -1) + w, g.gy - 1);
        const int Z0 = max(wave_min(act ? cz : 0x7fffffff) - w, 0);
        const int Z1 = min(wave_max(act ? cz : -1) + w, g.gz - 1);
        // the box corner all FP32 coordinates of this chunk are taken from
        const double ox = g.minx + X0 * g.csx, oy = g.miny + Y0 * g.csy, oz = g.minz + Z0 * g.csz;
        const double4 p = a.ph[s];
        const float px = (float)(p.x - ox), py = (float)(p.y - oy), pz = (float)(p.z - oz);
        const float skin = (float)g.skin;
        const float Ri = 2.0f * (float)p.w + skin;
        const float Ai = Ri * Ri * grow;  // own reach^2: the solid test, and i's side of max
        const float Rf = 2.0f * fmaxf((float)p.w, (float)g.hmax) + skin;  // any fluid pair of i
        const float cullf = act ? Rf * Rf * grow : -1.0f, culls = act ? Ai : -1.0f;
        const float csx = (float)g.csx, csy = (float)g.csy, csz = (float)g.csz;
        const float pad = (float)g.pad;
        // the lanes' bounding box widened by the largest reach of each kind
        // (a margin of 1e-3 over the scan's 1 + 2e-4: a culled candidate is
        // out of every lane's reach)
        const float big = 3.0e38f;
        const float bx0 = wave_minf(act ? px : big), bx1 = wave_maxf(act ? px : -big);
        const float by0 = wave_minf(act ? py : big), by1 = wave_maxf(act ? py : -big);
        const float bz0 = wave_minf(act ? pz : big), bz1 = wave_maxf(act ? pz : -big);
        const float mf = 1.001f * wave_maxf(act ? Rf : 0.0f), ms = 1.001f * wave_maxf(act ? Ri : 0.0f);
        auto in_box = [&](float x, float y, float z, float m) {
            return x >= bx0 - m && x <= bx1 + m && y >= by0 - m && y <= by1 + m && z >= bz0 - m && z <= bz1 + m;
        };
        LdsQueue qf{lq + wq, lj + wq, 0}, qs{lq + wq + kQueue, lj + wq + kQueue, 0};
        int c = 0, cs = 0;
        int64_t wo = s;  // G1: the store offset of the row's next column
        // membership scan of the first nb (<= 64) queued fluid candidates
        // (full: nb = 64)
        auto scan_fluid = [&](int nb, auto full) {
            constexpr bool FULL = decltype(full)::value;
            for (int k = 0; k < (FULL ? 64 : nb); k += 4) {
                // four broadcast reads in flight; entries past nb (within
                // the queue) are read but not used
                float4 q[4];
                int jj[4];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    q[v] = qf.q[k + v];
                    jj[v] = qf.j[k + v];
                }
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float ex = px - q[v].x, ey = py - q[v].y, ez = pz - q[v].z;
                    const float e2 = ex * ex + ey * ey + ez * ez;
                    bool in = act & (jj[v] != s) & (e2 <= (Ai > q[v].w ? Ai : q[v].w));
                    if constexpr (!FULL) in = in & (k + v < nb);
                    if constexpr (G1) {
                        if (in && c < supK) a.L.sup[wo] = jj[v];
                        wo += in ? (int64_t)a.nf : 0;
                        c += in;
                    } else {
                        const bool gin = (__ballot(in) & gmask) != 0;
                        if (writer && gin && c < supK) a.L.sup[row_at(gcode, c, a.nf)] = jj[v];
                        c += gin;
                    }
                }
            }
        };
        auto scan_solid = [&](int nb, auto full) {
            constexpr bool FULL = decltype(full)::value;
            for (int k = 0; k < (FULL ? 64 : nb); k += 4) {
                float4 q[4];
                int aa[4];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    q[v] = qs.q[k + v];
                    aa[v] = qs.j[k + v];
                }
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float ex = px - q[v].x, ey = py - q[v].y, ez = pz - q[v].z;
                    const bool in = (FULL || k + v < nb) & (ex * ex + ey * ey + ez * ez <= culls);
                    if (in && cs < fsupK) a.L.fsup[(int64_t)cs * a.nf + s] = aa[v];
                    cs += in;
                }
            }
        };
        // appends this lane's candidate (keep) in lane order; a full batch
        // of 64 is scanned and the rest moved to the queue's front
        auto push = [&](LdsQueue& Q, bool keep, float4 v, int j, auto&& scan) {
            const unsigned long long km = __ballot(keep);
            if (keep) {
                const int at = Q.n + lanes_below(km);
                Q.q[at] = v;
                Q.j[at] = j;
            }
            Q.n += __popcll(km);
            __builtin_amdgcn_wave_barrier();
            if (Q.n >= 64) {
                scan(64, std::true_type{});
                __builtin_amdgcn_wave_barrier();
                const float4 mv = Q.q[64 + (lane & 63)];
                const int mj = Q.j[64 + (lane & 63)];
                __builtin_amdgcn_wave_barrier();
                if (lane < Q.n - 64) {
                    Q.q[lane] = mv;
                    Q.j[lane] = mj;
                }
                Q.n -= 64;
                __builtin_amdgcn_wave_barrier();
            }
        };
        for (int zc = Z0; zc <= Z1; zc++) {
            const float dz = slab_dist(pz, (zc - Z0) * csz - pad, csz + 2.0f * pad);
            const float dz2 = dz * dz;
            if (__ballot(dz2 <= cullf) == 0) continue;
            for (int yc = Y0; yc <= Y1; yc++) {
                const float dy = slab_dist(py, (yc - Y0) * csy - pad, csy + 2.0f * pad);
                const float dzy2 = dz2 + dy * dy;
                if (__ballot(dzy2 <= cullf) == 0) continue;
                for (int xc = X0; xc <= X1; xc++) {
                    const float dx = slab_dist(px, (xc - X0) * csx - pad, csx + 2.0f * pad);
                    const float d2 = dzy2 + dx * dx;
                    const int h = g.cell0 + zc * gxy + yc * g.gx + xc;
                    if (__ballot(d2 <= cullf) != 0) {
                        const int st0 = __builtin_amdgcn_readfirstlane(a.fcs[h]);
                        const int en = __builtin_amdgcn_readfirstlane(a.fce[h]);
                        for (int base = st0; st0 >= 0 && base < en; base += 64) {
                            const int u = base + lane;
                            bool keep = false;
                            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            int j = 0;
                            if (u < en) {
                                j = a.fmember[u];
                                const double4 q = a.ph[j];
                                const float Rj = 2.0f * (float)q.w + skin;
                                v = make_float4((float)(q.x - ox), (float)(q.y - oy), (float)(q.z - oz),
                                                Rj * Rj * grow);
                                keep = in_box(v.x, v.y, v.z, mf);
                            }
                            push(qf, keep, v, j, scan_fluid);
                        }
                    }
                    if (__ballot(d2 <= culls) != 0) {
                        // the cell's static solids, then its active ones
                        for (int grid = 0; grid < (a.acs ? 2 : 1); grid++) {
                            const int* gcs = grid ? a.acs : a.scs;
                            const int* gce = grid ? a.ace : a.sce;
                            const int st1 = __builtin_amdgcn_readfirstlane(gcs[h]);
                            const int en = __builtin_amdgcn_readfirstlane(gce[h]);
                            for (int base = st1; st1 >= 0 && base < en; base += 64) {
                                const int u = base + lane;
                                bool keep = false;
                                float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                                int ai = 0;
                                if (u < en) {
                                    ai = grid ? a.active[a.aperm[u]] : a.sperm[u];
                                    const double4 q = a.spr[ai];
                                    v = make_float4((float)(q.x - ox), (float)(q.y - oy), (float)(q.z - oz), 0.0f);
                                    keep = in_box(v.x, v.y, v.z, ms);
                                }
                                push(qs, keep, v, ai, scan_solid);
                            }
                        }
                    }
                }
            }
        }
        // what is left in the queues
        if (qf.n > 0) scan_fluid(qf.n, std::false_type{});
        if (qs.n > 0) scan_solid(qs.n, std::false_type{});
        __builtin_amdgcn_wave_barrier();
        if (act) {
            a.L.sup_cnt[s] = c;
            a.L.fsup_cnt[s] = cs;
            a.L.grow[s] = gcode;
        }
        ovf = ovf || c > supK || cs > fsupK;
        cmax = max(cmax, c);
    }
    if (ovf) atomicOr(&a.st[ST_SUP_OVERFLOW], 1ull);
    wave_max_atomic(&a.st[ST_SUP_MAX], (unsigned long long)cmax);
}

// number of the 27 linear stencil offsets around cell (ax, ay, az) that give
// hash h, for a hash h of a real or spare cell (so in [0, ncell)); the
// reference visits a cell once per offset that reaches it
// (neighborSearch.cu:135-146)
__device__ __forceinline__ int stencil_count(const GridP& g, int ax, int ay, int az, int h) {
    const int base = az * g.gy * g.gx + ay * g.gx + ax;
    int m = 0;
    for (int zz = -1; zz <= 1; zz++)
        for (int yy = -1; yy <= 1; yy++)
            for (int xx = -1; xx <= 1; xx++)
                m += (base + zz * g.gy * g.gx + yy * g.gx + xx) == h;
    return m;
}

// Fluid neighbours of slot s as the reference's pair loops see them.  The
// reference lists (i, j), j > i, for each j its stencil walk from i finds
// within range (neighborSearch.cu:150-170), and the loops of particle i run
// over that half plus every (j, i), j < i, that j's walk found
// (integrate.cu:55-161).  Between two particles of real cells the stencil
// relation is symmetric and this is every particle of i's stencil within
// range; with a particle in the spare cell the two directions differ (its
// hash is the spare cell, its stencil is centred on its position), so every
// candidate is kept in its reference direction only.  Candidates: i's
// stencil cells, the spare cell and, for i in the spare cell, the cells whose
// stencils reach the spare cell.  A cell reached by several offsets (grids
// narrower than three cells in x or y) is counted once per offset.
__device__ void ff_walk(const SearchArgs& a, const double4 p, int s, int& c) {
    const GridP& g = grid_at(a.G, fluid_scene(a.G, s));
    const int c0 = g.cell0;  // cells below are the scene's own indices
    const int spare = g.ncell - 1;
    const int hs = (int)a.hslot[s] - c0;
    const int oi = a.forig[s];
    const double hi = p.w;
    int ax, ay, az;
    grid_pos(g, p.x, p.y, p.z, ax, ay, az);
    // candidate k of the walk (in order): the 27 stencil cells, the spare
    // cell, and for a particle in the spare cell the 27 cells whose stencils
    // reach it; -1 outside the grid.  Generated on the fly (no per-lane
    // array): a candidate equal to an earlier one is skipped where the
    // reference's walk would not repeat it (stencil cells only on grids
    // whose offsets collide, the others always).
    auto gen = [&](int k) -> int {
        int h;
        if (k < 27) {
            h = (az + k / 9 - 1) * g.gy * g.gx + (ay + (k / 3) % 3 - 1) * g.gx + (ax + k % 3 - 1);
        } else if (k == 27) {
            h = spare;
        } else {
            if (hs != spare) return -1;
            const int q = k - 28;
            h = spare - ((q / 9 - 1) * g.gy * g.gx + ((q / 3) % 3 - 1) * g.gx + (q % 3 - 1));
        }
        return (h < 0 || h >= g.ncell) ? -1 : h;
    };
    for (int k = 0; k < 55; k++) {
        const int h = gen(k);
        if (h < 0) continue;
        if (k >= 27 || g.collide != 0) {
            bool seen = false;
            for (int k2 = 0; k2 < k && !seen; k2++) seen = gen(k2) == h;
            if (seen) continue;
        }
        const int st0 = a.fcs[c0 + h];
        if (st0 < 0) continue;
        const int mh = stencil_count(g, ax, ay, az, h);  // offsets of i's stencil reaching h
        const bool sym = h != spare && hs != spare;
        for (int t = st0; t < a.fce[c0 + h]; t++) {
            const int j = a.fmember[t];
            if (j == s) continue;
            const double4 q = a.ph[j];
            if (!within(p.x, p.y, p.z, q.x, q.y, q.z, 2.0 * fmax(hi, q.w))) continue;
            int m = mh;
            if (a.forig[j] < oi && !sym) {
                int bx, by, bz;
                grid_pos(g, q.x, q.y, q.z, bx, by, bz);
                m = stencil_count(g, bx, by, bz, hs);  // offsets of j's stencil reaching i
            }
            for (; m > 0; m--) {
                if (c < a.L.K) a.L.nbr[(int64_t)c * a.nf + s] = j;
                c++;
            }
        }
    }
}

// The reference's per-lane stencil walk (neighborSearch.cu:135-214).
// COUNT: fluid neighbours into the ELL row, solid neighbours counted;
// FILL: solid neighbours written to the CSR list.
template <bool FILL>
__device__ void lane_walk(const SearchArgs& a, const double4 p, int s, int& c, int& cs,
                          int64_t& w) {
    const GridP& g = grid_at(a.G, fluid_scene(a.G, s));
    const double hi = p.w;
    if (!FILL) ff_walk(a, p, s, c);
    int ax, ay, az;
    grid_pos(g, p.x, p.y, p.z, ax, ay, az);
    for (int zz = -1; zz <= 1; zz++)
        for (int yy = -1; yy <= 1; yy++)
            for (int xx = -1; xx <= 1; xx++) {
                const int h = (az + zz) * g.gy * g.gx + (ay + yy) * g.gx + (ax + xx);
                if (h < 0 || h >= g.ncell) continue;
                for (int grid = 0; grid < (a.acs ? 2 : 1); grid++) {
                const int st1 = (grid ? a.acs : a.scs)[g.cell0 + h];
                if (st1 < 0) continue;
                const int en1 = (grid ? a.ace : a.sce)[g.cell0 + h];
                for (int t = st1; t < en1; t++) {
                    const int ai = grid ? a.active[a.aperm[t]] : a.sperm[t];
                    const double4 q = a.spr[ai];
                    if (!within(p.x, p.y, p.z, q.x, q.y, q.z, 2.0 * hi)) continue;
                    if (FILL) {
                        if (w < a.L.fs_cap) {
                            a.L.fs_a[w] = ai;
                            a.L.fs_own[w] = s;
                            a.L.fsF[w] = make_double4(0, 0, 0, 0);
                        }
                        w++;
                    } else {
                        cs++;
                    }
                }
                }
            }
}

// The per-step list pass, in two kernels.  Rows that must follow the
// reference stencil this step (walk) get their exact fluid neighbours in
// the ELL list, cnt = count; every other row keeps its superset for the pair
// loops (cnt = -1), which skip candidates outside the kernel support
// themselves.  Fluid–solid pairs are filtered exactly for every row.
// COUNT tests the candidates of each slot, records the members among its
// row's first 128 in a bit mask, and leaves the count of its workgroup's
// tile of slots in tsum.  FILL gives each tile its offset, the sum of the
// earlier tiles' counts (read by the whole workgroup at once, no waiting on
// other workgroups), scans the tile, and writes the CSR: offsets, then
// pairs.
__global__ void __launch_bounds__(kBS) k_filter_count(SearchArgs a) {
    __shared__ int lds[kBS / 64];
    const int tile = xcd_block();
    const int s = tile * blockDim.x + threadIdx.x;
    int c = 0, cs = 0;
    if (s < a.nf) {
        const double4 p = a.ph[s];
        const GridP& g = grid_at(a.G, fluid_scene(a.G, s));
        const unsigned spare = (unsigned)(g.cell0 + g.ncell - 1);  // the scene's spare cell
        const unsigned hs = a.hslot[s];
        const bool walk = g.walk_only || hs == spare || a.cspare[hs] != 0 || a.near_out[hs] != 0;
        unsigned long long m0 = 0, m1 = 0;
        if (walk) {
            int64_t w0 = 0;
            lane_walk<false>(a, p, s, c, cs, w0);
        } else {
            const int n2 = min(a.L.fsup_cnt[s], a.L.fsupK);
            const int* __restrict__ fsup = a.L.fsup + s;
            const int last = n2 > 0 ? n2 - 1 : 0;
            // candidates in batches of 8: indices, then positions, all in
            // flight together (entries past the row end re-read its last one
            // and are not used).  A row not walked has no stencil cell that
            // is the spare cell, so a solid that has left the domain since
            // the build is not its pair; otherwise the distance decides.
            constexpr int U = 8;
            for (int k = 0; k < n2; k += U) {
                int ai[U];
                double4 q[U];
                unsigned sk[U];
#pragma unroll
                for (int u = 0; u < U; u++) ai[u] = fsup[(int64_t)min(k + u, last) * a.nf];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    q[u] = a.spr[ai[u]];
                    sk[u] = a.skey[ai[u]];
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const bool in = (k + u < n2) & (sk[u] != spare) &
                                    within(p.x, p.y, p.z, q[u].x, q[u].y, q[u].z, 2.0 * p.w);
                    cs += in;
                    const int kk = k + u;
                    const unsigned long long bit = in ? 1ull << (kk & 63) : 0ull;
                    if (kk < 64) m0 |= bit;
                    else if (kk < 128) m1 |= bit;
                }
            }
        }
        a.L.cnt[s] = walk ? c : -1;
        a.L.fscnt[s] = cs;
        a.L.fsm[2 * (int64_t)s] = m0;
        a.L.fsm[2 * (int64_t)s + 1] = m1;
        a.fsp[s] = p;
    }
    wave_max_atomic(&a.st[ST_FF_MAX], (unsigned long long)c);
    if (c > a.L.K) atomicOr(&a.st[ST_FF_OVERFLOW], 1ull);
    const int inc = block_scan_int(cs, lds);
    if (threadIdx.x == blockDim.x - 1) a.tsum[tile] = inc;
}

// Pairs of one wavefront's slots are contiguous in the list: a wavefront
// whose pairs fit kFillBuf gathers them in LDS (each lane its own members
// from its mask, with its lane number as the pair's owner) and then stores
// them with consecutive lanes on consecutive entries; the others, and
// walked rows, store lane by lane.  Each pair's fluid slot goes to fs_own.
// 2048 pairs per wavefront holds the wavefronts beside a wall (a floor
// wavefront has about 20 pairs per lane); the buffer sets the kernel's
// occupancy (40 KiB per workgroup: 1024 pairs, 3072 and 4096 were slower).
constexpr int kFillBuf = 2048;
__global__ void __launch_bounds__(kBS) k_filter_fill(SearchArgs a, int zero_fs) {
    __shared__ int lds[kBS / 64];
    __shared__ long long part[kBS / 64];
    __shared__ int buf[kBS / 64][kFillBuf];
    __shared__ unsigned char obuf[kBS / 64][kFillBuf];
    const int tile = xcd_block();
    const int s = tile * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // the tile's offset: the earlier tiles' counts, summed by the workgroup
    long long pre = 0;
#pragma unroll 8
    for (int b = threadIdx.x; b < tile; b += blockDim.x) pre += a.tsum[b];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) pre += __shfl_xor(pre, d, 64);
    if (lane == 0) part[threadIdx.x >> 6] = pre;
    const bool live = s < a.nf;
    const int cs = live ? a.L.fscnt[s] : 0;
    const int inc = block_scan_int(cs, lds);  // its barriers also publish part[]
    long long base = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) base += part[w];
    int64_t w = base + inc - cs;
    if (live) {
        a.L.fsoff[s] = (int)w;
        if (s == a.nf - 1) {
            const int64_t tot = w + cs;
            a.L.fsoff[a.nf] = (int)tot;
            a.st[ST_FS_TOTAL] = (unsigned long long)tot;
            if (tot > a.L.fs_cap) atomicOr(&a.st[ST_FS_OVERFLOW], 1ull);
        }
    }
    const bool walked = live && cs > 0 && a.L.cnt[s] >= 0;
    const int n2 = (live && cs > 0 && !walked) ? min(a.L.fsup_cnt[s], a.L.fsupK) : 0;
    // the wavefront's pairs: [w0, w0 + T)
    const int64_t w0 = __shfl(w, 0, 64);
    const int64_t T = __shfl(w + cs, 63, 64) - w0;
    const bool staged = T > 0 && T <= kFillBuf && __ballot(walked || n2 > 128) == 0;
    if (walked) {  // its pairs from the stencil walk
        const double4 p = a.ph[s];
        int c2 = 0, cs2 = 0;
        lane_walk<true>(a, p, s, c2, cs2, w);
    }
    int* const wb = buf[threadIdx.x >> 6];
    unsigned char* const ob = obuf[threadIdx.x >> 6];
    if (n2 > 0) {
        const int* __restrict__ fsup = a.L.fsup + s;
        const unsigned long long m0 = a.L.fsm[2 * (int64_t)s], m1 = a.L.fsm[2 * (int64_t)s + 1];
        const int k0 = min(n2, 128);
        constexpr int V = 32;
        for (int k = 0; k < k0; k += V) {
            int ai[V];
#pragma unroll
            for (int u = 0; u < V; u++) ai[u] = fsup[(int64_t)min(k + u, k0 - 1) * a.nf];
#pragma unroll
            for (int u = 0; u < V; u++) {
                const int kk = k + u;
                const bool in = kk < k0 && (((kk < 64 ? m0 : m1) >> (kk & 63)) & 1ull);
                if (in) {
                    if (staged) {
                        wb[w - w0] = ai[u];
                        ob[w - w0] = (unsigned char)lane;
                    } else if (w < a.L.fs_cap) {
                        a.L.fs_a[w] = ai[u];
                        a.L.fs_own[w] = s;
                    }
                }
                w += in;
            }
        }
        // members past the first 128 candidates are tested again (never
        // staged)
        if (n2 > k0) {
            const double4 p = a.ph[s];
            for (int k = k0; k < n2; k++) {
                const int ai = fsup[(int64_t)k * a.nf];
                const double4 q = a.spr[ai];
                if (within(p.x, p.y, p.z, q.x, q.y, q.z, 2.0 * p.w)) {
                    if (w < a.L.fs_cap) {
                        a.L.fs_a[w] = ai;
                        a.L.fs_own[w] = s;
                    }
                    w++;
                }
            }
        }
    }
    if (staged) {
        // the lanes' LDS writes are complete and visible to the wavefront
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int i = lane; i < T; i += 64)
            if (w0 + i < a.L.fs_cap) {
                a.L.fs_a[w0 + i] = wb[i];
                a.L.fs_own[w0 + i] = s - lane + ob[i];
            }
    }
    // new pairs start with zero force where asked
    if (zero_fs && T > 0)
        for (int64_t i = w0 + lane; i < w0 + T; i += 64)
            if (i < a.L.fs_cap) a.L.fsF[i] = make_double4(0, 0, 0, 0);
}

// Exact fluid neighbour rows of every particle for read-back, into a
// separate ELL buffer of the superset's capacity (an exact row is a subset
// of its superset row): walked rows are copied, superset rows filtered with
// the reference's membership test on the positions of the last search.
// Pair-loop state is not touched.
__global__ void __launch_bounds__(kBS) k_exact_rows(SearchArgs a, int* __restrict__ out,
                                                    int* __restrict__ out_cnt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nf) return;
    const int cw = a.L.cnt[s];
    int c = 0;
    if (cw > a.L.K) {
        c = -1;  // the walked row overflowed its capacity
    } else if (cw >= 0) {
        for (int k = 0; k < cw; k++) out[(int64_t)k * a.nf + s] = a.L.nbr[(int64_t)k * a.nf + s];
        c = cw;
    } else {
        const double4 p = a.fsp[s];
        const int g = a.L.grow[s];  // the group's row
        const int n = min(a.L.sup_cnt[row_slot(g)], a.L.supK);
        for (int k = 0; k < n; k++) {
            const int j = a.L.sup[row_at(g, k, a.nf)];
            if (j == s) continue;
            const double4 q = a.fsp[j];
            if (within(p.x, p.y, p.z, q.x, q.y, q.z, 2.0 * fmax(p.w, q.w)))
                out[(int64_t)(c++) * a.nf + s] = j;
        }
    }
    out_cnt[s] = c;
}


// ------------------------------------------------ reverse (transposed) lists
// reverse CSR of a pair list by its second index (optionally only keys with
// a positive inverse mass); within a key the pair ids are ascending, as the
// reference's stable transpose sort gives (neighborSearch.cu:345-354)
__global__ void k_rev_count(const int* __restrict__ key, const int* __restrict__ total_p,
                            int64_t cap, const double4* __restrict__ kvel, int* __restrict__ cnt) {
    int64_t tot = *total_p;
    if (tot > cap) tot = cap;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot;
         p += (int64_t)gridDim.x * blockDim.x) {
        int a = key[p];
        if (kvel && !(kvel[a].w > 0.0)) continue;
        atomicAdd(&cnt[a], 1);
    }
}
// cursor[a] starts at the segment's first slot
__global__ void k_rev_fill(const int* __restrict__ key, const int* __restrict__ total_p, int64_t cap,
                           const double4* __restrict__ kvel, int* __restrict__ cursor,
                           int* __restrict__ idx) {
    int64_t tot = *total_p;
    if (tot > cap) tot = cap;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot;
         p += (int64_t)gridDim.x * blockDim.x) {
        int a = key[p];
        if (kvel && !(kvel[a].w > 0.0)) continue;
        // every counted key has a cursor inside [0, cap); the test only
        // keeps a key outside the scanned list from writing anywhere
        const int slot = atomicAdd(&cursor[a], 1);
        if (slot >= 0 && slot < cap) idx[slot] = (int)p;
    }
}
// segment of every listed key (all keys without a list) in ascending pair order
__global__ void k_rev_sort(const int* __restrict__ keys, int n, const int* __restrict__ off,
                           const int* __restrict__ cnt, int* __restrict__ idx) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int a = keys ? keys[t] : t;
    const int b = off[a], e = b + cnt[a];
    for (int i = b + 1; i < e; i++) {  // insertion sort: segments are short
        int v = idx[i];
        int k = i - 1;
        while (k >= b && idx[k] > v) {
            idx[k + 1] = idx[k];
            k--;
        }
        idx[k + 1] = v;
    }
}
// The reverse CSRs of two independent pair lists built together (the
// fluid–solid pairs by movable solid and the solid–solid contacts by
// partner): each kernel of the chain serves both lists, so the chain costs
// four launches instead of eight.  Each list keeps its own counts and
// cursors; its keys (keys[0, nkeys)) are the ones that can have pairs, in
// the keyed (sparse) form of the single-list build, with the same results.
struct RevList {
    const int* key;          // pair -> key
    const int* total;        // device-side pair count
    int64_t cap;             // pairs stored (0: the list is left out)
    const double4* kvel;     // given: only keys with kvel[k].w > 0
    const int* keys;
    int nkeys;
    int n_all;               // keys in all (off[n_all] = the total)
    int* cnt;
    int* off;
    int* idx;
    int* tmp;                // zero on entry and on exit
    int* cursor;
};
__device__ __forceinline__ int64_t rev_total(const RevList& R) {
    const int64_t t = *R.total;
    return t > R.cap ? R.cap : t;
}
__device__ __forceinline__ void rev_count_one(const RevList R) {
    if (R.cap <= 0) return;
    const int64_t tot = rev_total(R);
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int a = R.key[p];
        if (R.kvel && !(R.kvel[a].w > 0.0)) continue;
        atomicAdd(&R.tmp[a], 1);
    }
}
__global__ void k_rev_count2(RevList A, RevList B) {
    rev_count_one(A);
    rev_count_one(B);
}
// one workgroup per list: counts move from tmp to cnt (tmp left zero),
// off and cursor the exclusive scan over the list's keys
constexpr int kRevScanBS = 1024;
__global__ void __launch_bounds__(kRevScanBS) k_rev_scan2(RevList A, RevList B) {
    __shared__ int lds[kRevScanBS / 64];
    __shared__ int total;
    const RevList R = blockIdx.x == 0 ? A : B;
    if (R.cap <= 0) return;
    int carry = 0;
    for (int base = 0; base < R.nkeys; base += kRevScanBS) {
        const int r = base + (int)threadIdx.x;
        int key = -1, v = 0;
        if (r < R.nkeys) {
            key = R.keys[r];
            v = R.tmp[key];
        }
        const int inc = block_scan_int(v, lds);
        const int excl = carry + inc - v;
        if (r < R.nkeys) {
            R.off[key] = excl;
            R.cnt[key] = v;
            R.tmp[key] = 0;
            R.cursor[key] = excl;
        }
        if (threadIdx.x == kRevScanBS - 1) total = inc;
        __syncthreads();
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) R.off[R.n_all] = carry;
}
__device__ __forceinline__ void rev_fill_one(const RevList R) {
    if (R.cap <= 0) return;
    const int64_t tot = rev_total(R);
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int a = R.key[p];
        if (R.kvel && !(R.kvel[a].w > 0.0)) continue;
        const int slot = atomicAdd(&R.cursor[a], 1);
        if (slot >= 0 && slot < R.cap) R.idx[slot] = (int)p;
    }
}
__global__ void k_rev_fill2(RevList A, RevList B) {
    rev_fill_one(A);
    rev_fill_one(B);
}
// each listed key's segment in ascending pair order (keys of A, then of B)
__global__ void k_rev_sort2(RevList A, RevList B) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int na = A.cap > 0 ? A.nkeys : 0, nb = B.cap > 0 ? B.nkeys : 0;
    if (t >= na + nb) return;
    const RevList R = t < na ? A : B;
    if (t >= na) t -= na;
    const int a = R.keys[t];
    const int b = R.off[a], e = b + R.cnt[a];
    for (int i = b + 1; i < e; i++) {  // insertion sort: segments are short
        const int v = R.idx[i];
        int k = i - 1;
        while (k >= b && R.idx[k] > v) {
            R.idx[k + 1] = R.idx[k];
            k--;
        }
        R.idx[k + 1] = v;
    }
}

// dense variant of the keyed scan's bookkeeping: counts move from tmp to
// cnt (tmp left zero) and cursor = off
__global__ void k_rev_settle(int n, int* __restrict__ tmp, int* __restrict__ cnt, const int* __restrict__ off,
                             int* __restrict__ cursor) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        cnt[i] = tmp[i];
        tmp[i] = 0;
        cursor[i] = off[i];
    }
}

// ------------------------------------------------------- DEM contacts
// interactionSolid2Solid::save (deviceStructs.h:693-706): keep the previous
// list's offsets, partners and springs for the history lookup
__global__ void k_ss_save(const int* __restrict__ active, int nactive, const int* __restrict__ ssoff,
                          const int* __restrict__ sscnt, int ns, const int* __restrict__ ss_j,
                          const double* __restrict__ ssS, int64_t cap, int* __restrict__ hoff,
                          int* __restrict__ hcnt, int* __restrict__ h_j, double* __restrict__ hS) {
    int64_t tot = ssoff[ns];
    if (tot > cap) tot = cap;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nactive;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int a = active[t];
        hoff[a] = ssoff[a];
        hcnt[a] = sscnt[a];
    }
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
         i += (int64_t)gridDim.x * blockDim.x) {
        h_j[i] = ss_j[i];
#pragma unroll
        for (int k = 0; k < 9; k++) hS[9 * i + k] = ssS[9 * i + k];
    }
}


// The active grid of up to kSmallSort solids by counting: with the unique
// 64-bit keys K_i = (cell_i << 32) | i, the sorted position of solid i is
// #{j : K_j < K_i}, and its cell's range is [#{j : K_j < cell_i << 32},
// #{j : K_j < (cell_i + 1) << 32}).  Pass 1 counts these over tiles of j in
// parallel (one workgroup per (i tile, j tile), counts added atomically
// into zeroed accumulators) and resets the cells the previous sort filled;
// pass 2 places every solid, writes its cell's range (all members write the
// same values) and zeroes the accumulators again.
constexpr int kRankI = 64, kRankJ = 64;
__device__ __forceinline__ unsigned long long akey64(const int* active, const unsigned* skey, int t) {
    return ((unsigned long long)skey[active[t]] << 32) | (unsigned)t;
}
__global__ void __launch_bounds__(kRankI) k_active_rank(const int* __restrict__ active, int n,
                                                        const unsigned* __restrict__ skey,
                                                        const unsigned* __restrict__ sorted_prev,
                                                        int* __restrict__ cs, int* __restrict__ ce,
                                                        int* __restrict__ acc) {
    __shared__ unsigned long long kj[kRankJ];
    const int ni = (n + kRankI - 1) / kRankI;
    const int ti = blockIdx.x % ni, tj = blockIdx.x / ni;
    const int i = ti * kRankI + threadIdx.x;
    const int j0 = tj * kRankJ;
    // reset the previous sort's cells (once, by the first column of tiles)
    if (tj == 0 && i < n) {
        const unsigned k = sorted_prev[i];
        cs[k] = -1;
        ce[k] = -1;
    }
    for (int t = threadIdx.x; t < kRankJ; t += kRankI)
        kj[t] = j0 + t < n ? akey64(active, skey, j0 + t) : ~0ull;
    __syncthreads();
    if (i >= n) return;
    const unsigned long long ki = akey64(active, skey, i);
    const unsigned long long lo = ki & 0xffffffff00000000ull, hi = lo + (1ull << 32);
    int r = 0, cl = 0, ch = 0;
    const int m = min(kRankJ, n - j0);
    for (int t = 0; t < m; t++) {
        const unsigned long long x = kj[t];
        r += x < ki;
        cl += x < lo;
        ch += x < hi;
    }
    atomicAdd(&acc[3 * i], r);
    atomicAdd(&acc[3 * i + 1], cl);
    atomicAdd(&acc[3 * i + 2], ch);
}
__global__ void __launch_bounds__(kBS) k_active_place(const int* __restrict__ active, int n,
                                                      const unsigned* __restrict__ skey,
                                                      unsigned* __restrict__ sorted, int* __restrict__ perm,
                                                      int* __restrict__ cs, int* __restrict__ ce,
                                                      int* __restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned k = skey[active[i]];
    const int r = acc[3 * i];
    sorted[r] = k;
    perm[r] = i;
    cs[k] = acc[3 * i + 1];
    ce[k] = acc[3 * i + 2];
    acc[3 * i] = 0;
    acc[3 * i + 1] = 0;
    acc[3 * i + 2] = 0;
}

__global__ void k_active_keys(const int* __restrict__ active, int n, const unsigned* __restrict__ skey,
                              unsigned* __restrict__ akey) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) akey[k] = skey[active[k]];
}

// Pass 0 of the solid–solid search with 32 lanes per active solid: lane
// c < 27 counts the contacts in stencil offset c (each visit of a repeated
// cell counted, as the serial walk does), then the 32 lanes add up.
__global__ void __launch_bounds__(kBS) k_ss_count(Grids G, Solids so, const int* __restrict__ acs,
                                                  const int* __restrict__ ace, const int* __restrict__ aperm,
                                                  int* __restrict__ sscnt) {
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int k0 = gt >> 5, c = gt & 31;
    int n = 0;
    if (k0 < so.nactive && c < 27) {
        const int a = so.active[k0];
        const double4 p = so.pr[a];
        const double ra = so.eff[a];
        const int ca = so.clump[a];
        const bool fixa = so.vel[a].w < 1.e-10;
        const GridP& g = grid_at(G, solid_scene(G, a));
        int ax, ay, az;
        grid_pos(g, p.x, p.y, p.z, ax, ay, az);
        const int zz = c / 9 - 1, yy = (c / 3) % 3 - 1, xx = c % 3 - 1;
        const int hl = (az + zz) * g.gy * g.gx + (ay + yy) * g.gx + (ax + xx);
        if (hl >= 0 && hl < g.ncell) {
            const int h = g.cell0 + hl;
            const int st0 = acs[h];
            if (st0 >= 0) {
                const int en = ace[h];
                for (int t = st0; t < en; t++) {
                    const int b = so.active[aperm[t]];
                    if (b <= a) continue;
                    if (ca >= 0 && ca == so.clump[b]) continue;
                    if (fixa && so.vel[b].w < 1.e-10) continue;
                    const double4 q = so.pr[b];
                    n += within(p.x, p.y, p.z, q.x, q.y, q.z, so.eff[b] + ra) ? 1 : 0;
                }
            }
        }
    }
#pragma unroll
    for (int d = 16; d > 0; d >>= 1) n += __shfl_xor(n, d, 64);
    if (c == 0 && k0 < so.nactive) sscnt[so.active[k0]] = n;
}

// setSolid2SolidInteractionsKernel (neighborSearch.cu:222-312) over the
// non-dummy solids only (dummies are never contacts): pass 0 counts, pass 1
// fills and carries the springs of the same (i, j) over from the previous
// step's list.  The grid of active solids keeps the reference's order
// (stencil order, ascending index within a cell).
template <bool FILL>
__global__ void __launch_bounds__(kBS) k_ss_search(
    Grids G, Solids so, int ns, const int* __restrict__ acs, const int* __restrict__ ace,
    const int* __restrict__ aperm, int* __restrict__ sscnt, const int* __restrict__ ssoff,
    int64_t cap, int* __restrict__ ss_i, int* __restrict__ ss_j, double4* __restrict__ ssF,
    double4* __restrict__ ssT, double* __restrict__ ssS, const int* __restrict__ hoff,
    const int* __restrict__ hcnt, const int* __restrict__ h_j, const double* __restrict__ hS,
    unsigned long long* st) {
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (FILL && k0 == 0) {
        long long tot = ssoff[ns];
        st[ST_SS_TOTAL] = (unsigned long long)tot;
        if (tot > cap) atomicOr(&st[ST_SS_OVERFLOW], 1ull);
    }
    if (k0 >= so.nactive) return;
    const int a = so.active[k0];
    if (FILL && sscnt[a] == 0) return;
    const double4 p = so.pr[a];
    const double ra = so.eff[a];
    const int ca = so.clump[a];
    const bool fixa = so.vel[a].w < 1.e-10;
    const GridP& g = grid_at(G, solid_scene(G, a));
    int ax, ay, az;
    grid_pos(g, p.x, p.y, p.z, ax, ay, az);
    int c = 0;
    int64_t w = FILL ? (int64_t)ssoff[a] : 0;
    int hb = 0, he = 0;
    if (FILL) {
        hb = hoff[a];
        he = hb + hcnt[a];
    }
    for (int zz = -1; zz <= 1; zz++)
        for (int yy = -1; yy <= 1; yy++)
            for (int xx = -1; xx <= 1; xx++) {
                const int hl = (az + zz) * g.gy * g.gx + (ay + yy) * g.gx + (ax + xx);
                if (hl < 0 || hl >= g.ncell) continue;
                const int h = g.cell0 + hl;
                const int st0 = acs[h];
                if (st0 < 0) continue;
                const int en = ace[h];
                for (int t = st0; t < en; t++) {
                    const int b = so.active[aperm[t]];
                    if (b <= a) continue;
                    if (ca >= 0 && ca == so.clump[b]) continue;
                    if (fixa && so.vel[b].w < 1.e-10) continue;
                    const double4 q = so.pr[b];
                    if (!within(p.x, p.y, p.z, q.x, q.y, q.z, so.eff[b] + ra)) continue;
                    if (!FILL) {
                        c++;
                        continue;
                    }
                    if (w < cap) {
                        ss_i[w] = a;
                        ss_j[w] = b;
                        ssF[w] = make_double4(0, 0, 0, 0);
                        ssT[w] = make_double4(0, 0, 0, 0);
                        double sp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                        for (int k = hb; k < he && k < cap; k++)
                            if (h_j[k] == b) {
#pragma unroll
                                for (int m = 0; m < 9; m++) sp[m] = hS[9 * (int64_t)k + m];
                                break;
                            }
#pragma unroll
                        for (int m = 0; m < 9; m++) ssS[9 * w + m] = sp[m];
                    }
                    w++;
                }
            }
    if (!FILL) sscnt[a] = c;
}

// setSolidNormal (dataBase.cpp:348-413): n = normalize(-sum V_B gradW) over
// fixed–fixed or same-clump neighbours within 2h
__global__ void k_normals(Grids G, Solids so, int ns, const int* __restrict__ scs,
                          const int* __restrict__ sce, const int* __restrict__ sperm) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ns) return;
    const GridP& g = grid_at(G, solid_scene(G, a));
    const double h = g.hmax;  // the scene's largest fluid smoothing length
    const double4 p = so.pr[a];
    const int ca = so.clump[a];
    const double ima = so.vel[a].w;
    int ax, ay, az;
    grid_pos(g, p.x, p.y, p.z, ax, ay, az);
    d3 th{0, 0, 0};
    const double h13 = h / 1.3;
    const double Vwall = h13 * h13 * h13;
    for (int zz = -1; zz <= 1; zz++)
        for (int yy = -1; yy <= 1; yy++)
            for (int xx = -1; xx <= 1; xx++) {
                const int hl = (az + zz) * g.gy * g.gx + (ay + yy) * g.gx + (ax + xx);
                if (hl < 0 || hl >= g.ncell) continue;
                const int hh = g.cell0 + hl;
                const int st0 = scs[hh];
                if (st0 < 0) continue;
                const int en = sce[hh];
                for (int t = st0; t < en; t++) {
                    const int b = sperm[t];
                    const double4 q = so.pr[b];
                    d3 r{p.x - q.x, p.y - q.y, p.z - q.z};
                    const double rl = vlen(r);
                    if (rl > 2 * h) continue;
                    double V;
                    if (ima <= 1.e-10 && so.vel[b].w <= 1.e-10) {
                        V = Vwall;
                    } else if (ca >= 0 && ca == so.clump[b]) {
                        double d = 2.0 * q.w;
                        V = d * d * d;
                    } else {
                        continue;
                    }
                    th += (-V * grad_wendland_factor(rl, h)) * r;
                }
            }
    d3 n = vlen(th) > 1.e-10 ? unit(th) : d3{0, 0, 0};
    so.nrm[a] = make_double4(n.x, n.y, n.z, 0.0);
}

int cap_grid(int64_t n, int cap) {
    int b = blocks_for(n);
    return b > cap ? cap : b;
}

}  // namespace

// ---------------------------------------------------------------- wrappers
void hash_fluid(hipStream_t s, const Grids& G, const double4* ph, const int* orig, int n, int sub_bits,
                const unsigned* crank, const unsigned char* cspare, unsigned* key, unsigned* rank,
                unsigned* hslot, const double4* fxb, unsigned long long* bmax, int* near_out, int ncell,
                unsigned long long* ctl, long long* status) {
    static_assert(ST_WORDS <= kBS, "the first workgroup clears the status words");
    const int nb = near_out ? std::min(std::max(blocks_for(ncell), 1), 4096) : 1;
    hipLaunchKernelGGL(k_status_reset, dim3(nb), dim3(kBS), 0, s, (unsigned long long*)status, near_out, ncell);
    if (n > 0)
        hipLaunchKernelGGL(k_hash_fluid, dim3(blocks_for(n)), dim3(kBS), 0, s, G, ph, orig, n, sub_bits,
                           crank, cspare, key, rank, hslot, fxb, bmax, near_out, ctl,
                           (unsigned long long*)status);
}

void hash_points(hipStream_t s, const GridP& g, const double* xyz, int n, unsigned* key) {
    if (n > 0) hipLaunchKernelGGL(k_hash_points, dim3(blocks_for(n)), dim3(kBS), 0, s, g, xyz, n, key);
}

void hash_solid(hipStream_t s, const Grids& G, const double4* pr, int n, unsigned* key,
                long long* status) {
    if (n > 0)
        hipLaunchKernelGGL(k_hash_solid, dim3(blocks_for(n)), dim3(kBS), 0, s, G, pr, n, key,
                           (unsigned long long*)status);
}

void hash_live_decide(hipStream_t s, const Grids& G, const double4* pr, const int* live, int nlive,
                      unsigned* key, unsigned long long static_out, bool decide, const double4* sxb,
                      const unsigned long long* fbmax, int nf, int always, int always_sort,
                      unsigned long long* ctl, long long* status) {
    if (nlive <= 0 && !decide && static_out == 0) return;
    // one workgroup at least, so that a decision is taken without solids
    const int nb = std::max(1, blocks_for(nlive));
    hipLaunchKernelGGL(k_hash_live, dim3(nb), dim3(kBS), 0, s, G, pr, live, nlive, key, static_out,
                       (int)decide, sxb, fbmax, nf > 0 ? blocks_for(nf) : 0, always, always_sort, ctl,
                       (unsigned long long*)status);
}

void reorder_fluid(hipStream_t s, const FluidSet& f, const FluidSet& scratch, const int* perm,
                   int* slot_of, int n, bool mixed, int cur, const unsigned long long* ctl) {
    if (n <= 0) return;
    if (mixed)
        hipLaunchKernelGGL(k_reorder<float4>, dim3(blocks_for(n)), dim3(kBS), 0, s, f, scratch, perm,
                           slot_of, n, cur, ctl);
    else
        hipLaunchKernelGGL(k_reorder<double4>, dim3(blocks_for(n)), dim3(kBS), 0, s, f, scratch, perm,
                           slot_of, n, cur, ctl);
}

void after_sort(hipStream_t s, const FluidSet& f, const FluidSet& scratch, bool reorder, bool mixed, int cur,
                const unsigned* sorted, int n, int sub_bits, const int* perm, const int* slot_of,
                const unsigned* cinv, int* member, unsigned* hslot, const unsigned long long* ctl) {
    if (n <= 0) return;
    if (mixed)
        hipLaunchKernelGGL(k_after_sort<float4>, dim3(blocks_for(n)), dim3(kBS), 0, s, f, scratch, n, cur,
                           (int)reorder, sorted, sub_bits, perm, slot_of, cinv, member, hslot, ctl);
    else
        hipLaunchKernelGGL(k_after_sort<double4>, dim3(blocks_for(n)), dim3(kBS), 0, s, f, scratch, n, cur,
                           (int)reorder, sorted, sub_bits, perm, slot_of, cinv, member, hslot, ctl);
}

static int chunk_grid(int nchunk) {
    // one wavefront per chunk, 4 per workgroup; beyond the cap the waves
    // stride over the chunk list
    const int b = (nchunk + 3) / 4;
    return b < 1 ? 1 : (b > 65535 ? 65535 : b);
}

// ------------------------------------------------------- pair-loop tiles
// The union of a tile's candidates and its local rows, after a superset
// build.  One workgroup of kTileSlots threads per tile, one thread per slot.
// Slots are grouped in pages of 256 consecutive slots, and the union is
// kept as one 256-bit membership mask per page it touches, in an LDS hash
// table keyed by page (open addressing, linear probing):
//  1. every candidate of the tile's rows sets its bit (the particle itself
//     is left out: its term is an exact zero in continuity and momentum,
//     and the Shepard sum adds it apart);
//  2. the pages are ranked by page number, and each page's members get
//     their union positions from the counts of the pages before it and of
//     the mask words before theirs: the union in ascending slots (the pair
//     loops stage its records with coalesced loads) with no sort;
//  3. each row is written again as union positions, in its own order: a
//     page lookup and a population count per candidate;
//  4. the entries the pair loops read past a row's end get position 0.
// More than kTilePages pages, a union larger than kTileCap or a candidate
// not found leave the tile on its global rows (ucnt = -1).
constexpr int kTilePageBits = 8;                 // 256 slots per page
constexpr int kTilePageHash = 256;               // hash entries (power of two)
constexpr int kTilePages = kTilePageHash / 2;    // pages a tile may touch
constexpr int kTileBatch = 32;
// multiplicative (Fibonacci) hashing: the top 8 bits of key * 2^32 / phi
__device__ __forceinline__ unsigned page_hash(unsigned key) { return (key * 2654435761u) >> 24; }
static_assert(kTilePageHash == 256, "page_hash keeps 8 bits");

__global__ void __launch_bounds__(kTileSlots) k_tile_build(SearchArgs a) {
    if (a.ctl[CTL_REBUILD] == 0) return;
    __shared__ unsigned pkey[kTilePageHash];        // page + 1; 0 = empty
    __shared__ unsigned pbits[kTilePageHash][8];    // membership mask of the page's 256 slots
    __shared__ unsigned short wpos[kTilePageHash][8];  // union position of the mask word's first member
    __shared__ int prank[kTilePageHash + 1];        // pages by rank: members, then exclusive prefix
    __shared__ int npage, bad;
    const Tiles& T = a.T;
    const int b = xcd_block();
    const int2 tr = T.range[b];
    const int tid = threadIdx.x;
    for (int h = tid; h < kTilePageHash; h += blockDim.x) {
        pkey[h] = 0u;
#pragma unroll
        for (int w = 0; w < 8; w++) pbits[h][w] = 0u;
    }
    if (tid == 0) {
        npage = 0;
        bad = 0;
    }
    __syncthreads();
    const int s = tr.x + tid;
    const bool act = s < tr.y;
    int g = 0, n = 0;
    if (act) {
        g = a.L.grow[s];
        n = min(a.L.sup_cnt[row_slot(g)], a.L.supK);
    }
    // the row is read kTileBatch candidates at a time, all loads of a batch
    // in flight together (entries past the row end read as the particle
    // itself, which is skipped)
    constexpr int V = kTileBatch;
    auto batch = [&](int c0, int* jj) {
#pragma unroll
        for (int v = 0; v < V; v++) jj[v] = c0 + v < n ? a.L.sup[row_at(g, c0 + v, a.nf)] : s;
    };
    // the hash entry of page p, inserted if absent (-1: the table is full).
    // Most candidates' pages are in already: a plain read finds them, and
    // only an empty entry is claimed with an atomic (the lanes of a wave
    // would otherwise queue on the same entries)
    auto page_slot = [&](unsigned p, bool insert) -> int {
        const unsigned key = p + 1u;
        unsigned h = page_hash(key);
        for (int q = 0; q < kTilePageHash; q++) {
            unsigned old = pkey[h];
            if (insert && old == 0u) old = atomicCAS(&pkey[h], 0u, key);
            if (old == key) return (int)h;
            if (old == 0u) {
                if (!insert) return -1;
                atomicAdd(&npage, 1);
                return (int)h;
            }
            h = (h + 1u) & (kTilePageHash - 1);
        }
        return -1;
    };
    // 1. membership
    for (int c0 = 0; c0 < n; c0 += V) {
        int jj[V];
        batch(c0, jj);
        for (int v = 0; v < V; v++) {
            if (jj[v] == s) continue;
            const unsigned j = (unsigned)jj[v];
            const int h = page_slot(j >> kTilePageBits, true);
            if (h < 0) {
                bad = 1;
                continue;
            }
            // most members are set already (a candidate is in many rows)
            unsigned* word = &pbits[h][(j >> 5) & 7];
            const unsigned bit = 1u << (j & 31);
            if ((*(volatile unsigned*)word & bit) == 0u) atomicOr(word, bit);
        }
    }
    __syncthreads();
    if (bad || npage > kTilePages) {
        if (tid == 0) T.ucnt[b] = -1;
        return;
    }
    // 2. the rank of each page by page number, its member count by rank
    for (int h = tid; h < kTilePageHash; h += blockDim.x) {
        const unsigned k = pkey[h];
        if (k != 0u) {
            int r = 0;
            for (int h2 = 0; h2 < kTilePageHash; h2++) {
                const unsigned k2 = pkey[h2];
                r += (k2 != 0u) & (k2 < k);
            }
            int cnt = 0;
#pragma unroll
            for (int w = 0; w < 8; w++) cnt += __popc(pbits[h][w]);
            prank[r] = cnt;
            wpos[h][0] = (unsigned short)r;  // the rank until the offsets are in
        }
    }
    __syncthreads();
    if (tid == 0) {  // exclusive prefix over at most kTilePages ranks
        int acc = 0;
        for (int r = 0; r < npage; r++) {
            const int c = prank[r];
            prank[r] = acc;
            acc += c;
        }
        prank[npage] = acc;
    }
    __syncthreads();
    const int m = prank[npage];
    if (m > kTileCap) {
        if (tid == 0) T.ucnt[b] = -1;
        return;
    }
    int* uni = T.uni + (int64_t)b * kTileCap;
    for (int h = tid; h < kTilePageHash; h += blockDim.x) {
        const unsigned k = pkey[h];
        if (k == 0u) continue;
        int at = prank[wpos[h][0]];
        const unsigned base = (k - 1u) << kTilePageBits;
#pragma unroll
        for (int w = 0; w < 8; w++) {
            wpos[h][w] = (unsigned short)at;
            for (unsigned bits = pbits[h][w]; bits; bits &= bits - 1u)
                uni[at++] = (int)(base + 32u * w + (unsigned)__builtin_ctz(bits));
        }
    }
    __syncthreads();
    // 3. the local rows
    if (act) {
        int w = 0;
        for (int c0 = 0; c0 < n; c0 += V) {
            int jj[V];
            batch(c0, jj);
            for (int v = 0; v < V; v++) {
                if (jj[v] == s) continue;
                const unsigned j = (unsigned)jj[v];
                const int h = page_slot(j >> kTilePageBits, false);
                int pos = 0;
                if (h < 0) {
                    bad = 1;
                } else {
                    const unsigned word = (j >> 5) & 7;
                    pos = wpos[h][word] + __popc(pbits[h][word] & ((1u << (j & 31)) - 1u));
                }
                T.row[tile_row_at(w, s, a.nf)] = (unsigned short)tile_swz(pos);
                w++;
            }
        }
        // 4. the read-ahead entries, to the end of the third quad after the
        // row's last (w + 15 <= supK + kRowPad): the pair loops' sentinel
        // entry kTileCap, which lies outside every support
        for (int c = w; c < ((w + 3) & ~3) + 12; c++) T.row[tile_row_at(c, s, a.nf)] = (unsigned short)kTileCap;
        T.rcnt[s] = w;
    }
    __syncthreads();
    if (tid == 0) T.ucnt[b] = bad ? -1 : m;
}

void tile_rows(hipStream_t s, const SearchArgs& a) {
    if (a.nf > 0 && a.T.n > 0)
        hipLaunchKernelGGL(k_tile_build, dim3(a.T.n), dim3(kTileSlots), 0, s, a);
}

void fluid_superset(hipStream_t s, const SearchArgs& a, const int2* chunks, int nchunk) {
    if (a.nf <= 0) return;
    if (a.group == 1)
        hipLaunchKernelGGL(k_build<true>, dim3(chunk_grid(nchunk)), dim3(kBS), 0, s, a, chunks, nchunk);
    else
        hipLaunchKernelGGL(k_build<false>, dim3(chunk_grid(nchunk)), dim3(kBS), 0, s, a, chunks, nchunk);
}

void fluid_exact_rows(hipStream_t s, const SearchArgs& a, int* out, int* out_cnt) {
    if (a.nf > 0)
        hipLaunchKernelGGL(k_exact_rows, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a, out, out_cnt);
}

void fluid_filter(hipStream_t s, const SearchArgs& a, bool zero_fs) {
    if (a.nf <= 0) return;
    hipLaunchKernelGGL(k_filter_count, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a);
    hipLaunchKernelGGL(k_filter_fill, dim3(blocks_for(a.nf)), dim3(kBS), 0, s, a, (int)zero_fs);
}

// Reverse CSR of pairs by their key (a solid): count into tmp (zero on
// entry and on exit), scan, place with per-key cursors, then sort each
// segment ascending.  Only the listed keys can have pairs: a short list is
// scanned by one workgroup over the list, a long one over all n_all keys.
static void reverse_build(hipStream_t s, const int* key, const int* total_p, int64_t cap,
                          const double4* kvel, const int* keys, int nkeys, int n_all, int* cnt, int* off,
                          int* idx, int* tmp, int* cursor, int* scan_tmp) {
    if (n_all <= 0) return;
    const int gb = cap_grid(cap, 2048);
    if (cap > 0) hipLaunchKernelGGL(k_rev_count, dim3(gb), dim3(kBS), 0, s, key, total_p, cap, kvel, tmp);
    const bool sparse = nkeys <= kKeyScanMax;
    if (sparse) {
        scan_keys(s, keys, nkeys, tmp, cnt, off, cursor, n_all);
    } else {
        exclusive_scan(s, tmp, off, n_all, scan_tmp);
        hipLaunchKernelGGL(k_rev_settle, dim3(cap_grid(n_all, 2048)), dim3(kBS), 0, s, n_all, tmp, cnt,
                           (const int*)off, cursor);
    }
    if (cap > 0)
        hipLaunchKernelGGL(k_rev_fill, dim3(gb), dim3(kBS), 0, s, key, total_p, cap, kvel, cursor, idx);
    const int ns = sparse ? nkeys : n_all;
    if (ns > 0)
        hipLaunchKernelGGL(k_rev_sort, dim3(blocks_for(ns)), dim3(kBS), 0, s, sparse ? keys : nullptr, ns,
                           (const int*)off, (const int*)cnt, idx);
}

void reverse_movable(hipStream_t s, const int* pair_key, const int* pair_total, int64_t cap,
                     const double4* key_vel, const Solids& so, int ns, Lists& L, int* scan_tmp) {
    // movable solids (1/m > 0) are live
    reverse_build(s, pair_key, pair_total, cap, key_vel, so.live, so.nlive, ns, L.fsr_cnt, L.fsr_off,
                  L.fsr_idx, L.rev_tmp, L.rev_cur, scan_tmp);
}

void solid_save_history(hipStream_t s, const Solids& so, Lists& L, int ns) {
    if (ns <= 0) return;
    const int64_t n = L.ss_cap > so.nactive ? L.ss_cap : so.nactive;
    hipLaunchKernelGGL(k_ss_save, dim3(cap_grid(n > 0 ? n : 1, 2048)), dim3(kBS), 0, s, (const int*)so.active,
                       so.nactive, (const int*)L.ssoff, (const int*)L.sscnt, ns, (const int*)L.ss_j,
                       (const double*)L.ssS, L.ss_cap, L.hoff, L.hcnt, L.h_j, L.hS);
}

void active_grid(hipStream_t s, const Solids& so, const unsigned* skey, int ncell, int key_bits,
                 unsigned* akey, unsigned* akey_sorted, int* aperm, int* acs, int* ace,
                 unsigned* radix_tmp, int* rank_acc) {
    const int n = so.nactive;
    if (n > 0 && n <= kSmallSort) {
        const int ni = (n + kRankI - 1) / kRankI, nj = (n + kRankJ - 1) / kRankJ;
        hipLaunchKernelGGL(k_active_rank, dim3(ni * nj), dim3(kRankI), 0, s, (const int*)so.active, n, skey,
                           (const unsigned*)akey_sorted, acs, ace, rank_acc);
        hipLaunchKernelGGL(k_active_place, dim3(blocks_for(n)), dim3(kBS), 0, s, (const int*)so.active, n,
                           skey, akey_sorted, aperm, acs, ace, rank_acc);
        return;
    }
    if (n > 0) {
        hipLaunchKernelGGL(k_active_keys, dim3(blocks_for(n)), dim3(kBS), 0, s, (const int*)so.active, n,
                           skey, akey);
        radix_sort_iota(s, akey, n, key_bits, akey_sorted, aperm, radix_tmp);
    }
    cell_spans(s, akey_sorted, n, acs, ace, ncell, nullptr);
}

void solid_search(hipStream_t s, const Grids& g, const Solids& so, int ns, const int* acs,
                  const int* ace, const int* aperm, Lists& L, int* scan_tmp, long long* status) {
    if (ns <= 0) return;
    const int nb = blocks_for(so.nactive > 0 ? so.nactive : 1);
    if (so.nactive > 0)
        hipLaunchKernelGGL(k_ss_count, dim3(blocks_for(32 * (int64_t)so.nactive)), dim3(kBS), 0, s, g, so, acs,
                           ace, aperm, L.sscnt);
    // only active solids have contacts: offsets are kept for them (and the
    // total at ns)
    if (so.nactive <= kKeyScanMax)
        scan_keys(s, so.active, so.nactive, L.sscnt, nullptr, L.ssoff, nullptr, ns);
    else
        exclusive_scan(s, L.sscnt, L.ssoff, ns, scan_tmp);
    hipLaunchKernelGGL(k_ss_search<true>, dim3(nb), dim3(kBS), 0, s, g, so, ns, acs, ace, aperm, L.sscnt,
                       (const int*)L.ssoff, L.ss_cap, L.ss_i, L.ss_j, L.ssF, L.ssT, L.ssS,
                       (const int*)L.hoff, (const int*)L.hcnt, (const int*)L.h_j, (const double*)L.hS,
                       (unsigned long long*)status);
}

void reverse_solid(hipStream_t s, const Solids& so, Lists& L, int ns, int* scan_tmp) {
    // partners are active solids
    reverse_build(s, L.ss_j, L.ssoff + ns, L.ss_cap, nullptr, so.active, so.nactive, ns, L.ssr_cnt, L.ssr_off,
                  L.ssr_idx, L.rev_tmp, L.rev_cur, scan_tmp);
}

void reverse_both(hipStream_t s, const int* pair_key, const int* pair_total, int64_t cap,
                  const double4* key_vel, const Solids& so, Lists& L, int ns, int* scan_tmp) {
    if (ns <= 0) return;
    if (so.nlive > kKeyScanMax || so.nactive > kKeyScanMax) {
        reverse_movable(s, pair_key, pair_total, cap, key_vel, so, ns, L, scan_tmp);
        reverse_solid(s, so, L, ns, scan_tmp);
        return;
    }
    const RevList A{pair_key, pair_total, cap, key_vel, so.live, so.nlive, ns,
                    L.fsr_cnt, L.fsr_off, L.fsr_idx, L.rev_tmp2, L.rev_cur2};
    const RevList B{L.ss_j, L.ssoff + ns, L.ss_cap, nullptr, so.active, so.nactive, ns,
                    L.ssr_cnt, L.ssr_off, L.ssr_idx, L.rev_tmp, L.rev_cur};
    const int gb = cap_grid(cap > L.ss_cap ? cap : L.ss_cap, 2048);
    hipLaunchKernelGGL(k_rev_count2, dim3(gb), dim3(kBS), 0, s, A, B);
    hipLaunchKernelGGL(k_rev_scan2, dim3(2), dim3(kRevScanBS), 0, s, A, B);
    hipLaunchKernelGGL(k_rev_fill2, dim3(gb), dim3(kBS), 0, s, A, B);
    const int nk = (cap > 0 ? so.nlive : 0) + (L.ss_cap > 0 ? so.nactive : 0);
    if (nk > 0) hipLaunchKernelGGL(k_rev_sort2, dim3(blocks_for(nk)), dim3(kBS), 0, s, A, B);
}

void solid_normals(hipStream_t s, const Grids& G, const Solids& so, int ns, const int* scs,
                   const int* sce, const int* sperm) {
    if (ns > 0)
        hipLaunchKernelGGL(k_normals, dim3(blocks_for(ns)), dim3(kBS), 0, s, G, so, ns, scs, sce, sperm);
}

}  // namespace dsph
