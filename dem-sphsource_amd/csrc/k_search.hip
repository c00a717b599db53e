// k_search.hip — hashed-grid neighbour search on the reference's grid.
//
// Follows core/neighborSearch.cu (hash :19-36, fluid lists :115-220, solid
// contact lists with spring history :222-312, driver :314-381) and the
// solid-normal estimate of core/dataBase.cpp:348-413.
//
// Bit-exactness: the hash and every distance test are evaluated in FP64
// with FMA contraction off, IEEE division and correctly rounded sqrt, so the
// cell of a particle and the membership of a pair are the reference's.  The
// distance test skips the sqrt when d^2 is outside [R^2(1-1e-14),
// R^2(1+1e-14)] (sqrt_rn is monotone, so the answer is the same).
//
// MI355X design: one thread per particle in cell-sorted order (a wave covers
// ~64 consecutive particles of one or two cells, so its 27-cell stencils and
// neighbour gathers coincide); the reference's inflated cells (cell =
// 2 max effective radius over all species) are pruned per lane with an exact
// box-distance test, so a fluid particle only walks the cells that can hold a
// neighbour.  The fluid–fluid list is the full (both-direction) list written
// once in stencil order into an ELL table — no transpose sort.
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

// squared distance from p to cell h's box, enlarged by pad (conservative)
__device__ __forceinline__ double box_dist2(const GridP& g, int h, double px, double py,
                                            double pz) {
    const int gxy = g.gx * g.gy;
    const int cz = h / gxy, rem = h - cz * gxy;
    const int cy = rem / g.gx, cx = rem - cy * g.gx;
    auto axis = [&](double p, double mn, double cs, int c) {
        double lo = mn + c * cs - g.pad, hi = mn + (c + 1) * cs + g.pad;
        double d = lo - p;
        double e = p - hi;
        d = d > e ? d : e;
        return d > 0 ? d : 0.0;
    };
    double ax = axis(px, g.minx, g.csx, cx), ay = axis(py, g.miny, g.csy, cy),
           az = axis(pz, g.minz, g.csz, cz);
    return ax * ax + ay * ay + az * az;
}

__device__ __forceinline__ void wave_max_atomic(unsigned long long* addr, unsigned long long v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        unsigned long long u = __shfl_xor(v, d, 64);
        v = v > u ? v : u;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(addr, v);
}

__global__ void k_status_reset(unsigned long long* st) {
    int i = threadIdx.x;
    // overflow flags are sticky until read; counters are per search
    if (i == ST_OUT_OF_DOMAIN || i == ST_FF_MAX || i == ST_FS_TOTAL || i == ST_SS_TOTAL) st[i] = 0;
}

__global__ void k_hash_fluid(GridP g, const double4* __restrict__ ph, const int* __restrict__ orig,
                             int n, unsigned* __restrict__ key, unsigned long long* st) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long out = 0;
    if (s < n) {
        double4 p = ph[s];
        int h = point_hash(g, p.x, p.y, p.z);
        key[orig[s]] = (unsigned)h;
        out = h == g.ncell - 1;
    }
    unsigned long long m = __ballot(out != 0);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&st[ST_OUT_OF_DOMAIN], (unsigned long long)__popcll(m));
}

__global__ void k_hash_solid(GridP g, const double4* __restrict__ pr, int n,
                             unsigned* __restrict__ key, unsigned long long* st) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long out = 0;
    if (s < n) {
        double4 p = pr[s];
        int h = point_hash(g, p.x, p.y, p.z);
        key[s] = (unsigned)h;
        out = h == g.ncell - 1;
    }
    unsigned long long m = __ballot(out != 0);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&st[ST_OUT_OF_DOMAIN], (unsigned long long)__popcll(m));
}

template <class R4>
__global__ void k_reorder(FluidSet src, FluidSet dst, const int* __restrict__ perm,
                          int* __restrict__ slot_of, int* __restrict__ member, int n, int cur) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int o = perm[t];
    int old = slot_of[o];
    dst.ph[t] = src.ph[old];
    dst.vr[t] = src.vr[old];
    dst.drho[cur][t] = src.drho[cur][old];
    dst.cn[t] = src.cn[old];
    dst.sumw0[t] = src.sumw0[old];
    dst.acc[t] = src.acc[old];
    dst.orig[t] = o;
    ((R4*)dst.pv[cur])[t] = ((const R4*)src.pv[cur])[old];
    ((R4*)dst.pq[cur])[t] = ((const R4*)src.pq[cur])[old];
    slot_of[o] = t;
    member[t] = t;
}

__global__ void k_member(const int* __restrict__ perm, const int* __restrict__ slot_of,
                         int* __restrict__ member, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) member[t] = slot_of[perm[t]];
}

// Per-lane stencil walk of one fluid particle (the reference's own loop,
// neighborSearch.cu:135-214) — used for particles in the spare cell, whose
// stencils differ lane by lane.  COUNT: append fluid neighbours to the ELL
// row and count solid neighbours; FILL: write the solid neighbours.
template <bool FILL>
__device__ void lane_search(const GridP& g, const double4 p, int s, int nf,
                            const int* __restrict__ fcs, const int* __restrict__ fce,
                            const int* __restrict__ fmember, const double4* __restrict__ ph,
                            const int* __restrict__ scs, const int* __restrict__ sce,
                            const int* __restrict__ sperm, const double4* __restrict__ spr,
                            double hmax, int K, int* __restrict__ nbr, int& c, int& cs,
                            int64_t w, int64_t cap, int* __restrict__ fs_i, int* __restrict__ fs_a,
                            double4* __restrict__ fsF) {
    const double hi = p.w;
    int ax, ay, az;
    grid_pos(g, p.x, p.y, p.z, ax, ay, az);
    for (int zz = -1; zz <= 1; zz++)
        for (int yy = -1; yy <= 1; yy++)
            for (int xx = -1; xx <= 1; xx++) {
                const int h = (az + zz) * g.gy * g.gx + (ay + yy) * g.gx + (ax + xx);
                if (h < 0 || h >= g.ncell) continue;
                if (!FILL) {
                    const int st0 = fcs[h];
                    if (st0 >= 0)
                        for (int t = st0; t < fce[h]; t++) {
                            const int j = fmember ? fmember[t] : t;
                            if (j == s) continue;
                            const double4 q = ph[j];
                            if (within(p.x, p.y, p.z, q.x, q.y, q.z, 2.0 * fmax(hi, q.w))) {
                                if (c < K) nbr[(int64_t)c * nf + s] = j;
                                c++;
                            }
                        }
                }
                const int st1 = scs[h];
                if (st1 < 0) continue;
                for (int t = st1; t < sce[h]; t++) {
                    const int a = sperm[t];
                    const double4 q = spr[a];
                    if (!within(p.x, p.y, p.z, q.x, q.y, q.z, 2.0 * hi)) continue;
                    if (FILL) {
                        if (w < cap) {
                            fs_i[w] = s;
                            fs_a[w] = a;
                            fsF[w] = make_double4(0, 0, 0, 0);
                        }
                        w++;
                    } else {
                        cs++;
                    }
                }
            }
}

// Cell blocks: every cell holding fluid is cut into blocks of <= 64
// consecutive sorted particles, one wavefront each.
__global__ void k_blk_count(const int* __restrict__ cs, const int* __restrict__ ce, int ncell,
                            int* __restrict__ nb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int st = cs[c];
    nb[c] = st >= 0 ? (ce[c] - st + 63) >> 6 : 0;
}
__global__ void k_blk_fill(const int* __restrict__ cs, const int* __restrict__ ce, int ncell,
                           const int* __restrict__ off, int2* __restrict__ blocks) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int st = cs[c];
    if (st < 0) return;
    const int n = ce[c] - st;
    const int o = off[c];
    for (int b = 0; b * 64 < n; b++) blocks[o + b] = make_int2(c, st + 64 * b);
}

// setFluidInteractionsKernel (neighborSearch.cu:115-220), one wavefront per
// cell block.  All lanes of a block share the cell, so the 27-cell stencil
// walk is wave-uniform: cell ranges and candidate records are loaded once
// for the wave, and a stencil cell is skipped when no lane's box distance can
// reach it (ballot).  Output: the full fluid–fluid list in stencil order
// (ELL row of the particle), the fluid–solid count.
template <bool FILL>
__global__ void __launch_bounds__(kBS) k_fluid_blocks(
    GridP g, const double4* __restrict__ ph, int nf, const int* __restrict__ fcs,
    const int* __restrict__ fce, const int* __restrict__ fmember, const int* __restrict__ scs,
    const int* __restrict__ sce, const int* __restrict__ sperm, const double4* __restrict__ spr,
    double hmax, int K, int* __restrict__ nbr, int* __restrict__ cnt, int* __restrict__ fscnt,
    const int2* __restrict__ blocks, const int* __restrict__ nblk_p, const int* __restrict__ fsoff,
    int64_t cap, int* __restrict__ fs_i, int* __restrict__ fs_a, double4* __restrict__ fsF,
    unsigned long long* st) {
    const int lane = threadIdx.x & 63;
    const int wv = xcd_block() * (kBS / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwv = gridDim.x * (kBS / 64);
    const int nblk = *nblk_p;
    unsigned long long cmax = 0;
    bool ovf = false;
    for (int b = wv; b < nblk; b += nwv) {
        const int2 blk = blocks[b];
        const int cell = __builtin_amdgcn_readfirstlane(blk.x);
        const int t0 = __builtin_amdgcn_readfirstlane(blk.y);
        const int nmem = min(64, fce[cell] - t0);
        const bool act = lane < nmem;
        const int t = t0 + (act ? lane : 0);
        const int s = fmember ? fmember[t] : t;
        const double4 p = ph[s];
        int c = 0, cs = 0;
        int64_t w = 0;
        if (FILL) {
            const int mine = act ? fscnt[s] : 0;
            if (__ballot(mine > 0) == 0) continue;
            w = fsoff[s];
        }
        if (cell == g.ncell - 1) {  // spare cell: stencils differ per lane
            if (act)
                lane_search<FILL>(g, p, s, nf, fcs, fce, fmember, ph, scs, sce, sperm, spr, hmax, K,
                                  nbr, c, cs, w, cap, fs_i, fs_a, fsF);
        } else {
            const int gxy = g.gx * g.gy;
            const int cz = cell / gxy, rem = cell - cz * gxy;
            const int cy = rem / g.gx, cx = rem - cy * g.gx;
            const double hi = p.w;
            const double Rf = 2.0 * fmax(hi, hmax) * (1.0 + 1e-9);
            const double Rs = 2.0 * hi * (1.0 + 1e-9);
            for (int zz = -1; zz <= 1; zz++)
                for (int yy = -1; yy <= 1; yy++)
                    for (int xx = -1; xx <= 1; xx++) {
                        const int h = (cz + zz) * gxy + (cy + yy) * g.gx + (cx + xx);
                        if (h < 0 || h >= g.ncell) continue;
                        const double bd2 = (h == g.ncell - 1) ? 0.0 : box_dist2(g, h, p.x, p.y, p.z);
                        if (!FILL && __ballot(act && bd2 <= Rf * Rf) != 0) {
                            const int st0 = __builtin_amdgcn_readfirstlane(fcs[h]);
                            if (st0 >= 0) {
                                const int en = __builtin_amdgcn_readfirstlane(fce[h]);
                                for (int u = st0; u < en; u++) {
                                    const int j = __builtin_amdgcn_readfirstlane(fmember ? fmember[u] : u);
                                    const double4 q = ph[j];
                                    if (act && j != s &&
                                        within(p.x, p.y, p.z, q.x, q.y, q.z, 2.0 * fmax(hi, q.w))) {
                                        if (c < K) nbr[(int64_t)c * nf + s] = j;
                                        c++;
                                    }
                                }
                            }
                        }
                        if (__ballot(act && bd2 <= Rs * Rs) != 0) {
                            const int st1 = __builtin_amdgcn_readfirstlane(scs[h]);
                            if (st1 < 0) continue;
                            const int en = __builtin_amdgcn_readfirstlane(sce[h]);
                            for (int u = st1; u < en; u++) {
                                const int a = __builtin_amdgcn_readfirstlane(sperm[u]);
                                const double4 q = spr[a];
                                if (!(act && within(p.x, p.y, p.z, q.x, q.y, q.z, 2.0 * hi))) continue;
                                if (FILL) {
                                    if (w < cap) {
                                        fs_i[w] = s;
                                        fs_a[w] = a;
                                        fsF[w] = make_double4(0, 0, 0, 0);
                                    }
                                    w++;
                                } else {
                                    cs++;
                                }
                            }
                        }
                    }
        }
        if (!FILL && act) {
            cnt[s] = c;
            fscnt[s] = cs;
        }
        cmax = cmax > (unsigned long long)c ? cmax : (unsigned long long)c;
        ovf = ovf || c > K;
    }
    if (!FILL) {
        wave_max_atomic(&st[ST_FF_MAX], cmax);
        if (ovf) atomicOr(&st[ST_FF_OVERFLOW], 1ull);
    }
}

__global__ void k_fs_status(const int* __restrict__ fsoff, int nf, int64_t cap,
                            unsigned long long* st) {
    const long long tot = fsoff[nf];
    st[ST_FS_TOTAL] = (unsigned long long)tot;
    if (tot > cap) st[ST_FS_OVERFLOW] |= 1ull;
}

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
__global__ void k_rev_fill(const int* __restrict__ key, const int* __restrict__ total_p, int64_t cap,
                           const double4* __restrict__ kvel, const int* __restrict__ off,
                           int* __restrict__ cursor, int* __restrict__ idx) {
    int64_t tot = *total_p;
    if (tot > cap) tot = cap;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot;
         p += (int64_t)gridDim.x * blockDim.x) {
        int a = key[p];
        if (kvel && !(kvel[a].w > 0.0)) continue;
        int slot = off[a] + atomicAdd(&cursor[a], 1);
        idx[slot] = (int)p;
    }
}
__global__ void k_rev_sort(const int* __restrict__ off, int n, int* __restrict__ idx) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    int b = off[a], e = off[a + 1];
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

// interactionSolid2Solid::save (deviceStructs.h:693-706): keep the previous
// list's offsets, partners and springs for the history lookup
__global__ void k_ss_save(const int* __restrict__ ssoff, int ns, const int* __restrict__ ss_j,
                          const double* __restrict__ ssS, int64_t cap, int* __restrict__ hoff,
                          int* __restrict__ h_j, double* __restrict__ hS) {
    int64_t tot = ssoff[ns];
    if (tot > cap) tot = cap;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= ns;
         i += (int64_t)gridDim.x * blockDim.x)
        hoff[i] = ssoff[i];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
         i += (int64_t)gridDim.x * blockDim.x) {
        h_j[i] = ss_j[i];
#pragma unroll
        for (int k = 0; k < 9; k++) hS[9 * i + k] = ssS[9 * i + k];
    }
}

// setSolid2SolidInteractionsKernel (neighborSearch.cu:222-312): pass 0
// counts, pass 1 fills + carries the springs of the same (i, j) over
template <bool FILL>
__global__ void __launch_bounds__(kBS) k_ss_search(
    GridP g, Solids so, int ns, const int* __restrict__ scs, const int* __restrict__ sce,
    const int* __restrict__ sperm, int* __restrict__ sscnt, const int* __restrict__ ssoff,
    int64_t cap, int* __restrict__ ss_i, int* __restrict__ ss_j, double4* __restrict__ ssF,
    double4* __restrict__ ssT, double* __restrict__ ssS, const int* __restrict__ hoff,
    const int* __restrict__ h_j, const double* __restrict__ hS, unsigned long long* st) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (FILL && a == 0) {
        long long tot = ssoff[ns];
        st[ST_SS_TOTAL] = (unsigned long long)tot;
        if (tot > cap) atomicOr(&st[ST_SS_OVERFLOW], 1ull);
    }
    if (a >= ns) return;
    if (so.dummy[a] == 1) {
        if (!FILL) sscnt[a] = 0;
        return;
    }
    if (FILL && sscnt[a] == 0) return;
    const double4 p = so.pr[a];
    const double ra = so.eff[a];
    const int ca = so.clump[a];
    const bool fixa = so.vel[a].w < 1.e-10;
    int ax, ay, az;
    grid_pos(g, p.x, p.y, p.z, ax, ay, az);
    int c = 0;
    int64_t w = FILL ? (int64_t)ssoff[a] : 0;
    int hb = 0, he = 0;
    if (FILL) {
        hb = hoff[a];
        he = hoff[a + 1];
    }
    for (int zz = -1; zz <= 1; zz++)
        for (int yy = -1; yy <= 1; yy++)
            for (int xx = -1; xx <= 1; xx++) {
                const int h = (az + zz) * g.gy * g.gx + (ay + yy) * g.gx + (ax + xx);
                if (h < 0 || h >= g.ncell) continue;
                const int st0 = scs[h];
                if (st0 < 0) continue;
                const int en = sce[h];
                for (int t = st0; t < en; t++) {
                    const int b = sperm[t];
                    if (b <= a) continue;
                    if (so.dummy[b] == 1) continue;
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
__global__ void k_normals(GridP g, Solids so, int ns, const int* __restrict__ scs,
                          const int* __restrict__ sce, const int* __restrict__ sperm, double h) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ns) return;
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
                const int hh = (az + zz) * g.gy * g.gx + (ay + yy) * g.gx + (ax + xx);
                if (hh < 0 || hh >= g.ncell) continue;
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

}  // namespace

// ---------------------------------------------------------------- wrappers
void hash_fluid(hipStream_t s, const GridP& g, const double4* ph, const int* orig, int n,
                unsigned* key, long long* status) {
    hipLaunchKernelGGL(k_status_reset, dim3(1), dim3(ST_WORDS), 0, s, (unsigned long long*)status);
    if (n > 0)
        hipLaunchKernelGGL(k_hash_fluid, dim3(blocks_for(n)), dim3(kBS), 0, s, g, ph, orig, n, key,
                           (unsigned long long*)status);
}

void hash_solid(hipStream_t s, const GridP& g, const double4* pr, int n, unsigned* key,
                long long* status) {
    if (n > 0)
        hipLaunchKernelGGL(k_hash_solid, dim3(blocks_for(n)), dim3(kBS), 0, s, g, pr, n, key,
                           (unsigned long long*)status);
}

void reorder_fluid(hipStream_t s, const FluidSet& src, const FluidSet& dst, const int* perm,
                   int* slot_of, int* member, int n, bool mixed, int cur) {
    if (n <= 0) return;
    if (mixed)
        hipLaunchKernelGGL(k_reorder<float4>, dim3(blocks_for(n)), dim3(kBS), 0, s, src, dst, perm,
                           slot_of, member, n, cur);
    else
        hipLaunchKernelGGL(k_reorder<double4>, dim3(blocks_for(n)), dim3(kBS), 0, s, src, dst, perm,
                           slot_of, member, n, cur);
}

void member_slots(hipStream_t s, const int* perm, const int* slot_of, int* member, int n) {
    if (n > 0)
        hipLaunchKernelGGL(k_member, dim3(blocks_for(n)), dim3(kBS), 0, s, perm, slot_of, member, n);
}

void fluid_blocks(hipStream_t s, const int* fcs, const int* fce, int ncell, int* nb, int* off,
                  int2* blocks, int* scan_tmp) {
    hipLaunchKernelGGL(k_blk_count, dim3(blocks_for(ncell)), dim3(kBS), 0, s, fcs, fce, ncell, nb);
    exclusive_scan(s, nb, off, ncell, scan_tmp);
    hipLaunchKernelGGL(k_blk_fill, dim3(blocks_for(ncell)), dim3(kBS), 0, s, fcs, fce, ncell,
                       (const int*)off, blocks);
}

static int wave_grid(int nf) {
    // about one wavefront per cell block (4 per workgroup); waves stride over
    // the block list, which lives on the device
    int b = (nf + 127) / 128;
    return b < 1 ? 1 : (b > 16384 ? 16384 : b);
}

void fluid_search(hipStream_t s, const GridP& g, const FluidSet& f, int nf, const int* fcs,
                  const int* fce, const int* fmember, const int* scs, const int* sce,
                  const int* sperm, const Solids& so, double hmax, Lists& L, const int2* blocks,
                  const int* nblk, long long* status) {
    if (nf <= 0) return;
    hipLaunchKernelGGL(k_fluid_blocks<false>, dim3(wave_grid(nf)), dim3(kBS), 0, s, g, f.ph, nf, fcs,
                       fce, fmember, scs, sce, sperm, so.pr, hmax, L.K, L.nbr, L.cnt, L.fscnt, blocks,
                       nblk, (const int*)nullptr, (int64_t)0, (int*)nullptr, (int*)nullptr,
                       (double4*)nullptr, (unsigned long long*)status);
}

void fluid_solid_fill(hipStream_t s, const GridP& g, const FluidSet& f, int nf, const int* fcs,
                      const int* fce, const int* fmember, const int* scs, const int* sce,
                      const int* sperm, const Solids& so, Lists& L, const int2* blocks,
                      const int* nblk, long long* status) {
    if (nf <= 0) return;
    hipLaunchKernelGGL(k_fs_status, dim3(1), dim3(1), 0, s, (const int*)L.fsoff, nf, L.fs_cap,
                       (unsigned long long*)status);
    hipLaunchKernelGGL(k_fluid_blocks<true>, dim3(wave_grid(nf)), dim3(kBS), 0, s, g, f.ph, nf, fcs,
                       fce, fmember, scs, sce, sperm, so.pr, 0.0, L.K, L.nbr, L.cnt, L.fscnt, blocks,
                       nblk, (const int*)L.fsoff, L.fs_cap, L.fs_i, L.fs_a, L.fsF,
                       (unsigned long long*)status);
}

static void reverse_build(hipStream_t s, const int* key, const int* total_p, int64_t cap,
                          const double4* kvel, int nkeys, int* cnt, int* off, int* idx,
                          int* scan_tmp) {
    if (nkeys <= 0) return;
    fill_int(s, cnt, nkeys, 0);
    int gb = blocks_for(cap);
    if (gb > 2048) gb = 2048;
    if (cap > 0) hipLaunchKernelGGL(k_rev_count, dim3(gb), dim3(kBS), 0, s, key, total_p, cap, kvel, cnt);
    exclusive_scan(s, cnt, off, nkeys, scan_tmp);
    fill_int(s, cnt, nkeys, 0);
    if (cap > 0)
        hipLaunchKernelGGL(k_rev_fill, dim3(gb), dim3(kBS), 0, s, key, total_p, cap, kvel,
                           (const int*)off, cnt, idx);
    hipLaunchKernelGGL(k_rev_sort, dim3(blocks_for(nkeys)), dim3(kBS), 0, s, (const int*)off, nkeys, idx);
}

void reverse_movable(hipStream_t s, const int* pair_key, const int* pair_total, int64_t cap,
                     const double4* key_vel, int nkeys, int* cnt, int* off, int* idx,
                     int* scan_tmp) {
    reverse_build(s, pair_key, pair_total, cap, key_vel, nkeys, cnt, off, idx, scan_tmp);
}

void solid_save_history(hipStream_t s, Lists& L, int ns) {
    if (ns <= 0) return;
    int gb = blocks_for(L.ss_cap > ns + 1 ? L.ss_cap : ns + 1);
    if (gb > 2048) gb = 2048;
    hipLaunchKernelGGL(k_ss_save, dim3(gb), dim3(kBS), 0, s, (const int*)L.ssoff, ns,
                       (const int*)L.ss_j, (const double*)L.ssS, L.ss_cap, L.hoff, L.h_j, L.hS);
}

void solid_search(hipStream_t s, const GridP& g, const Solids& so, int ns, const int* scs,
                  const int* sce, const int* sperm, Lists& L, int* scan_tmp, long long* status) {
    if (ns <= 0) return;
    hipLaunchKernelGGL(k_ss_search<false>, dim3(blocks_for(ns)), dim3(kBS), 0, s, g, so, ns, scs, sce,
                       sperm, L.sscnt, (const int*)L.ssoff, L.ss_cap, L.ss_i, L.ss_j, L.ssF, L.ssT,
                       L.ssS, (const int*)L.hoff, (const int*)L.h_j, (const double*)L.hS,
                       (unsigned long long*)status);
    exclusive_scan(s, L.sscnt, L.ssoff, ns, scan_tmp);
    hipLaunchKernelGGL(k_ss_search<true>, dim3(blocks_for(ns)), dim3(kBS), 0, s, g, so, ns, scs, sce,
                       sperm, L.sscnt, (const int*)L.ssoff, L.ss_cap, L.ss_i, L.ss_j, L.ssF, L.ssT,
                       L.ssS, (const int*)L.hoff, (const int*)L.h_j, (const double*)L.hS,
                       (unsigned long long*)status);
}

void reverse_solid(hipStream_t s, Lists& L, int ns, int* scan_tmp) {
    reverse_build(s, L.ss_j, L.ssoff + ns, L.ss_cap, nullptr, ns, L.ssr_cnt, L.ssr_off, L.ssr_idx,
                  scan_tmp);
}

void solid_normals(hipStream_t s, const GridP& g, const Solids& so, int ns, const int* scs,
                   const int* sce, const int* sperm, double h) {
    if (ns > 0)
        hipLaunchKernelGGL(k_normals, dim3(blocks_for(ns)), dim3(kBS), 0, s, g, so, ns, scs, sce,
                           sperm, h);
}

}  // namespace dsph
