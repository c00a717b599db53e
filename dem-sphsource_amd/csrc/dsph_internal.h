// dsph_internal.h — device data layout and kernel-launch interface shared by
// the HIP translation units of libdsph.  Not part of the public ABI.
//
// Layout in HBM (all per-particle arrays are SoA of 16/32-byte records so a
// neighbour gather is one aligned vector load):
//   fluid  — kept in cell-sorted ("slot") order, two copies ping-ponged on
//            reorder: ph (x,y,z,h) f64x4, vr (vx,vy,vz,rho0) f64x4,
//            drho f64 (Jacobi ping-pong), cn (c,nu) f64x2, sumw0 f64, acc f64x4,
//            orig (slot -> caller index) i32; neighbour records in the pair
//            precision (see FluidSet).
//   solids — caller order: pr (x,y,z,r), vel (vx,vy,vz,1/m), acc, w, torque,
//            normal (f64x4), eff radius f64, material/cluster/clump/dummy i32.
//   lists  — fluid–fluid: Verlet superset rows (both directions) in ELL
//            column-major layout sup[k*NF + slot], plus exact rows nbr for
//            particles walked on the reference stencil; fluid–solid and
//            solid–solid: CSR by pointed particle with device-side totals.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dsph.h"
#include "dsph_math.h"

namespace dsph {

// status words (mirrors include/dsph.h)
enum {
    ST_FF_OVERFLOW = 0,
    ST_FS_OVERFLOW = 1,
    ST_SS_OVERFLOW = 2,
    ST_OUT_OF_DOMAIN = 3,
    ST_FF_MAX = 4,
    ST_FS_TOTAL = 5,
    ST_SS_TOTAL = 6,
    ST_SUP_OVERFLOW = 7,
    ST_REBUILDS = 8,
    ST_SUP_MAX = 9,        // longest (group) superset row of this search's build
    ST_NONFINITE = 10,     // fluid + live solids with a NaN/inf position at this search
    ST_WORDS = 12
};

// device control words of the neighbour search
enum {
    CTL_DISP_F = 0,   // max fluid displacement since the last superset build (f64 bits)
    CTL_DISP_S = 1,   // max solid displacement (f64 bits)
    CTL_FORCE = 2,    // host request: rebuild at the next search
    CTL_REBUILD = 3,  // this search rebuilds the superset
    CTL_NEAR_SPARE = 4,    // some fluid particle's stencil reaches the spare cell, or it is in it
    CTL_SORT = 5,          // this search sorts the grids (rebuild or exact walks)
    CTL_TICKET = 6,        // blocks of the solid hash done (the last one decides)
    CTL_WORDS = 7
};

// the reference's spatialGrid (deviceStructs.h:976-1014) of one scene.  A
// batch of independent scenes keeps one grid per scene; its cells occupy the
// global cell indices [cell0, cell0 + ncell), the spare cell last.  Hashes
// stored on the device are global (cell0 + the scene's own hash); stencils
// and spare tests are evaluated in the scene's own cell coordinates, so each
// scene's lists are those of the scene run alone.
struct GridP {
    double minx, miny, minz;
    double maxx, maxy, maxz;
    double csx, csy, csz;
    int gx, gy, gz, ncell;
    int cell0;      // first global cell index of this scene
    int collide;    // gx < 3 or gy < 3: some of the 27 linear stencil offsets coincide
    int sw;         // half-width in cells of the superset build stencil
    int walk_only;  // the superset stencil's offsets are not distinct: walk every row
    double pad;     // conservative enlargement of cell boxes for pruning
    double hmax;    // largest fluid smoothing length of the scene
    double skin;    // Verlet skin of the scene's fluid superset (absolute)
    double iskin2;  // 1 / skin^2 (1e300 without a skin)
};

// the grids of all scenes and the scene of every particle
struct Grids {
    const GridP* g = nullptr;       // per scene, device
    const int* fscene = nullptr;    // fluid slot -> scene (slot order)
    const int* sscene = nullptr;    // solid -> scene
    int nscene = 1;
    GridP g0{};                     // scene 0 by value (the one-scene case)
};
__device__ __forceinline__ const GridP& grid_at(const Grids& G, int scene) {
    return G.nscene == 1 ? G.g0 : G.g[scene];
}
__device__ __forceinline__ int fluid_scene(const Grids& G, int slot) {
    return G.nscene == 1 ? 0 : G.fscene[slot];
}
__device__ __forceinline__ int solid_scene(const Grids& G, int i) {
    return G.nscene == 1 ? 0 : G.sscene[i];
}

// Per-scene time steps of a batch (dsph_set_scene_time_steps): with sdt
// set, a phase's dt argument is its factor (1/2, 1, the fluid gap) and
// particle i of scene scene[i] steps by that factor times sdt[scene[i]]
// (the product of the one-scene run: factor and step commute exactly);
// without, by dt itself.
__device__ __forceinline__ double step_of(const double* sdt, const int* scene, int i, double dt) {
    return sdt ? dt * sdt[scene[i]] : dt;
}

struct FluidSet {
    double4* ph = nullptr;              // x, y, z, h
    double4* vr = nullptr;              // vx, vy, vz, rho0 (FP64 master)
    double* drho[2] = {nullptr, nullptr};  // FP64 master, Jacobi ping-pong
    double2* cn = nullptr;              // c, nu
    double* sumw0 = nullptr;
    double* alpha = nullptr;            // artificial viscosity coefficient
    int* scene = nullptr;               // scene of each slot (batches)
    double4* acc = nullptr;
    int* orig = nullptr;                // slot -> caller index
    // neighbour records, ping-ponged with drho so the Jacobi update never
    // overwrites what neighbours read.  Two quads per slot:
    //   FP64  (64 B): (x, y, z, table index), (v, dRho) as f64x4;
    //   mixed (32 B): (x, y, z as 32-bit fixed point, table index),
    //                 (v, dRho) as f32x4.
    void* rec[2] = {nullptr, nullptr};
    int* tid = nullptr;           // property-table index of each slot
    // fluid property table, shared by all slots (and both sets): one entry
    // per distinct (h, c, rho0, nu) of the scene(s), (h, c, m, rho0) as
    // f64x4 / f32x4 in tab and nu in tabnu
    const void* tab = nullptr;
    const void* tabnu = nullptr;
    // fixed-point map of the mixed records: q = rint((x - qo) * qi), one unit
    // = qr metres per axis
    double qo[3] = {0, 0, 0};
    double qi[3] = {0, 0, 0};
    float qr[3] = {0, 0, 0};
};

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; this
// maps the hardware block id to a logical one so each XCD (and its L2) works
// on a contiguous 1/8 of the cell-sorted particles.
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

struct Solids {
    double4* pr = nullptr;   // x, y, z, radius
    double4* vel = nullptr;  // vx, vy, vz, inverse mass
    double4* acc = nullptr;
    double4* w = nullptr;
    double4* tq = nullptr;
    double4* nrm = nullptr;
    double* eff = nullptr;
    int* mat = nullptr;
    int* clus = nullptr;
    int* clump = nullptr;
    int* dummy = nullptr;
    int* active = nullptr;  // indices of non-dummy solids
    int nactive = 0;
    // the solids whose state the DEM phases can change: non-dummies, and
    // any dummy with mass, a velocity or a clump (wall dummies are static)
    int* live = nullptr;
    int nlive = 0;
    // the mixed path's fluid–solid records, two 16-byte quads per solid:
    // (x, y, z in the fluid's fixed-point map, radius as f32) and (normal
    // as f32, flag); flag 0: a static solid with zero velocity and
    // acceleration, whose record holds until the solids are uploaded, their
    // normals computed or a state loaded; flag 1: the pair terms read the
    // solid's FP64 state (solids the DEM phases move, and solids outside
    // the fixed-point map)
    uint4* mrec = nullptr;
};

struct Clumps {
    double4* pos = nullptr;
    double4* vel = nullptr;  // w = inverse mass
    double4* acc = nullptr;
    double4* w = nullptr;
    double4* tq = nullptr;
    double4* q = nullptr;    // q0, q1, q2, q3
    double* invI = nullptr;  // 6 per clump (body frame)
    int* start = nullptr;
    int* end = nullptr;
};

// Every fluid row array (nbr, sup) is allocated with kRowPad columns past its
// capacity: the pair loops read row indices ahead of use, up to 2 batches
// past a row's end, without clamping.
constexpr int kRowPad = 16;

// grow[slot] names the slot r whose storage holds the slot's superset row;
// bits 28+ hold lg, the row's interleave: a group of 2^lg consecutive slots
// r, r + 1, ... (all of one group) stores its shared row across their
// storage, column c at (c >> lg) * nf + r + (c & (2^lg - 1)), so the
// group's lanes read dense index lines.  lg = 0: column c at c * nf + r.
// Slot counts stay below 2^26 (the pair loops' 32-bit record offsets).
constexpr int kRowSlotBits = 28;
__host__ __device__ __forceinline__ int row_slot(int g) { return g & ((1 << kRowSlotBits) - 1); }
__host__ __device__ __forceinline__ int row_lg(int g) { return g >> kRowSlotBits; }
__host__ __device__ __forceinline__ int64_t row_at(int g, int c, int nf) {
    const int lg = row_lg(g);
    return (int64_t)(c >> lg) * nf + row_slot(g) + (c & ((1 << lg) - 1));
}

struct Lists {
    // fluid–fluid ELL
    int K = 0;
    int* nbr = nullptr;
    int* cnt = nullptr;
    // fluid–solid CSR by fluid slot
    int64_t fs_cap = 0;
    int* fscnt = nullptr;
    int* fsoff = nullptr;  // NF + 1
    int* fs_a = nullptr;   // solid index
    int* fs_own = nullptr; // fluid slot of the pair
    double4* fsF = nullptr;
    // the pairs' terms of the coming fluid pass (one R, or an R4, per pair),
    // written by the pair kernel before the pass reads them
    void* fsT = nullptr;
    // reverse fluid–solid for movable solids: CSR by solid
    int* fsr_cnt = nullptr;
    int* fsr_off = nullptr;  // NS + 1
    int* fsr_idx = nullptr;  // fs pair ids, capacity fs_cap
    // Verlet supersets, ELL column-major like nbr (rebuilt only when needed)
    int supK = 0;
    int* sup = nullptr;
    int* sup_cnt = nullptr;
    // superset rows are shared by groups of consecutive sorted particles
    // (the union of their supersets): grow[slot] = the slot whose row and
    // count hold the group's row
    int* grow = nullptr;
    int fsupK = 0;
    int* fsup = nullptr;
    int* fsup_cnt = nullptr;
    // members of the fluid–solid list among the first 128 candidates of
    // each slot's solid superset row (two words per slot)
    unsigned long long* fsm = nullptr;
    // solid–solid CSR by solid index (the reference's own list)
    int64_t ss_cap = 0;
    int* sscnt = nullptr;
    int* ssoff = nullptr;  // NS + 1
    int* ss_i = nullptr;
    int* ss_j = nullptr;
    double4* ssF = nullptr;
    double4* ssT = nullptr;
    double* ssS = nullptr;  // 9 per contact: sliding, rolling, torsion
    // Offsets of the solid-keyed CSRs (ssoff, hoff, ssr_off, fsr_off) are
    // kept only for solids that can have pairs (the active ones; movable
    // ones for fsr), plus the total at index NS: a segment is
    // [off[i], off[i] + cnt[i]).
    // previous step's list (spring history)
    int* hoff = nullptr;    // NS + 1
    int* hcnt = nullptr;    // NS
    int* h_j = nullptr;
    double* hS = nullptr;
    // reverse solid–solid: CSR by pointing solid
    // reverse-CSR scratch: per-solid counts (zero between builds) and cursors
    int* rev_tmp = nullptr;
    int* rev_cur = nullptr;
    int* rev_tmp2 = nullptr;  // the second list of a fused build (reverse_both)
    int* rev_cur2 = nullptr;
    int* ssr_cnt = nullptr;
    int* ssr_off = nullptr;
    int* ssr_idx = nullptr;
};

// Tiles of the mixed pair loops: runs of up to kTileSlots consecutive slots
// of one scene.  Slots are in Morton order after every superset build, so a
// tile is a compact block of particles, and the union of its rows'
// candidates is a few times its size.  At every build each tile gets that
// union (uni, ascending slots, at most kTileCap) and each of its rows the
// same candidates in the same order, minus the particle itself, as LDS
// entries of the union (row, u16; rcnt entries), four consecutive entries of
// a row in one 8-byte quad, quads ELL column-major: entry c of slot s at
// ((c / 4) * nf + s) * 4 + c % 4.  The pair loops stage the union's records
// in LDS once per tile and walk the local rows there.  ucnt < 0: the union
// did not fit, and the tile walks its global rows.
constexpr int kTileSlots = 512;
constexpr int kTileCap = 2432;  // 76 KiB of mixed records in LDS
// LDS entry of union position p: p with its bits 0-2 XOR bits 3-5, a
// permutation within each aligned block of 64 entries (kTileCap, the
// sentinel, maps to itself).  A 16-byte LDS read serves eight lanes per
// pass, one per bank group (entry mod 8); neighbouring slots' rows step
// through the union a few positions per slot, and a step of 2, 4 or 8 put
// several lanes of a pass on one bank group, which the XOR spreads.
static_assert(kTileCap % 64 == 0, "the swizzle permutes aligned blocks of 64 entries");
__host__ __device__ constexpr int tile_swz(int p) { return p ^ ((p >> 3) & 7); }
// entries per slot of the local rows: the superset capacity, the pad read
// ahead, rounded up to whole quads
__host__ __device__ constexpr int tile_row_cols(int supK) { return (supK + kRowPad + 3) / 4 * 4; }
__host__ __device__ __forceinline__ int64_t tile_row_at(int c, int s, int nf) {
    return ((int64_t)(c >> 2) * nf + s) * 4 + (c & 3);
}
struct Tiles {
    int n = 0;                         // tiles (0: the tiled loops are off)
    const int2* range = nullptr;       // [first slot, end slot) of each tile
    int* uni = nullptr;                // n * kTileCap slots
    int* ucnt = nullptr;               // n
    unsigned short* row = nullptr;     // tile_row_cols(supK) * nf
    int* rcnt = nullptr;               // nf
};

struct Bonds {
    int n = 0;
    int* i = nullptr;
    int* j = nullptr;
    int* on = nullptr;
    double4* nrm = nullptr;
    double4* pt = nullptr;
    double4* Fs = nullptr;
    double4* Tb = nullptr;
    double* Fn = nullptr;
    double* Tt = nullptr;
};

// ---- launch wrappers (implemented in the .hip files) -----------------------
// generic
// gate: when non-null, the kernels do nothing unless *gate != 0 (a device
// flag, so the decision can be taken inside a captured step)
using Gate = const unsigned long long*;
void fill_int(hipStream_t s, int* p, int64_t n, int v, Gate gate = nullptr);
void copy_int(hipStream_t s, int* dst, const int* src, int64_t n);
// exclusive scan of n ints into out[0..n] (out[n] = total); tmp >= scan_tmp_size(n)
int64_t scan_tmp_size(int64_t n);
void exclusive_scan(hipStream_t s, const int* in, int* out, int64_t n, int* tmp, Gate gate = nullptr);
// exclusive scan over a list of indices (ascending; all others count 0), in
// one workgroup: off[keys[r]] = sum of the counts of keys[0..r), off[n_all]
// = total; with cnt_out the counts move from cnt (left zero) to cnt_out;
// with cursor, cursor[key] = off[key]
void scan_keys(hipStream_t s, const int* keys, int nkeys, int* cnt, int* cnt_out, int* off, int* cursor,
               int n_all);
// lists at most this long are scanned by scan_keys (one workgroup)
constexpr int kKeyScanMax = 16384;
// stable LSD radix sort of (key, index) pairs; keys < 2^bits; vals out = perm
int64_t radix_tmp_size(int64_t n);
void radix_sort_iota(hipStream_t s, const unsigned* keys_in, int64_t n, int bits,
                     unsigned* keys_out, int* perm_out, unsigned* tmp, Gate gate = nullptr);
// start/end of each cell over sorted keys; with inv, the keys are cell
// ranks and inv maps a rank to its cell
void cell_spans(hipStream_t s, const unsigned* sorted, int64_t n, int* start, int* end,
                int ncell, const unsigned* inv, Gate gate = nullptr, int shift = 0);

// search (k_search.hip)
struct SearchArgs {
    Grids G;
    int ncell;              // cells of all scenes
    int group;              // particles sharing one superset row (1, 2, 4 or 8)
    int nf, ns;
    const double4* ph;      // fluid position + h, slot order
    const int* fcs;         // fluid cell ranges over sorted positions
    const int* fce;
    const int* fmember;     // slot of each sorted position
    const int* forig;       // caller index of each fluid slot
    const unsigned* hslot;  // reference hash of each fluid slot
    const unsigned char* cspare;  // per cell: its 27-offset stencil reaches the spare cell
    const int* near_out;    // per cell: within one cell of an out-of-domain fluid particle (this step)
    // the solids the fluid sees, as two grids: the static solids (wall
    // dummies, sorted once) and the active grid of the non-dummies (sorted
    // every step for the DEM search); without the second (acs null) the
    // first holds every solid
    const int* scs;         // cell ranges and solid indices of the first grid
    const int* sce;
    const int* sperm;
    const int* acs;         // active grid: cell ranges over positions whose
    const int* ace;         // solid is active[aperm[position]]
    const int* aperm;
    const int* active;
    const unsigned* skey;   // reference hash of each solid
    const double4* spr;
    double4* fsp;           // fluid positions of the last fluid search, slot order
    double4* fxb;           // positions at the last build: fluid slots
    double4* sxb;           // and solids (the live ones change)
    const int* slive;
    int nslive;
    Lists L;
    unsigned long long* ctl;
    unsigned long long* st;
    // fluid–solid count of each tile of 256 slots (the list pass)
    int* tsum;
    // the pair-loop tiles, rebuilt with the superset (n = 0: none)
    Tiles T;
};
// also the reference hash of every slot, and CTL_NEAR_SPARE when a particle's
// cell has a stencil reaching the spare cell; with fxb, the largest squared
// displacement since the last build (in skin^2 units) of each workgroup of
// 256 particles into bmax (blocks_for(n) words)
// The fluid sort key: the cell's Morton rank above the Morton code of the
// particle's sub-cell (sub_bits / 3 bits per axis), so a cell's members are
// contiguous and themselves in z-order.
// With near_out (fluid steps), the cells within one cell of each fluid
// particle out of the domain are marked (the array is cleared first); a
// particle whose build position was out of the domain and that is back in
// it requests a rebuild.
void hash_fluid(hipStream_t s, const Grids& G, const double4* ph, const int* orig, int n, int sub_bits,
                const unsigned* crank, const unsigned char* cspare, unsigned* key_orig,
                unsigned* rank_orig, unsigned* hslot, const double4* fxb, unsigned long long* bmax,
                int* near_out, int ncell, unsigned long long* ctl, long long* status);
// calculatePointHash of n caller points (interleaved xyz doubles) on one
// grid (a scene-local grid: cell0 = 0), into key
void hash_points(hipStream_t s, const GridP& g, const double* xyz, int n, unsigned* key);
// the device address of the allocation-free entry points' scratch on the
// current device (api_grid.hip); 0 on success
int noalloc_scratch_setup(int device);
// the reference hash of every solid (and the out-of-domain count)
void hash_solid(hipStream_t s, const Grids& G, const double4* pr, int n, unsigned* key,
                long long* status);
// The per-step solid hash: only the live solids move, so only their hashes
// are recomputed (the others' were set at upload; static_out of them are
// out of the domain).  With decide, also the live solids' displacement
// since the last build, and then the rebuild decision, taken by the last
// workgroup to finish once every displacement and out-of-domain count is
// in: CTL_REBUILD when d_f + max(d_f, d_s) >= skin (displacements in units
// of each scene's skin) or always != 0, and CTL_SORT = rebuild, exact walks
// or always_sort.
void hash_live_decide(hipStream_t s, const Grids& G, const double4* pr, const int* live, int nlive,
                      unsigned* key, unsigned long long static_out, bool decide, const double4* sxb,
                      const unsigned long long* fbmax, int nf, int always, int always_sort,
                      unsigned long long* ctl, long long* status);
void reorder_fluid(hipStream_t s, const FluidSet& f, const FluidSet& scratch, const int* perm,
                   int* slot_of, int n, bool mixed, int cur, const unsigned long long* ctl);
// after a sort (CTL_SORT): member[t] = slot of sorted position t and its
// hash; with reorder, after a rebuild (CTL_REBUILD) the reordered set is
// copied back from scratch
void after_sort(hipStream_t s, const FluidSet& f, const FluidSet& scratch, bool reorder, bool mixed, int cur,
                const unsigned* sorted, int n, int sub_bits, const int* perm, const int* slot_of,
                const unsigned* cinv, int* member, unsigned* hslot, const unsigned long long* ctl);
// chunks: (first, end) sorted positions of <= 64 consecutive particles of
// one scene, one wavefront each; also stores the build positions (fxb, and
// sxb of the live solids)
void fluid_superset(hipStream_t s, const SearchArgs& a, const int2* chunks, int nchunk);
// after a superset build (CTL_REBUILD): each tile's union and local rows
void tile_rows(hipStream_t s, const SearchArgs& a);
// exact fluid neighbour rows of all particles for read-back (ELL, stride nf,
// capacity max(K, supK)); out_cnt = -1 marks a walked row that overflowed
void fluid_exact_rows(hipStream_t s, const SearchArgs& a, int* out, int* out_cnt);
void fluid_filter(hipStream_t s, const SearchArgs& a, bool zero_fs);
// fluid–solid pairs by movable solid (fsr_*)
void reverse_movable(hipStream_t s, const int* pair_key, const int* pair_off_total, int64_t cap,
                     const double4* key_vel, const Solids& so, int ns, Lists& L, int* scan_tmp);
void solid_save_history(hipStream_t s, const Solids& so, Lists& L, int ns);
// the active grid; up to kSmallSort solids by counting ranks (resetting
// only the cells the previous sort filled; rank_acc: 3 zeroed ints per
// active solid, zero again afterwards)
constexpr int kSmallSort = 4096;
void active_grid(hipStream_t s, const Solids& so, const unsigned* skey, int ncell, int key_bits,
                 unsigned* akey, unsigned* akey_sorted, int* aperm, int* acs, int* ace,
                 unsigned* radix_tmp, int* rank_acc);
void solid_search(hipStream_t s, const Grids& G, const Solids& so, int ns, const int* acs,
                  const int* ace, const int* aperm, Lists& L, int* scan_tmp, long long* status);
void reverse_solid(hipStream_t s, const Solids& so, Lists& L, int ns, int* scan_tmp);
// both of the above in one chain of four launches (the fluid–solid list as
// reverse_movable takes it)
void reverse_both(hipStream_t s, const int* pair_key, const int* pair_total, int64_t cap,
                  const double4* key_vel, const Solids& so, Lists& L, int ns, int* scan_tmp);
// wall normals with the smoothing length of each solid's scene (grid hmax)
void solid_normals(hipStream_t s, const Grids& G, const Solids& so, int ns, const int* scs,
                   const int* sce, const int* sperm);

// fluid (k_fluid.hip)
// the fluid properties when every fluid particle shares them
struct UniformFluid {
    double h = 0, c = 0, m = 0, rho0 = 0, nu = 0;
};
struct FluidArgs {
    FluidSet f;
    int uni;           // all fluid particles share one (h, c, rho0, nu): u
    UniformFluid u;
    int nf;
    int bs;   // workgroup size of the pair loops (tunable, <= 512)
    int cur;  // current drho buffer
    Lists L;
    Solids so;
    double gx, gy, gz;
    uint32_t flags;
    Tiles T;  // n > 0: the mixed pair loops run tiled
    const double* sdt = nullptr;  // per-scene time steps (step_of, scenes f.scene), or none
    int ns = 0;  // solids (0: no fluid–solid pairs)
};
void fluid_records(hipStream_t s, const FluidArgs& a, int buf, bool mixed);
void density_reinit(hipStream_t s, const FluidArgs& a, int i_step, bool mixed);
// writes 1-cur; drift != 0 also moves each particle by drift v
void continuity(hipStream_t s, const FluidArgs& a, double dt, bool mixed, double drift);
// kick != 0 also kicks and drifts each particle into the other buffer (flip cur after)
void momentum(hipStream_t s, const FluidArgs& a, bool mixed, bool write_fs_force, double kick);
void fluid_solid_force(hipStream_t s, const FluidArgs& a, bool mixed);
// the solids' mixed fluid–solid records (Solids::mrec) from their state
void solid_records(hipStream_t s, const Solids& so, int ns, const FluidSet& f);

// Frames (k_frame.hip): the arrays of one VTU frame in a device staging
// block, FP32 (Int32 for indices), caller order
struct FrameOut {
    float *fpos, *fh, *fvel, *frho, *fp;                     // fluid: 3, 1, 3, 1, 1 per particle
    float *spos, *srad;                                      // solid: 3, 1
    int *sclus, *sclump;                                     //        1, 1
    float *svel, *sw, *snrm;                                 //        3, 3, 3
    float* pairs;                                            // pair:  3, 3, then (int) 1, 1, as
                                                             //        consecutive arrays of the n pairs
    long long* head;                                         // [0]: n, pairs packed (-1: list overflowed)
};
// packs the species (DSPH_FRAME_* bits) of the current state
void frame_pack(hipStream_t s, unsigned species, const FluidSet& f, int nf, int cur, const Solids& so, int ns,
                const Lists& L, int64_t pair_cap, const FrameOut& o);

// DEM (k_dem.hip)
struct DemArgs {
    Solids so;
    int ns;
    Clumps c;
    int nc;
    Lists L;
    Bonds b;
    const double4* wpt;
    const double4* wn;
    const int* wmat;
    int nw;
    // walls of scene k are [woff[k], woff[k+1]); a solid meets only its
    // scene's walls and keeps 9 history doubles per wall of its scene
    const int* woff;
    const int* sscene;
    int nscene;
    int nwmax;      // most walls of any scene
    double* whist;  // 9 * ns * nwmax
    // cylinder walls, laid out like the infinite walls: (point, radius) and
    // (unit axis, 0) and (angular velocity, 0) per wall, walls of scene k
    // are [yoff[k], yoff[k+1]), 9 history doubles per solid and wall
    const double4* ypt;
    const double4* yax;
    const double4* yw;
    const int* ymat;
    int ny;
    const int* yoff;
    int nymax;
    double* yhist;  // 9 * ns * nymax
    ContactTables tb;
    double gx, gy, gz;
    uint32_t flags;
    const double* sdt = nullptr;  // per-scene time steps (step_of), or none
};
void dem_contacts(hipStream_t s, const DemArgs& a, double dt);
void dem_bonds(hipStream_t s, const DemArgs& a, double dt);
// contact sums of the live solids, then their infinite and cylinder walls
void dem_solid_sum(hipStream_t s, const DemArgs& a, double dt);
// kick and drift of the live solids in one pass (scenes without clumps)
void dem_solid_kick_drift(hipStream_t s, const DemArgs& a, double dt, bool kick_first);
void dem_clump_sum(hipStream_t s, const DemArgs& a);
void dem_solid_kick(hipStream_t s, const DemArgs& a, double dt);
void dem_solid_drift(hipStream_t s, const DemArgs& a, double dt);
void dem_clump_kick(hipStream_t s, const DemArgs& a, double dt);
void dem_clump_drift(hipStream_t s, const DemArgs& a, double dt);
// bonds from the initial contact list (addBondData): flag[k] = contact k
// becomes a bond (capacity ss_cap); then, with off the exclusive scan of the
// flags, the bonds written at their offsets (b sized for the total)
void dem_bond_flags(hipStream_t s, const DemArgs& a, int* flag);
void dem_bond_fill(hipStream_t s, const DemArgs& a, const int* off, const Bonds& b);
void dem_external_force(hipStream_t s, const DemArgs& a, int idx, double fx, double fy, double fz);
void dem_global_damping(hipStream_t s, const DemArgs& a, int idx, double cd);

inline int blocks_for(int64_t n, int bs = 256) {
    int64_t b = (n + bs - 1) / bs;
    return (int)(b < 1 ? 1 : b);
}

}  // namespace dsph
