// api.hip — the C ABI of include/dsph.h: context, device memory, the phase
// sequence of one timestep, HIP-graph replay and read-back.
//
// Replaces the reference's DeviceData ownership and phase drivers:
// dataBase::buildDeviceData / setSpatialGrids (core/dataBase.cpp:415-449),
// neighborSearch (core/neighborSearch.cu:314-381), fluidIntegrate
// (core/integrate.cu:385-421), the DEM phase drivers (:940-998),
// solverBase::initialize/update (core/solverBase.cpp:13-67) and the
// uploadState read-backs (core/deviceStructs.h).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dsph.h"
#include "dsph_internal.h"

using namespace dsph;

namespace {

struct ApiError {
    int code;
    std::string msg;
};

#define HIPCHK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess)                                                           \
            throw ApiError{DSPH_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)}; \
    } while (0)

struct StepGraph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int dcur_after = 0;
};

// kernel-family timers (HIP events on the context stream)
// (the frame copy's on the frames' copy stream)
enum Fam { F_GRID = 0, F_FSEARCH, F_FSLIST, F_SSEARCH, F_REINIT, F_CONT, F_MOM, F_FSTREAM, F_DEM, F_FPACK,
           F_FCOPY, F_NFAM };
const char* kFamNames[F_NFAM] = {"grid_hash_sort", "fluid_search",  "fluid_solid_list",
                                 "solid_search",   "density_reinit", "continuity",
                                 "momentum",       "fluid_stream",  "dem",
                                 "frame_pack",     "frame_copy"};

}  // namespace

namespace {
struct Merged;
}

struct dsph_ctx {
    // the scene as uploaded (the static arrays dsph_append_scene uploads
    // again), or null after a batch
    std::unique_ptr<Merged> host;
    int dev = 0;
    hipStream_t st = nullptr;
    int prec = DSPH_PREC_MIXED;
    uint32_t flags = 0;
    std::string err;
    bool uploaded = false, initialized = false;
    int nf = 0, ns = 0, nc = 0, nw = 0;
    // scenes: one grid each; particles of scene k are the caller indices
    // [foff[k], foff[k+1]) (fluid), [soff[k], soff[k+1]) (solid), ...
    int nscene = 1;
    std::vector<GridP> grids;
    GridP* dgrids = nullptr;
    int ncell = 0;  // cells of all scenes
    std::vector<int> foff, soff, coff, woff;
    std::vector<int> h_fscene, h_sscene;  // scene of each caller index
    int* sscene = nullptr;                // device, solid -> scene
    int* dwoff = nullptr;                 // device, walls of each scene
    int nwmax = 0;                        // most walls of one scene
    int key_bits = 1;
    double hmax = 0;  // largest fluid smoothing length of all scenes
    double fix_lmax = 0;  // longest extent of the fixed-point map's domain box
    double gv[3] = {0, 0, -9.81};
    // fluid
    FluidSet fset[2];
    int dcur = 0;  // current dRho / pair-record buffer (Jacobi ping-pong)
    int* fslot = nullptr;
    int* fmember = nullptr;
    unsigned* fkey = nullptr;         // reference hash per caller index
    unsigned* frank = nullptr;        // sort key: Morton rank of that cell
    unsigned* fkey_sorted = nullptr;  // sorted ranks
    int* fperm = nullptr;             // sorted position -> caller index
    unsigned* crank = nullptr;        // cell -> Morton rank (spare cell last)
    unsigned* cinv = nullptr;         // Morton rank -> cell
    int* fcs = nullptr;
    int* fce = nullptr;
    int2* fchunks = nullptr;    // superset build chunks: (first, end) sorted positions
    int nfchunks = 0;
    unsigned* hslot = nullptr;  // reference hash per fluid slot
    unsigned char* cspare = nullptr;  // per cell: its stencil reaches the spare cell
    int* near_out = nullptr;          // per cell: near an out-of-domain fluid particle (this step)
    // the static solids' grid (wall dummies never move: sorted once, at
    // upload); with a moving dummy every search sorts all solids instead
    // the solids' hashes are set for all of them (the static ones keep
    // theirs); static_out: static solids out of the domain
    bool skey_valid = false;
    unsigned long long* dispmax = nullptr;  // per-workgroup fluid displacement maxima
    int* rank_acc = nullptr;                // active-grid rank accumulators (zero between sorts)
    int* tsum = nullptr;                    // fluid–solid count per tile of 256 slots
    unsigned long long static_out = 0;
    int* dcs = nullptr;
    int* dce = nullptr;
    int* dperm = nullptr;
    bool all_solid_sort = false;
    double4* fxb = nullptr;     // fluid positions at the last superset build
    double4* fsp = nullptr;     // fluid positions at the last fluid search
    double4* sxb = nullptr;     // solid positions at the last superset build
    unsigned long long* ctl = nullptr;
    double skin_h = 0.3;        // superset skin / max fluid h
    int tune_bs = 256;          // workgroup of the fluid pair loops
    int tune_bmem = 64;         // particles per superset build chunk (one wavefront)
    int tune_group = 1;         // particles sharing one superset row
    int tune_tile = 1;          // mixed pair loops over LDS tiles (1) or gathering from HBM (0)
    Tiles T;                    // the tiles while they are on (tiles_on)
    bool fluid_uniform = false; // every fluid particle has the same (h, c, rho0, nu)
    UniformFluid ufluid;
    int sub_bits = 9;           // fluid sort key: sub-cell Morton bits below the cell rank
    // active (non-dummy) solid grid for the DEM contact search
    unsigned* akey = nullptr;
    unsigned* akey_sorted = nullptr;
    int* aperm = nullptr;
    int* acs = nullptr;
    int* ace = nullptr;
    // solids
    Solids so;
    unsigned* skey = nullptr;
    unsigned* skey_sorted = nullptr;
    int* sperm = nullptr;
    int* scs = nullptr;
    int* sce = nullptr;
    Clumps cl;
    double4* wpt = nullptr;
    double4* wn = nullptr;
    int* wmat = nullptr;
    double* whist = nullptr;
    // cylinder walls: (point, radius), (unit axis), (angular velocity)
    double4* ypt = nullptr;
    double4* yax = nullptr;
    double4* yw = nullptr;
    int* ymat = nullptr;
    int ny = 0;
    int* dyoff = nullptr;                 // device, cylinders of each scene
    std::vector<int> yoff;
    int nymax = 0;                        // most cylinders of one scene
    double* yhist = nullptr;
    ContactTables tb{};
    int nmat = 0, ncomb = 0;
    Lists L;
    Bonds B;
    unsigned* radix_tmp = nullptr;
    int* scan_tmp = nullptr;
    unsigned long long* status = nullptr;
    int reorder_period = 1;
    double graph_dt = -1.0;  // time step baked into the captured graphs
    double* sdt = nullptr;   // per-scene time steps of a batch (device), or none
    std::vector<void*> allocs;
    std::map<uint64_t, StepGraph> graphs;
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[F_NFAM];
    size_t ev_used[F_NFAM] = {0};
    // device copy of the state taken before a graph-replayed run, restored
    // when a list overflows during it (one buffer per state array)
    std::vector<void*> backup;
    std::vector<size_t> backup_bytes;
    // work buffer of dsph_hash_sort (grown on demand)
    unsigned* hs_work = nullptr;
    int64_t hs_words = 0;
    // VTU frames (dsph_frame_begin / dsph_frame_write): per slot a staging
    // block on the device and its pinned host copy (kept across uploads),
    // copied on a stream of their own
    struct FrameSlot {
        void* dev = nullptr;   // in allocs
        size_t dev_cap = 0;
        void* host = nullptr;  // pinned (hipHostMalloc)
        size_t cap = 0;        // bytes allocated
        hipEvent_t packed = nullptr, copied = nullptr;
        bool pending = false;
        uint32_t species = 0;
        int32_t step = 0;
        double time = 0.0;
        int64_t nf = 0, ns = 0, pcap = 0;
    } frames[DSPH_FRAME_SLOTS];
    hipStream_t cst = nullptr;  // the frames' copy stream (created at the first frame)

    bool mixed() const { return prec == DSPH_PREC_MIXED; }
};

namespace {

template <class T>
T* dalloc(dsph_ctx* c, size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess)
        throw ApiError{DSPH_E_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed"};
    c->allocs.push_back(p);
    return (T*)p;
}

// frees an allocation of this context (a pointer it does not hold, e.g.
// one already released with the whole scene, is left alone)
void dfree(dsph_ctx* c, void* p) {
    if (!p) return;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
    if (it == c->allocs.end()) return;
    c->allocs.erase(it);
    (void)hipFree(p);
}

template <class T>
void h2d(dsph_ctx* c, T* d, const T* h, size_t n) {
    if (n) HIPCHK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, c->st));
}
template <class T>
void d2h(dsph_ctx* c, T* h, const T* d, size_t n) {
    if (n) HIPCHK(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, c->st));
}

void free_all(dsph_ctx* c) {
    for (auto& kv : c->graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    c->graphs.clear();
    if (c->cst) (void)hipStreamSynchronize(c->cst);  // frame copies read staging blocks
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    c->T = Tiles{};
    c->sdt = nullptr;
    c->backup.clear();
    c->backup_bytes.clear();
    c->hs_work = nullptr;
    c->hs_words = 0;
    for (auto& f : c->frames) {
        f.dev = nullptr;
        f.dev_cap = 0;
        f.pending = false;
    }
}

int fail(dsph_ctx* c, const ApiError& e) {
    if (c) c->err = e.msg;
    return e.code;
}

// --------------------------------------------------------------- timing
struct FamTimer {
    dsph_ctx* c;
    int f;
    hipStream_t s;
    FamTimer(dsph_ctx* c_, int f_, hipStream_t s_ = nullptr) : c(c_), f(f_), s(s_ ? s_ : c_->st) {
        if (!c->timing) return;
        if (c->ev_used[f] == c->ev[f].size()) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            c->ev[f].push_back({a, b});
        }
        HIPCHK(hipEventRecord(c->ev[f][c->ev_used[f]].first, s));
    }
    ~FamTimer() {
        if (!c->timing) return;
        (void)hipEventRecord(c->ev[f][c->ev_used[f]].second, s);
        c->ev_used[f]++;
    }
};

// --------------------------------------------------------------- capacities
void alloc_ff(dsph_ctx* c, int K) {
    dfree(c, c->L.nbr);
    c->L.K = K;
    c->L.nbr = dalloc<int>(c, (size_t)(K + kRowPad) * std::max(c->nf, 1));
}

// ------------------------------------------------------------------ tiles
// The pair-loop tiles run for the mixed precision with the reorder on (a
// tile must be a block of consecutive slots that is compact in space).
// the tiled loops' sentinel record (fixed-point origin) lies at least the
// domain box's longest extent from every particle in it, which must clear
// every support (2 h_max) with room to spare
bool tiles_on(const dsph_ctx* c) {
    return c->tune_tile != 0 && c->mixed() && c->reorder_period > 0 && c->nf > 0 &&
           c->fix_lmax >= 4.0 * c->hmax;
}

void free_tiles(dsph_ctx* c) {
    dfree(c, (void*)c->T.range);
    dfree(c, c->T.uni);
    dfree(c, c->T.ucnt);
    dfree(c, c->T.row);
    dfree(c, c->T.rcnt);
    c->T = Tiles{};
}

// local rows of the tiles, (supK + kRowPad) columns like the superset
void alloc_tile_rows(dsph_ctx* c) {
    if (c->T.n == 0) return;
    dfree(c, c->T.row);
    c->T.row = dalloc<unsigned short>(c, (size_t)tile_row_cols(c->L.supK) * std::max(c->nf, 1));
    fill_int(c->st, c->T.ucnt, c->T.n, -1);  // no tile is staged before the next build
}

// Runs of up to kTileSlots consecutive slots of one scene (scene k holds
// the slots [foff[k], foff[k+1])).  No tile is built until the next
// superset build (ucnt = -1: the tile walks its global rows).
void alloc_tiles(dsph_ctx* c) {
    free_tiles(c);
    if (!tiles_on(c)) return;
    std::vector<int2> r;
    for (int k = 0; k < c->nscene; k++)
        for (int t = c->foff[k]; t < c->foff[k + 1]; t += kTileSlots)
            r.push_back(make_int2(t, std::min(t + kTileSlots, c->foff[k + 1])));
    if (r.empty()) return;
    Tiles& T = c->T;
    T.n = (int)r.size();
    int2* range = dalloc<int2>(c, r.size());
    h2d(c, range, r.data(), r.size());
    T.range = range;
    T.uni = dalloc<int>(c, (size_t)T.n * kTileCap);
    T.ucnt = dalloc<int>(c, T.n);
    T.rcnt = dalloc<int>(c, c->nf);
    fill_int(c->st, T.rcnt, c->nf, 0);
    alloc_tile_rows(c);
    HIPCHK(hipStreamSynchronize(c->st));
}

void alloc_sup(dsph_ctx* c, int K, int KS) {
    dfree(c, c->L.sup);
    dfree(c, c->L.fsup);
    c->L.supK = K;
    c->L.fsupK = KS;
    c->L.sup = dalloc<int>(c, (size_t)(K + kRowPad) * std::max(c->nf, 1));
    c->L.fsup = dalloc<int>(c, (size_t)KS * std::max(c->nf, 1));
    alloc_tile_rows(c);
}

// more superset capacity with the current rows kept: the ELL is
// column-major with stride nf, so the first supK columns copy as one block
// (for the tiles' local rows also the pad columns, which hold positions
// the pair loops read ahead)
void grow_sup_keep(dsph_ctx* c, int K) {
    if (K <= c->L.supK) return;
    const size_t nf = (size_t)std::max(c->nf, 1);
    int* fresh = dalloc<int>(c, (size_t)(K + kRowPad) * nf);
    HIPCHK(hipMemcpyAsync(fresh, c->L.sup, sizeof(int) * (size_t)c->L.supK * nf, hipMemcpyDeviceToDevice,
                          c->st));
    if (c->T.n > 0) {
        // the quads of the old capacity are a prefix of the new array
        unsigned short* rows = dalloc<unsigned short>(c, (size_t)tile_row_cols(K) * nf);
        HIPCHK(hipMemcpyAsync(rows, c->T.row, sizeof(unsigned short) * (size_t)tile_row_cols(c->L.supK) * nf,
                              hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        dfree(c, c->T.row);
        c->T.row = rows;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(c, c->L.sup);
    c->L.sup = fresh;
    c->L.supK = K;
}

void alloc_fs(dsph_ctx* c, int64_t cap) {
    dfree(c, c->L.fs_a);
    dfree(c, c->L.fs_own);
    dfree(c, c->L.fsF);
    dfree(c, c->L.fsT);
    dfree(c, c->L.fsr_idx);
    c->L.fs_cap = cap;
    c->L.fs_a = dalloc<int>(c, cap);
    c->L.fs_own = dalloc<int>(c, cap);
    c->L.fsF = dalloc<double4>(c, cap);
    c->L.fsT = dalloc<double4>(c, cap);
    c->L.fsr_idx = dalloc<int>(c, cap);
    // owners stay in range until a list is written
    HIPCHK(hipMemsetAsync(c->L.fs_own, 0, (size_t)cap * sizeof(int), c->st));
}

// grows the solid–solid arrays, keeping the saved history (offsets are
// separate arrays and unaffected)
void alloc_ss(dsph_ctx* c, int64_t cap, bool keep_history) {
    int* old_hj = c->L.h_j;
    double* old_hS = c->L.hS;
    int64_t old_cap = c->L.ss_cap;
    for (void* p : {(void*)c->L.ss_i, (void*)c->L.ss_j, (void*)c->L.ssF, (void*)c->L.ssT,
                    (void*)c->L.ssS, (void*)c->L.ssr_idx})
        dfree(c, p);
    c->L.ss_cap = cap;
    c->L.ss_i = dalloc<int>(c, cap);
    c->L.ss_j = dalloc<int>(c, cap);
    c->L.ssF = dalloc<double4>(c, cap);
    c->L.ssT = dalloc<double4>(c, cap);
    c->L.ssS = dalloc<double>(c, 9 * cap);
    c->L.ssr_idx = dalloc<int>(c, cap);
    c->L.h_j = dalloc<int>(c, cap);
    c->L.hS = dalloc<double>(c, 9 * cap);
    if (keep_history && old_hj) {
        int64_t n = std::min(old_cap, cap);
        HIPCHK(hipMemcpyAsync(c->L.h_j, old_hj, n * sizeof(int), hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->L.hS, old_hS, 9 * n * sizeof(double), hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    dfree(c, old_hj);
    dfree(c, old_hS);
}

void invalidate_graphs(dsph_ctx* c) {
    for (auto& kv : c->graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    c->graphs.clear();
}

// the bond arrays for nb bonds (replacing any earlier ones)
void alloc_bonds(dsph_ctx* c, int nb) {
    Bonds& B = c->B;
    for (void* p : {(void*)B.i, (void*)B.j, (void*)B.on, (void*)B.nrm, (void*)B.pt, (void*)B.Fs, (void*)B.Tb,
                    (void*)B.Fn, (void*)B.Tt})
        dfree(c, p);
    B = Bonds();
    B.n = nb;
    B.i = dalloc<int>(c, nb);
    B.j = dalloc<int>(c, nb);
    B.on = dalloc<int>(c, nb);
    B.nrm = dalloc<double4>(c, nb);
    B.pt = dalloc<double4>(c, nb);
    B.Fs = dalloc<double4>(c, nb);
    B.Tb = dalloc<double4>(c, nb);
    B.Fn = dalloc<double>(c, nb);
    B.Tt = dalloc<double>(c, nb);
}

// ------------------------------------------------------------------ state
// Every device array a later step (or a read-back) reads from an earlier
// one, in a fixed order: the fluid in its slot order with both record
// buffers, the hashed grids and sort results, the Verlet supersets with
// their build positions and control words, every list (the fluid-solid
// list and its forces carry over the DEM-only steps of a sub-cycled run,
// the contact list carries the spring history), solids, clumps, wall
// springs, bonds, the contact tables and the status words.  Scratch whose
// contents no step inherits (the reorder copy, sort and scan buffers,
// per-workgroup maxima) and arrays fixed at upload are left out.
struct StateArr {
    void* p;
    size_t bytes;
};
std::vector<StateArr> state_arrays(dsph_ctx* c) {
    std::vector<StateArr> v;
    auto add = [&](const void* p, size_t n, size_t elem) {
        if (p && n) v.push_back({const_cast<void*>(p), n * elem});
    };
    const size_t nf = c->nf, ns = c->ns, nc = c->nc, nb = c->B.n, ncell = c->ncell;
    const size_t na = c->so.nactive;
    const FluidSet& f = c->fset[0];
    const size_t rec = 2 * (c->mixed() ? sizeof(float4) : sizeof(double4));
    add(f.ph, nf, 32);
    add(f.vr, nf, 32);
    add(f.drho[0], nf, 8);
    add(f.drho[1], nf, 8);
    add(f.cn, nf, 16);
    add(f.alpha, nf, 8);
    add(f.sumw0, nf, 8);
    add(f.acc, nf, 32);
    add(f.orig, nf, 4);
    add(f.scene, nf, 4);
    add(f.tid, nf, 4);
    add(f.rec[0], nf, rec);
    add(f.rec[1], nf, rec);
    for (const void* p : {(const void*)c->fslot, (const void*)c->fmember, (const void*)c->fkey,
                          (const void*)c->frank, (const void*)c->fkey_sorted, (const void*)c->fperm,
                          (const void*)c->hslot})
        add(p, nf, 4);
    add(c->fcs, ncell, 4);
    add(c->fce, ncell, 4);
    add(c->near_out, ncell, 4);
    add(c->fxb, nf, 32);
    add(c->fsp, nf, 32);
    add(c->sxb, ns, 32);
    add(c->ctl, CTL_WORDS, 8);
    add(c->status, ST_WORDS, 8);
    add(c->akey, na, 4);
    add(c->akey_sorted, na, 4);
    add(c->aperm, na, 4);
    add(c->acs, ncell, 4);
    add(c->ace, ncell, 4);
    const Solids& so = c->so;
    for (const void* p : {(const void*)so.pr, (const void*)so.vel, (const void*)so.acc, (const void*)so.w,
                          (const void*)so.tq, (const void*)so.nrm})
        add(p, ns, 32);
    add(c->skey, ns, 4);
    add(c->skey_sorted, ns, 4);
    add(c->sperm, ns, 4);
    add(c->scs, ncell, 4);
    add(c->sce, ncell, 4);
    const Clumps& cl = c->cl;
    for (const void* p : {(const void*)cl.pos, (const void*)cl.vel, (const void*)cl.acc, (const void*)cl.w,
                          (const void*)cl.tq, (const void*)cl.q})
        add(p, nc, 32);
    add(c->whist, 9 * ns * (size_t)c->nwmax, 8);
    add(c->yhist, 9 * ns * (size_t)c->nymax, 8);
    add(c->tb.hertz, 8 * (size_t)c->ncomb, 8);
    add(c->tb.linear, 11 * (size_t)c->ncomb, 8);
    add(c->tb.bonded, 6 * (size_t)c->ncomb, 8);
    const Lists& L = c->L;
    add(L.cnt, nf, 4);
    add(L.nbr, (size_t)(L.K + kRowPad) * nf, 4);
    add(L.fscnt, nf, 4);
    add(L.fsoff, nf + 1, 4);
    add(L.fs_a, (size_t)L.fs_cap, 4);
    add(L.fs_own, (size_t)L.fs_cap, 4);
    add(L.fsF, (size_t)L.fs_cap, 32);
    add(L.fsr_idx, (size_t)L.fs_cap, 4);
    add(L.fsr_cnt, ns, 4);
    add(L.fsr_off, ns + 1, 4);
    add(L.sup, (size_t)(L.supK + kRowPad) * nf, 4);
    add(L.sup_cnt, nf, 4);
    add(L.grow, nf, 4);
    add(L.fsup, (size_t)L.fsupK * nf, 4);
    add(L.fsup_cnt, nf, 4);
    if (c->T.n > 0) {
        add(c->T.uni, (size_t)c->T.n * kTileCap, 4);
        add(c->T.ucnt, (size_t)c->T.n, 4);
        add(c->T.row, (size_t)tile_row_cols(L.supK) * nf, 2);
        add(c->T.rcnt, nf, 4);
    }
    add(L.fsm, 2 * nf, 8);
    add(L.ss_i, (size_t)L.ss_cap, 4);
    add(L.ss_j, (size_t)L.ss_cap, 4);
    add(L.ssF, (size_t)L.ss_cap, 32);
    add(L.ssT, (size_t)L.ss_cap, 32);
    add(L.ssS, 9 * (size_t)L.ss_cap, 8);
    add(L.ssr_idx, (size_t)L.ss_cap, 4);
    add(L.sscnt, ns, 4);
    add(L.ssoff, ns + 1, 4);
    add(L.hoff, ns + 1, 4);
    add(L.hcnt, ns, 4);
    add(L.h_j, (size_t)L.ss_cap, 4);
    add(L.hS, 9 * (size_t)L.ss_cap, 8);
    add(L.ssr_cnt, ns, 4);
    add(L.ssr_off, ns + 1, 4);
    const Bonds& B = c->B;
    add(B.i, nb, 4);
    add(B.j, nb, 4);
    add(B.on, nb, 4);
    add(B.nrm, nb, 32);
    add(B.pt, nb, 32);
    add(B.Fs, nb, 32);
    add(B.Tb, nb, 32);
    add(B.Fn, nb, 8);
    add(B.Tt, nb, 8);
    return v;
}

// host-side state that goes with the device arrays of a snapshot
struct HostState {
    int dcur;
    int skey_valid;
};

void free_backup(dsph_ctx* c) {
    for (void* p : c->backup) dfree(c, p);
    c->backup.clear();
    c->backup_bytes.clear();
}

// device copy of the state (buffers reallocated when a capacity changed)
HostState snapshot_device(dsph_ctx* c) {
    const std::vector<StateArr> v = state_arrays(c);
    bool same = v.size() == c->backup.size();
    for (size_t k = 0; same && k < v.size(); k++) same = v[k].bytes == c->backup_bytes[k];
    if (!same) {
        HIPCHK(hipStreamSynchronize(c->st));
        free_backup(c);
        for (const StateArr& a : v) {
            c->backup.push_back(dalloc<char>(c, a.bytes));
            c->backup_bytes.push_back(a.bytes);
        }
    }
    for (size_t k = 0; k < v.size(); k++)
        HIPCHK(hipMemcpyAsync(c->backup[k], v[k].p, v[k].bytes, hipMemcpyDeviceToDevice, c->st));
    return HostState{c->dcur, (int)c->skey_valid};
}

void restore_device(dsph_ctx* c, const HostState& h) {
    const std::vector<StateArr> v = state_arrays(c);
    if (v.size() != c->backup.size()) throw ApiError{DSPH_E_STATE, "internal error: state snapshot out of date"};
    for (size_t k = 0; k < v.size(); k++) {
        if (v[k].bytes != c->backup_bytes[k])
            throw ApiError{DSPH_E_STATE, "internal error: state snapshot out of date"};
        HIPCHK(hipMemcpyAsync(v[k].p, c->backup[k], v[k].bytes, hipMemcpyDeviceToDevice, c->st));
    }
    c->dcur = h.dcur;
    c->skey_valid = h.skey_valid != 0;
}

// ----------------------------------------------------------------- phases
FluidArgs fluid_args(dsph_ctx* c) {
    FluidArgs a;
    a.f = c->fset[0];
    a.uni = c->fluid_uniform;
    a.u = c->ufluid;
    a.nf = c->nf;
    a.bs = c->tune_bs;
    a.cur = c->dcur;
    a.L = c->L;
    a.so = c->so;
    a.gx = c->gv[0];
    a.gy = c->gv[1];
    a.gz = c->gv[2];
    a.flags = c->flags;
    a.T = c->T;
    a.sdt = c->sdt;
    a.ns = c->ns;
    return a;
}

DemArgs dem_args(dsph_ctx* c) {
    DemArgs a;
    a.so = c->so;
    a.ns = c->ns;
    a.c = c->cl;
    a.nc = c->nc;
    a.L = c->L;
    a.b = c->B;
    a.wpt = c->wpt;
    a.wn = c->wn;
    a.wmat = c->wmat;
    a.nw = c->nw;
    a.woff = c->dwoff;
    a.sscene = c->sscene;
    a.nscene = c->nscene;
    a.nwmax = c->nwmax;
    a.whist = c->whist;
    a.ypt = c->ypt;
    a.yax = c->yax;
    a.yw = c->yw;
    a.ymat = c->ymat;
    a.ny = c->ny;
    a.yoff = c->dyoff;
    a.nymax = c->nymax;
    a.yhist = c->yhist;
    a.tb = c->tb;
    a.gx = c->gv[0];
    a.gy = c->gv[1];
    a.gz = c->gv[2];
    a.flags = c->flags;
    a.sdt = c->sdt;
    return a;
}

// Chunks of the superset build: runs of <= tune_bmem consecutive sorted
// particles of one scene.  Scene k's cells rank in [cell0_k, cell0_k +
// ncell_k), so its particles hold the sorted positions [foff[k], foff[k+1]).
void build_chunks(dsph_ctx* c) {
    std::vector<int2> ch;
    const int m = c->tune_bmem;
    for (int k = 0; k < c->nscene; k++)
        for (int t = c->foff[k]; t < c->foff[k + 1]; t += m)
            ch.push_back(make_int2(t, std::min(t + m, c->foff[k + 1])));
    if (c->fchunks) {
        HIPCHK(hipStreamSynchronize(c->st));
        dfree(c, c->fchunks);
    }
    c->fchunks = dalloc<int2>(c, std::max<size_t>(ch.size(), 1));
    h2d(c, c->fchunks, ch.data(), ch.size());
    HIPCHK(hipStreamSynchronize(c->st));
    c->nfchunks = (int)ch.size();
}

// Half-width of the superset build stencil: every particle within the
// superset radius 2 h_max + skin of a particle lies within sw cells of that
// particle's cell on each axis.  The (2 sw + 1)^3 linear offsets are distinct
// while gx, gy >= 2 sw + 1; narrower grids (or very large skins) walk the
// reference stencil for every row instead.
void superset_stencil(dsph_ctx* c) {
    for (GridP& g : c->grids) {
        g.skin = c->skin_h * g.hmax;
        g.iskin2 = g.skin > 0.0 ? 1.0 / (g.skin * g.skin) : 1e300;
        const double R = (2.0 + c->skin_h) * g.hmax * (1.0 + 1e-6);
        const double cmin = std::min({g.csx, g.csy, g.csz});
        int w = cmin > 0 ? (int)std::ceil(R / cmin) : 1;
        w = std::max(w, 1);
        g.sw = std::min(w, 8);
        g.walk_only = w > 8 || g.gx < 2 * g.sw + 1 || g.gy < 2 * g.sw + 1;
    }
    if (c->dgrids) h2d(c, c->dgrids, c->grids.data(), c->grids.size());
}

// the grids as the kernels see them
Grids grids_view(dsph_ctx* c) {
    Grids G;
    G.g = c->dgrids;
    G.fscene = c->fset[0].scene;
    G.sscene = c->sscene;
    G.nscene = c->nscene;
    G.g0 = c->grids[0];
    return G;
}

SearchArgs search_args(dsph_ctx* c) {
    SearchArgs a;
    a.G = grids_view(c);
    a.ncell = c->ncell;
    a.group = c->tune_group;
    a.nf = c->nf;
    a.ns = c->ns;
    a.ph = c->fset[0].ph;
    a.fcs = c->fcs;
    a.fce = c->fce;
    a.fmember = c->fmember;
    a.forig = c->fset[0].orig;
    a.hslot = c->hslot;
    a.cspare = c->cspare;
    a.near_out = c->near_out;
    if (c->all_solid_sort) {
        a.scs = c->scs;
        a.sce = c->sce;
        a.sperm = c->sperm;
        a.acs = a.ace = a.aperm = a.active = nullptr;
    } else {
        a.scs = c->dcs;
        a.sce = c->dce;
        a.sperm = c->dperm;
        a.acs = c->acs;
        a.ace = c->ace;
        a.aperm = c->aperm;
        a.active = c->so.active;
    }
    a.skey = c->skey;
    a.spr = c->so.pr;
    a.fsp = c->fsp;
    a.fxb = c->fxb;
    a.sxb = c->sxb;
    a.slive = c->so.live;
    a.nslive = c->so.nlive;
    a.tsum = c->tsum;
    a.L = c->L;
    a.ctl = c->ctl;
    a.st = c->status;
    a.T = c->T;
    return a;
}

// neighborSearch (neighborSearch.cu:314-381): the reference hashes of both
// species every step; on fluid steps the Verlet superset (rebuilt when
// needed, with the sorted-order reorder) and the exact lists; DEM contacts
// every step.  The cell-sorted grids are only read by a superset build and
// by exact stencil walks, so the sorts run on the steps that have one (a
// device-side decision, CTL_SORT); the reference grid of any step is
// reconstructed from its hashes on read-back.
// zero_fs: the new fluid-solid pairs start with zero force, as the
// reference's fresh list does; a full step skips it (its fluid phase writes
// every pair's force before anything reads them)
void neighbor_search(dsph_ctx* c, int i_step, int gap, bool save_history, bool zero_fs = true) {
    hipStream_t s = c->st;
    long long* stw = (long long*)c->status;
    const bool fluid_step = i_step % gap == 0 && c->nf > 0;
    {
        FamTimer t(c, F_GRID);
        const Grids G = grids_view(c);
        hash_fluid(s, G, c->fset[0].ph, c->fset[0].orig, c->nf, c->sub_bits, c->crank, c->cspare, c->fkey,
                   c->frank, c->hslot, fluid_step ? c->fxb : nullptr, c->dispmax,
                   fluid_step ? c->near_out : nullptr, c->ncell, c->ctl, stw);
        bool walk_only = false;
        for (const GridP& g : c->grids) walk_only = walk_only || g.walk_only;
        // only the live solids move: after one full pass the others' hashes
        // stay valid (static_out of them are out of the domain)
        const bool full = !c->skey_valid;
        if (full) hash_solid(s, G, c->so.pr, c->ns, c->skey, stw);
        hash_live_decide(s, G, c->so.pr, c->so.live, full ? 0 : c->so.nlive, c->skey, full ? 0 : c->static_out,
                         fluid_step, c->sxb, c->dispmax, c->nf, c->skin_h <= 0.0, walk_only, c->ctl, stw);
        c->skey_valid = true;
        if (fluid_step) {
            const Gate gate = c->ctl + CTL_SORT;
            // by cell rank, then sub-cell code, then index: a stable LSD radix
            // sort of the keys (linear in the particle count however full a
            // cell or the spare cell is); the cell ranges come from the
            // sorted keys
            radix_sort_iota(s, c->frank, c->nf, c->key_bits + c->sub_bits, c->fkey_sorted, c->fperm,
                            c->radix_tmp, gate);
            cell_spans(s, c->fkey_sorted, c->nf, c->fcs, c->fce, c->ncell, c->cinv, gate, c->sub_bits);
            if (c->all_solid_sort) {
                if (c->ns > 0)
                    radix_sort_iota(s, c->skey, c->ns, c->key_bits, c->skey_sorted, c->sperm, c->radix_tmp,
                                    gate);
                cell_spans(s, c->skey_sorted, c->ns, c->scs, c->sce, c->ncell, nullptr, gate);
            }
            const bool reorder = c->reorder_period > 0;
            if (reorder)
                reorder_fluid(s, c->fset[0], c->fset[1], c->fperm, c->fslot, c->nf, c->mixed(),
                              c->dcur, c->ctl);
            // slots of the sorted positions and the copy-back of a reorder,
            // in one pass
            after_sort(s, c->fset[0], c->fset[1], reorder, c->mixed(), c->dcur, c->fkey_sorted, c->nf,
                       c->sub_bits, c->fperm, c->fslot, c->cinv, c->fmember, c->hslot, c->ctl);
        }
    }
    if (c->ns > 0) {
        // the active solids' grid: the DEM search, and the fluid's view of
        // the moving solids
        FamTimer t(c, F_SSEARCH);
        active_grid(s, c->so, c->skey, c->ncell, c->key_bits, c->akey, c->akey_sorted, c->aperm,
                    c->acs, c->ace, c->radix_tmp, c->rank_acc);
    }
    if (fluid_step) {
        const SearchArgs sa = search_args(c);
        {
            FamTimer t(c, F_FSEARCH);
            fluid_superset(s, sa, c->fchunks, c->nfchunks);
            tile_rows(s, sa);
        }
        FamTimer t(c, F_FSLIST);
        fluid_filter(s, sa, zero_fs);
    }
    if (c->ns > 0) {
        FamTimer t(c, F_SSEARCH);
        if (save_history) solid_save_history(s, c->so, c->L, c->ns);
        solid_search(s, grids_view(c), c->so, c->ns, c->acs, c->ace, c->aperm, c->L, c->scan_tmp, stw);
        // the reverse lists of the contacts and, on fluid steps, of the new
        // fluid–solid pairs (by movable solid), built together
        if (fluid_step)
            reverse_both(s, c->L.fs_a, c->L.fsoff + c->nf, c->L.fs_cap, c->so.vel, c->so, c->L, c->ns,
                         c->scan_tmp);
        else
            reverse_solid(s, c->so, c->L, c->ns, c->scan_tmp);
    }
}

void force_rebuild(dsph_ctx* c) {
    unsigned long long one = 1;
    HIPCHK(hipMemcpyAsync(c->ctl + CTL_FORCE, &one, sizeof(one), hipMemcpyHostToDevice, c->st));
}

void read_status(dsph_ctx* c, unsigned long long* w) {
    HIPCHK(hipMemcpyAsync(w, c->status, ST_WORDS * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
}

// a search that met a NaN/inf position (the hash kernels count them): the
// state cannot recover from it, so the step reports it instead of going on
void check_finite(const unsigned long long* w) {
    if (w[ST_NONFINITE] == 0) return;
    throw ApiError{DSPH_E_NONFINITE, std::to_string(w[ST_NONFINITE]) +
                                         " particle(s) with a NaN or infinite position; reload a "
                                         "saved state or upload the scene again"};
}

// search with the reference's grow-on-demand semantics (setNum realloc,
// deviceStructs.h:483-491, 668-691): on overflow the lists grow and the
// search is repeated without a second reorder or history save
void search_checked(dsph_ctx* c, int i_step, int gap) {
    neighbor_search(c, i_step, gap, true);
    for (int attempt = 0; attempt < 8; attempt++) {
        unsigned long long w[ST_WORDS];
        read_status(c, w);
        check_finite(w);
        const bool ovf = w[ST_FF_OVERFLOW] | w[ST_FS_OVERFLOW] | w[ST_SS_OVERFLOW] | w[ST_SUP_OVERFLOW];
        if (!ovf) {
            // superset rows change only at builds, which graph-replayed steps
            // cannot grow: rows above 3/4 of the capacity get room for half
            // as much again (the rows are kept, no new build)
            if (w[ST_SUP_MAX] * 4 > (unsigned long long)c->L.supK * 3) {
                grow_sup_keep(c, (int)(w[ST_SUP_MAX] * 3 / 2 + 16));
                invalidate_graphs(c);
            }
            return;
        }
        if (w[ST_SUP_OVERFLOW])
            alloc_sup(c, (int)std::max<unsigned long long>(w[ST_SUP_MAX] * 3 / 2 + 16,
                                                           (unsigned long long)c->L.supK * 3 / 2 + 16),
                      c->L.fsupK * 3 / 2 + 16);
        if (w[ST_FF_OVERFLOW])
            alloc_ff(c, (int)std::max<unsigned long long>(w[ST_FF_MAX] + 16, (unsigned long long)c->L.K * 2));
        if (w[ST_FS_OVERFLOW]) alloc_fs(c, (int64_t)(w[ST_FS_TOTAL] * 3 / 2 + 1024));
        if (w[ST_SS_OVERFLOW]) alloc_ss(c, (int64_t)(w[ST_SS_TOTAL] * 3 / 2 + 1024), true);
        unsigned long long zero[3] = {0, 0, 0};
        HIPCHK(hipMemcpyAsync(c->status, zero, sizeof(zero), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->status + ST_SUP_OVERFLOW, zero, sizeof(zero[0]), hipMemcpyHostToDevice, c->st));
        invalidate_graphs(c);
        force_rebuild(c);
        neighbor_search(c, i_step, gap, false);
    }
    throw ApiError{DSPH_E_CAPACITY, "neighbor list capacity could not be satisfied"};
}

void solid_before(dsph_ctx* c, double dt) {
    FamTimer t(c, F_DEM);
    DemArgs a = dem_args(c);
    if (c->nc == 0) {
        dem_solid_kick_drift(c->st, a, 0.5 * dt, true);
        return;
    }
    dem_solid_kick(c->st, a, 0.5 * dt);
    dem_clump_kick(c->st, a, 0.5 * dt);
    dem_clump_drift(c->st, a, 0.5 * dt);
    dem_solid_drift(c->st, a, 0.5 * dt);
}

void solid_after(dsph_ctx* c, double dt) {
    FamTimer t(c, F_DEM);
    DemArgs a = dem_args(c);
    if (c->nc == 0) {
        dem_solid_kick_drift(c->st, a, 0.5 * dt, false);
        return;
    }
    dem_solid_drift(c->st, a, 0.5 * dt);
    dem_clump_drift(c->st, a, 0.5 * dt);
    dem_clump_kick(c->st, a, 0.5 * dt);
    dem_solid_kick(c->st, a, 0.5 * dt);
}

void contacts(dsph_ctx* c, double dt) {
    FamTimer t(c, F_DEM);
    DemArgs a = dem_args(c);
    dem_contacts(c->st, a, dt);
    dem_bonds(c->st, a, dt);
    dem_solid_sum(c->st, a, dt);
    dem_clump_sum(c->st, a);
}

void fluid_integrate(dsph_ctx* c, double dt, int i_step, int gap) {
    if (c->nf <= 0) return;
    const bool mx = c->mixed();
    const double dtf = dt * double(gap);
    if (i_step % (5 * gap) == 0 || i_step == 1) {
        FamTimer t(c, F_REINIT);
        density_reinit(c->st, fluid_args(c), i_step, mx);
    }
    if (i_step % gap == 0) {
        {
            // with the half-step drift that follows it
            FamTimer t(c, F_CONT);
            continuity(c->st, fluid_args(c), 0.5 * dtf, mx, 0.5 * dtf);
        }
        c->dcur ^= 1;
        {
            // with the kick and drift that follow it, into the other buffer
            FamTimer t(c, F_MOM);
            momentum(c->st, fluid_args(c), mx, true, dtf);
        }
        c->dcur ^= 1;
        {
            FamTimer t(c, F_CONT);
            continuity(c->st, fluid_args(c), 0.5 * dtf, mx, 0.0);
        }
        c->dcur ^= 1;
    } else {
        FamTimer t(c, F_MOM);
        fluid_solid_force(c->st, fluid_args(c), mx);
    }
}

void step_body(dsph_ctx* c, int i_step, double dt, int gap) {
    neighbor_search(c, i_step, gap, true, false);
    solid_before(c, dt);
    fluid_integrate(c, dt, i_step, gap);
    contacts(c, dt);
    solid_after(c, dt);
}

uint64_t step_key(dsph_ctx* c, int i_step, int gap) {
    uint64_t k = 0;
    k |= (uint64_t)(i_step % gap == 0);
    k |= (uint64_t)(i_step % (5 * gap) == 0 || i_step == 1) << 1;
    k |= (uint64_t)(i_step == 1) << 2;
    k |= (uint64_t)c->dcur << 5;
    k |= (uint64_t)(uint32_t)gap << 8;
    return k;
}

// ------------------------------------------------------------- validation
void require(bool cond, const std::string& msg) {
    if (!cond) throw ApiError{DSPH_E_INVALID, msg};
}

}  // namespace

namespace {
// calculatePointHash (neighborSearch.cu:19-36) on the host, with the
// device's operations (k_search.hip point_hash): the scene's own cell, or
// its spare cell outside [min, max) or at the rounding edge of max
int host_point_hash(const GridP& g, double x, double y, double z) {
    if (!(g.minx <= x && x < g.maxx && g.miny <= y && y < g.maxy && g.minz <= z && z < g.maxz))
        return g.ncell - 1;
    const int ix = (int)((x - g.minx) / g.csx), iy = (int)((y - g.miny) / g.csy),
              iz = (int)((z - g.minz) / g.csz);
    const int h = iz * g.gy * g.gx + iy * g.gx + ix;
    return (h >= 0 && h < g.ncell - 1) ? h : g.ncell - 1;
}

// Slot order of the fluid: cells in Morton (z-order) order of their (x, y, z)
// indices, members of a cell in ascending index.  A cell's members stay
// contiguous, as the reference's hash sort keeps them, while spatially close
// cells get close ranks, so the neighbours of a run of consecutive particles
// are a compact block that stays in one XCD's L2.  The spare cell ranks last.
uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}
void morton_cell_ranks(const GridP& g, std::vector<unsigned>& rank, std::vector<unsigned>& inv) {
    const int64_t nreal = (int64_t)g.ncell - 1;
    std::vector<std::pair<uint64_t, unsigned>> code(nreal);
    for (int64_t h = 0; h < nreal; h++) {
        const int64_t x = h % g.gx, y = (h / g.gx) % g.gy, z = h / ((int64_t)g.gx * g.gy);
        code[h] = {spread3(x) | spread3(y) << 1 | spread3(z) << 2, (unsigned)h};
    }
    std::sort(code.begin(), code.end());
    rank.assign(g.ncell, 0);
    inv.assign(g.ncell, 0);
    for (int64_t r = 0; r < nreal; r++) {
        rank[code[r].second] = (unsigned)r;
        inv[r] = code[r].second;
    }
    rank[nreal] = (unsigned)nreal;
    inv[nreal] = (unsigned)nreal;
}

// The reference emits each list segment in stencil order (z, y, x over the 27
// cells around the particle's cell, neighborSearch.cu:188-214) with cell
// members in ascending index.  The device keeps list rows in the order of
// the last superset build; read-back re-sorts them by (stencil rank on the
// reference grid, index) from the current reference hashes.
void stencil_hashes(const GridP& g, double x, double y, double z, int H[27]) {
    const int ax = (int)((x - g.minx) / g.csx);
    const int ay = (int)((y - g.miny) / g.csy);
    const int az = (int)((z - g.minz) / g.csz);
    int k = 0;
    for (int zz = -1; zz <= 1; zz++)
        for (int yy = -1; yy <= 1; yy++)
            for (int xx = -1; xx <= 1; xx++) {
                const int h = (az + zz) * g.gy * g.gx + (ay + yy) * g.gx + (ax + xx);
                H[k++] = (h < 0 || h >= g.ncell) ? -1 : g.cell0 + h;  // global cell
            }
}
int stencil_rank(const int H[27], int h) {
    for (int k = 0; k < 27; k++)
        if (H[k] == h) return k;
    return 27;
}

// one read-back row entry: the neighbour's cell, its index, a payload
struct RowEnt {
    int key, idx;
    int64_t id;
    int rank;
};
// Puts a row in the reference's order.  When stencil offsets coincide (grids
// narrower than three cells in x or y) a cell is visited once per offset and
// its members appear once per visit: the row then holds that many copies of
// each entry, emitted one copy per visit.  Entries off the stencil go last.
void order_row(const GridP& g, const int H[27], std::vector<RowEnt>& row) {
    for (auto& e : row) e.rank = stencil_rank(H, e.key);
    auto lt = [](const RowEnt& x, const RowEnt& y) {
        if (x.rank != y.rank) return x.rank < y.rank;
        if (x.idx != y.idx) return x.idx < y.idx;
        return x.id < y.id;
    };
    if (!g.collide) {
        std::sort(row.begin(), row.end(), lt);
        return;
    }
    std::sort(row.begin(), row.end(), [](const RowEnt& x, const RowEnt& y) {
        return x.idx != y.idx ? x.idx < y.idx : x.id < y.id;
    });
    std::vector<RowEnt> out;
    std::vector<char> used(row.size(), 0);
    for (int k = 0; k < 27; k++) {
        if (H[k] < 0) continue;
        int last = -1;
        for (size_t e = 0; e < row.size(); e++)
            if (!used[e] && row[e].key == H[k] && row[e].idx != last) {
                used[e] = 1;
                last = row[e].idx;
                out.push_back(row[e]);
            }
    }
    for (size_t e = 0; e < row.size(); e++)
        if (!used[e]) out.push_back(row[e]);
    row.swap(out);
}

struct FluidHost {
    std::vector<int> orig, slot_of;
    std::vector<double4> ph;
    std::vector<unsigned> fkey;  // reference hash per caller index
};
FluidHost fluid_host(dsph_ctx* c) {
    const int nf = c->nf;
    FluidHost f;
    f.orig.resize(nf);
    f.ph.resize(nf);
    f.fkey.resize(nf);
    d2h(c, f.orig.data(), c->fset[0].orig, nf);
    d2h(c, f.ph.data(), c->fsp, nf);  // stencils of the last search
    d2h(c, f.fkey.data(), c->fkey, nf);
    HIPCHK(hipStreamSynchronize(c->st));
    f.slot_of.resize(nf);
    for (int s2 = 0; s2 < nf; s2++) f.slot_of[f.orig[s2]] = s2;
    return f;
}

// fluid–solid pairs in reference order (CSR by fluid caller index) as pair
// ids into the device arrays
std::vector<int64_t> fs_reference_order(dsph_ctx* c, std::vector<int>& fluid_idx,
                                        std::vector<int>& solid_idx, std::vector<int>* count) {
    const int nf = c->nf;
    const FluidHost fh = fluid_host(c);
    std::vector<int> off(nf + 1);
    d2h(c, off.data(), c->L.fsoff, nf + 1);
    HIPCHK(hipStreamSynchronize(c->st));
    int64_t tot = off[nf];
    if (tot > c->L.fs_cap) throw ApiError{DSPH_E_CAPACITY, "fluid-solid list overflowed"};
    // the offsets must rise from 0 to the total and the partners be solids
    bool ok = tot >= 0 && off[0] == 0;
    for (int s2 = 0; ok && s2 < nf; s2++) ok = off[s2] <= off[s2 + 1];
    if (!ok) throw ApiError{DSPH_E_HIP, "internal error: inconsistent fluid-solid list"};
    std::vector<int> a(tot);
    std::vector<unsigned> skey(c->ns);
    d2h(c, a.data(), c->L.fs_a, tot);
    d2h(c, skey.data(), c->skey, c->ns);
    HIPCHK(hipStreamSynchronize(c->st));
    for (int64_t k = 0; k < tot; k++)
        if (a[k] < 0 || a[k] >= c->ns) throw ApiError{DSPH_E_HIP, "internal error: inconsistent fluid-solid list"};
    std::vector<int64_t> ids;
    ids.reserve(tot);
    fluid_idx.clear();
    solid_idx.clear();
    if (count) count->assign(nf, 0);
    std::vector<RowEnt> row;
    for (int i = 0; i < nf; i++) {
        const int s2 = fh.slot_of[i];
        int H[27];
        const GridP& g = c->grids[c->h_fscene[i]];
        stencil_hashes(g, fh.ph[s2].x, fh.ph[s2].y, fh.ph[s2].z, H);
        row.clear();
        for (int k = off[s2]; k < off[s2 + 1]; k++)
            row.push_back({(int)skey[a[k]], a[k], (int64_t)k, 0});
        order_row(g, H, row);
        for (auto& e : row) {
            ids.push_back(e.id);
            fluid_idx.push_back(i);
            solid_idx.push_back(e.idx);
        }
        if (count) (*count)[i] = (int)row.size();
    }
    return ids;
}
}  // namespace

// ==========================================================================
namespace {
// ---------------------------------------------------------------- batches
// Several independent scenes in one context: their arrays are concatenated in
// the order given (so caller indices of scene k follow those of scene k-1),
// clump and cluster ids and pebble ranges are shifted to stay distinct, and
// each scene keeps its own hashed grid (setSpatialGrids, dataBase.cpp:415-432)
// at its own range of global cell indices.
struct Merged {
    std::vector<double> fpos, fvel, fh, frho0, fdrho, fc, fnu, falpha;
    std::vector<double> spos, svel, sw, snrm, srad, seff, sinvm;
    std::vector<int> smat, sclus, sclump, sdum;
    std::vector<double> cpos, cvel, cw, cq, cinvI, cinvm;
    std::vector<int> cstart, cend;
    std::vector<double> wpt, wn;
    std::vector<int> wmat;
    std::vector<double> ypt, yax, yw, yrad;
    std::vector<int> ymat, yoff;
    std::vector<int> fscene, sscene;           // caller order
    std::vector<int> foff, soff, coff, woff;   // nscene + 1 each
    std::vector<GridP> grids;
    bool any_normal = false;
    dsph_scene flat{};
};

template <class T>
void put(std::vector<T>& v, const T* src, size_t n, size_t width, T dflt) {
    if (src)
        v.insert(v.end(), src, src + n * width);
    else
        v.insert(v.end(), n * width, dflt);
}

void validate_scene(const dsph_scene* s) {
    require(s->n_fluid >= 0 && s->n_solid >= 0 && s->n_clump >= 0 && s->n_wall >= 0,
            "negative particle count");
    require(s->n_fluid == 0 || (s->fluid_pos && s->fluid_h && s->fluid_rho0 && s->fluid_c && s->fluid_nu),
            "fluid arrays missing");
    require(s->n_solid == 0 || (s->solid_pos && s->solid_radius && s->solid_eff_radius && s->solid_inv_mass),
            "solid arrays missing");
    require(s->n_clump == 0 || (s->clump_centroid && s->clump_inv_inertia && s->clump_inv_mass &&
                                s->clump_pebble_start && s->clump_pebble_end),
            "clump arrays missing");
    require(s->n_wall == 0 || (s->wall_point && s->wall_normal), "wall arrays missing");
    require(s->n_cylinder >= 0 && (s->n_cylinder == 0 || (s->cylinder_point && s->cylinder_axis &&
                                                          s->cylinder_radius)),
            "cylinder wall arrays missing");
    for (int k = 0; k < s->n_cylinder; k++) {
        const double* ax = s->cylinder_axis + 3 * k;
        require(s->cylinder_radius[k] > 0.0 && ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2] > 0.0,
                "cylinder wall needs a positive radius and a non-zero axis");
        const int m = s->cylinder_material ? s->cylinder_material[k] : 0;
        require(m >= 0 && m < s->n_material, "cylinder wall material outside [0, n_material)");
    }
    require(s->n_material >= 1, "n_material must be >= 1");
    for (int k = 0; k < 3; k++) require(s->domain_size[k] > 0, "domain size must be positive");
    const int nmat = s->n_material;
    for (int i = 0; i < s->n_solid; i++) {
        int m = s->solid_material ? s->solid_material[i] : 0;
        int d = s->solid_dummy ? s->solid_dummy[i] : 0;
        // an invalid material pair would index past the contact tables (the
        // reference reads out of bounds there, deviceStructs.h:960-963)
        require(d == 1 || (m >= 0 && m < nmat),
                "solid " + std::to_string(i) + " has material outside [0, n_material)");
    }
    for (int i = 0; i < s->n_wall; i++) {
        int m = s->wall_material ? s->wall_material[i] : 0;
        require(m >= 0 && m < nmat, "wall material outside [0, n_material)");
    }
    for (int k = 0; k < s->n_clump; k++)
        require(s->clump_pebble_start[k] >= 0 && s->clump_pebble_end[k] <= s->n_solid &&
                    s->clump_pebble_start[k] <= s->clump_pebble_end[k],
                "clump pebble range out of bounds");
    for (int i = 0; i < s->n_fluid; i++)
        require(s->fluid_h[i] > 0 && std::isfinite(s->fluid_h[i]), "fluid smoothing length must be finite and > 0");
    // positions and sizes build the grid: NaN or inf there is rejected here
    // (a position that turns non-finite while stepping is reported by the
    // step instead, DSPH_E_NONFINITE)
    auto finite3 = [](const double* p, int n) {
        for (int64_t i = 0; i < 3 * (int64_t)n; i++)
            if (!std::isfinite(p[i])) return false;
        return true;
    };
    require(s->n_fluid == 0 || finite3(s->fluid_pos, s->n_fluid), "a fluid position is not finite");
    require(s->n_solid == 0 || finite3(s->solid_pos, s->n_solid), "a solid position is not finite");
    for (int i = 0; i < s->n_solid; i++)
        require(std::isfinite(s->solid_radius[i]) && std::isfinite(s->solid_eff_radius[i]),
                "a solid radius is not finite");
}

// the grid of one scene: cell = 2 max(all effective radii), G = int(L / cell)
// (at least 1), cell size L / G per axis, plus the spare cell
GridP scene_grid(const dsph_scene* s, int cell0) {
    double mf = 0, ms = 0;
    for (int i = 0; i < s->n_fluid; i++) mf = std::max(mf, s->fluid_h[i]);
    for (int i = 0; i < s->n_solid; i++) ms = std::max(ms, s->solid_eff_radius[i]);
    const double cs = 2 * std::max(mf, ms);
    GridP g{};
    g.minx = s->domain_origin[0];
    g.miny = s->domain_origin[1];
    g.minz = s->domain_origin[2];
    g.maxx = s->domain_origin[0] + s->domain_size[0];
    g.maxy = s->domain_origin[1] + s->domain_size[1];
    g.maxz = s->domain_origin[2] + s->domain_size[2];
    g.gx = s->domain_size[0] > cs ? int(s->domain_size[0] / cs) : 1;
    g.gy = s->domain_size[1] > cs ? int(s->domain_size[1] / cs) : 1;
    g.gz = s->domain_size[2] > cs ? int(s->domain_size[2] / cs) : 1;
    g.csx = s->domain_size[0] / double(g.gx);
    g.csy = s->domain_size[1] / double(g.gy);
    g.csz = s->domain_size[2] / double(g.gz);
    const long long ncell = (long long)g.gx * g.gy * g.gz + 1;
    require(ncell < (1LL << 30), "grid has too many cells");
    g.ncell = (int)ncell;
    g.cell0 = cell0;
    g.collide = g.gx < 3 || g.gy < 3;
    const double ext = std::max({std::fabs(g.minx), std::fabs(g.maxx), std::fabs(g.miny), std::fabs(g.maxy),
                                 std::fabs(g.minz), std::fabs(g.maxz)});
    g.pad = 1e-9 * (1.0 + ext);
    g.hmax = mf;
    return g;
}

// a scene's contact tables with the reference's defaults where none are given
// (restitution 1, bond multiplier 1, kn/ks 2; hostStructs.h:290-346)
std::vector<double> contact_tables(const dsph_scene* s) {
    const size_t ncb = (size_t)s->n_material * (s->n_material + 1) / 2;
    std::vector<double> t(25 * ncb, 0.0);
    double* hz = t.data();
    double* li = hz + 8 * ncb;
    double* bo = li + 11 * ncb;
    for (size_t k = 0; k < ncb; k++) {
        hz[2 * ncb + k] = 1.0;
        bo[k] = 1.0;
        bo[2 * ncb + k] = 2.0;
    }
    if (s->hertz) std::copy(s->hertz, s->hertz + 8 * ncb, hz);
    if (s->linear) std::copy(s->linear, s->linear + 11 * ncb, li);
    if (s->bonded) std::copy(s->bonded, s->bonded + 6 * ncb, bo);
    return t;
}

void merge_scenes(const dsph_scene* sc, int k, Merged& m) {
    require(k >= 1, "a batch needs at least one scene");
    const std::vector<double> tables0 = contact_tables(sc);
    m.foff = {0};
    m.soff = {0};
    m.coff = {0};
    m.woff = {0};
    m.yoff = {0};
    int cell0 = 0, cluster_base = 0;
    for (int b = 0; b < k; b++) {
        const dsph_scene* s = sc + b;
        validate_scene(s);
        require(s->n_material == sc[0].n_material, "scenes of a batch share n_material");
        require(b == 0 || contact_tables(s) == tables0, "scenes of a batch share the contact tables");
        const size_t nf = s->n_fluid, ns = s->n_solid, nc = s->n_clump, nw = s->n_wall;
        put(m.fpos, s->fluid_pos, nf, 3, 0.0);
        put(m.fvel, s->fluid_vel, nf, 3, 0.0);
        put(m.fh, s->fluid_h, nf, 1, 0.0);
        put(m.frho0, s->fluid_rho0, nf, 1, 0.0);
        put(m.fdrho, s->fluid_drho, nf, 1, 0.0);
        put(m.fc, s->fluid_c, nf, 1, 0.0);
        put(m.fnu, s->fluid_nu, nf, 1, 0.0);
        put(m.falpha, s->fluid_alpha, nf, 1, 0.0);
        m.fscene.insert(m.fscene.end(), nf, b);
        const int sbase = m.soff.back();
        put(m.spos, s->solid_pos, ns, 3, 0.0);
        put(m.svel, s->solid_vel, ns, 3, 0.0);
        put(m.sw, s->solid_angvel, ns, 3, 0.0);
        put(m.snrm, s->solid_normal, ns, 3, 0.0);
        m.any_normal = m.any_normal || (s->solid_normal && ns > 0);
        put(m.srad, s->solid_radius, ns, 1, 0.0);
        put(m.seff, s->solid_eff_radius, ns, 1, 0.0);
        put(m.sinvm, s->solid_inv_mass, ns, 1, 0.0);
        put(m.smat, s->solid_material, ns, 1, 0);
        put(m.sdum, s->solid_dummy, ns, 1, 0);
        int cmax = -1;
        for (size_t i = 0; i < ns; i++) {
            const int cl = s->solid_cluster ? s->solid_cluster[i] : -1;
            m.sclus.push_back(cl >= 0 ? cl + cluster_base : -1);
            cmax = std::max(cmax, cl);
            const int cp = s->solid_clump ? s->solid_clump[i] : -1;
            m.sclump.push_back(cp >= 0 ? cp + m.coff.back() : -1);
        }
        cluster_base += cmax + 1;
        m.sscene.insert(m.sscene.end(), ns, b);
        put(m.cpos, s->clump_centroid, nc, 3, 0.0);
        put(m.cvel, s->clump_vel, nc, 3, 0.0);
        put(m.cw, s->clump_angvel, nc, 3, 0.0);
        for (size_t i = 0; i < nc; i++) {
            if (s->clump_orientation)
                m.cq.insert(m.cq.end(), s->clump_orientation + 4 * i, s->clump_orientation + 4 * i + 4);
            else
                m.cq.insert(m.cq.end(), {1.0, 0.0, 0.0, 0.0});
            m.cstart.push_back(s->clump_pebble_start[i] + sbase);
            m.cend.push_back(s->clump_pebble_end[i] + sbase);
        }
        put(m.cinvI, s->clump_inv_inertia, nc, 6, 0.0);
        put(m.cinvm, s->clump_inv_mass, nc, 1, 0.0);
        put(m.wpt, s->wall_point, nw, 3, 0.0);
        put(m.wn, s->wall_normal, nw, 3, 0.0);
        put(m.wmat, s->wall_material, nw, 1, 0);
        const size_t ny = s->n_cylinder;
        put(m.ypt, s->cylinder_point, ny, 3, 0.0);
        put(m.yax, s->cylinder_axis, ny, 3, 0.0);
        put(m.yw, s->cylinder_angvel, ny, 3, 0.0);
        put(m.yrad, s->cylinder_radius, ny, 1, 0.0);
        put(m.ymat, s->cylinder_material, ny, 1, 0);
        m.yoff.push_back(m.yoff.back() + (int)ny);
        m.grids.push_back(scene_grid(s, cell0));
        require((long long)cell0 + m.grids.back().ncell < (1LL << 31) - 1, "batch has too many cells");
        cell0 += m.grids.back().ncell;
        m.foff.push_back(m.foff.back() + (int)nf);
        m.soff.push_back(m.soff.back() + (int)ns);
        m.coff.push_back(m.coff.back() + (int)nc);
        m.woff.push_back(m.woff.back() + (int)nw);
    }
    dsph_scene& f = m.flat;
    f = sc[0];
    f.n_fluid = m.foff.back();
    f.fluid_pos = m.fpos.data();
    f.fluid_vel = m.fvel.data();
    f.fluid_h = m.fh.data();
    f.fluid_rho0 = m.frho0.data();
    f.fluid_drho = m.fdrho.data();
    f.fluid_c = m.fc.data();
    f.fluid_nu = m.fnu.data();
    f.fluid_alpha = m.falpha.data();
    f.n_solid = m.soff.back();
    f.solid_pos = m.spos.data();
    f.solid_vel = m.svel.data();
    f.solid_angvel = m.sw.data();
    f.solid_normal = m.any_normal ? m.snrm.data() : nullptr;
    f.solid_radius = m.srad.data();
    f.solid_eff_radius = m.seff.data();
    f.solid_inv_mass = m.sinvm.data();
    f.solid_material = m.smat.data();
    f.solid_cluster = m.sclus.data();
    f.solid_clump = m.sclump.data();
    f.solid_dummy = m.sdum.data();
    f.n_clump = m.coff.back();
    f.clump_centroid = m.cpos.data();
    f.clump_vel = m.cvel.data();
    f.clump_angvel = m.cw.data();
    f.clump_orientation = m.cq.data();
    f.clump_inv_inertia = m.cinvI.data();
    f.clump_inv_mass = m.cinvm.data();
    f.clump_pebble_start = m.cstart.data();
    f.clump_pebble_end = m.cend.data();
    f.n_wall = m.woff.back();
    f.wall_point = m.wpt.data();
    f.wall_normal = m.wn.data();
    f.wall_material = m.wmat.data();
    f.n_cylinder = m.yoff.back();
    f.cylinder_point = m.ypt.data();
    f.cylinder_axis = m.yax.data();
    f.cylinder_angvel = m.yw.data();
    f.cylinder_radius = m.yrad.data();
    f.cylinder_material = m.ymat.data();
}
}  // namespace

namespace dsph {
int ctx_device(dsph_ctx* c) { return c->dev; }
hipStream_t ctx_stream(dsph_ctx* c) { return c->st; }
void ctx_set_error(dsph_ctx* c, const char* msg) { c->err = msg; }
unsigned* ctx_work(dsph_ctx* c, int64_t words) {
    if (c->hs_words >= words) return c->hs_work;
    try {
        if (c->hs_work) {
            HIPCHK(hipStreamSynchronize(c->st));
            dfree(c, c->hs_work);
            c->hs_work = nullptr;
            c->hs_words = 0;
        }
        c->hs_work = dalloc<unsigned>(c, (size_t)words);
        c->hs_words = words;
        return c->hs_work;
    } catch (const ApiError& e) {
        c->err = e.msg;
        return nullptr;
    }
}
}  // namespace dsph

extern "C" {

const char* dsph_version(void) { return "dsph-mi355x 0.1 (gfx950)"; }

const char* dsph_last_error(const dsph_ctx* c) { return c ? c->err.c_str() : "null context"; }

int dsph_create(int device_index, int precision, dsph_ctx** out) {
    if (!out) return DSPH_E_INVALID;
    *out = nullptr;
    if (precision != DSPH_PREC_FP64 && precision != DSPH_PREC_MIXED) return DSPH_E_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DSPH_E_NODEVICE;
    if (device_index < 0 || device_index >= n) return DSPH_E_NODEVICE;
    dsph_ctx* c = new dsph_ctx();
    c->dev = device_index;
    c->prec = precision;
    if (hipSetDevice(device_index) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess) {
        delete c;
        return DSPH_E_HIP;
    }
    // the address of the allocation-free entry points' scratch on this
    // device (a failure here is reported by those entry points)
    (void)noalloc_scratch_setup(device_index);
    *out = c;
    return DSPH_OK;
}

int dsph_destroy(dsph_ctx* c) {
    if (!c) return DSPH_OK;
    (void)hipSetDevice(c->dev);
    (void)hipStreamSynchronize(c->st);
    free_all(c);
    for (int f = 0; f < F_NFAM; f++)
        for (auto& e : c->ev[f]) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    for (auto& f : c->frames) {
        if (f.host) (void)hipHostFree(f.host);
        if (f.packed) (void)hipEventDestroy(f.packed);
        if (f.copied) (void)hipEventDestroy(f.copied);
    }
    if (c->cst) (void)hipStreamDestroy(c->cst);
    (void)hipStreamDestroy(c->st);
    delete c;
    return DSPH_OK;
}

int dsph_set_flags(dsph_ctx* c, uint32_t flags) {
    if (!c) return DSPH_E_INVALID;
    c->flags = flags;
    invalidate_graphs(c);
    return DSPH_OK;
}

int dsph_set_reorder_period(dsph_ctx* c, int32_t p) {
    if (!c || p < 0) return DSPH_E_INVALID;
    c->reorder_period = p;
    invalidate_graphs(c);
    if (c->uploaded && (c->T.n > 0) != tiles_on(c)) {
        try {
            HIPCHK(hipStreamSynchronize(c->st));
            alloc_tiles(c);
            force_rebuild(c);
        } catch (const ApiError& e) {
            return fail(c, e);
        }
    }
    return DSPH_OK;
}

int dsph_set_skin(dsph_ctx* c, double skin_over_h) {
    if (!c || !(skin_over_h >= 0.0)) return DSPH_E_INVALID;
    c->skin_h = skin_over_h;
    if (c->uploaded) superset_stencil(c);
    invalidate_graphs(c);
    if (c->uploaded) {
        try {
            force_rebuild(c);
        } catch (const ApiError& e) {
            return fail(c, e);
        }
    }
    return DSPH_OK;
}

int dsph_set_superset_capacity(dsph_ctx* c, int32_t ff, int32_t fs) {
    if (!c || !c->uploaded) return DSPH_E_STATE;
    if (ff < 0 || fs < 0) return DSPH_E_INVALID;
    try {
        HIPCHK(hipStreamSynchronize(c->st));
        alloc_sup(c, ff > 0 ? ff : c->L.supK, fs > 0 ? fs : c->L.fsupK);
        invalidate_graphs(c);
        force_rebuild(c);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_set_tuning(dsph_ctx* c, int32_t key, int32_t value) {
    if (!c) return DSPH_E_INVALID;
    if (key == DSPH_TUNE_PAIR_BLOCK) {
        if (value < 64 || value > 512 || value % 64) return DSPH_E_INVALID;
        c->tune_bs = value;
    } else if (key == DSPH_TUNE_GROUP) {
        if (value != 1 && value != 2 && value != 4 && value != 8) return DSPH_E_INVALID;
        c->tune_group = value;
        if (c->uploaded) {
            try {
                force_rebuild(c);
            } catch (const ApiError& e) {
                return fail(c, e);
            }
        }
    } else if (key == DSPH_TUNE_BUILD_BLOCK) {
        if (value < 1 || value > 64) return DSPH_E_INVALID;
        c->tune_bmem = value;
        if (c->uploaded) {
            try {
                build_chunks(c);
                force_rebuild(c);
            } catch (const ApiError& e) {
                return fail(c, e);
            }
        }
    } else if (key == DSPH_TUNE_TILE) {
        if (value != 0 && value != 1) return DSPH_E_INVALID;
        c->tune_tile = value;
        if (c->uploaded) {
            try {
                HIPCHK(hipStreamSynchronize(c->st));
                alloc_tiles(c);
                force_rebuild(c);
            } catch (const ApiError& e) {
                return fail(c, e);
            }
        }
    } else {
        return DSPH_E_INVALID;
    }
    invalidate_graphs(c);
    return DSPH_OK;
}

int dsph_get_tuning(dsph_ctx* c, int32_t key, int32_t* value) {
    if (!c || !value) return DSPH_E_INVALID;
    if (key == DSPH_TUNE_PAIR_BLOCK)
        *value = c->tune_bs;
    else if (key == DSPH_TUNE_BUILD_BLOCK)
        *value = c->tune_bmem;
    else if (key == DSPH_TUNE_GROUP)
        *value = c->tune_group;
    else if (key == DSPH_TUNE_TILE)
        *value = c->tune_tile;
    else
        return DSPH_E_INVALID;
    return DSPH_OK;
}

int dsph_set_gravity(dsph_ctx* c, const double g[3]) {
    if (!c || !g) return DSPH_E_INVALID;
    c->gv[0] = g[0];
    c->gv[1] = g[1];
    c->gv[2] = g[2];
    invalidate_graphs(c);
    return DSPH_OK;
}

int dsph_upload_scene(dsph_ctx* c, const dsph_scene* s) { return dsph_upload_batch(c, s, 1); }

int dsph_upload_batch(dsph_ctx* c, const dsph_scene* scenes, int32_t n_scenes) {
    if (!c || !scenes || n_scenes < 1) return DSPH_E_INVALID;
    try {
        HIPCHK(hipSetDevice(c->dev));
        Merged M;
        merge_scenes(scenes, n_scenes, M);
        const dsph_scene* s = &M.flat;
        const int nmat = s->n_material;

        invalidate_graphs(c);
        free_all(c);
        c->host.reset();
        c->L = Lists();
        c->B = Bonds();
        c->so = Solids();
        c->cl = Clumps();
        c->fset[0] = FluidSet();
        c->fset[1] = FluidSet();
        c->fchunks = nullptr;
        c->nfchunks = 0;
        c->dcur = 0;
        c->initialized = false;
        const int nf = c->nf = s->n_fluid;
        const int ns = c->ns = s->n_solid;
        const int nc = c->nc = s->n_clump;
        const int nw = c->nw = s->n_wall;
        c->gv[0] = s->gravity[0];
        c->gv[1] = s->gravity[1];
        c->gv[2] = s->gravity[2];

        // ---- grids: one per scene, setSpatialGrids (dataBase.cpp:415-432)
        c->nscene = n_scenes;
        c->grids = M.grids;
        c->ncell = M.grids.back().cell0 + M.grids.back().ncell;
        c->foff = M.foff;
        c->soff = M.soff;
        c->coff = M.coff;
        c->woff = M.woff;
        c->h_fscene = M.fscene;
        c->h_sscene = M.sscene;
        c->hmax = 0;
        for (const GridP& gg : c->grids) c->hmax = std::max(c->hmax, gg.hmax);
        c->dgrids = dalloc<GridP>(c, n_scenes);
        superset_stencil(c);  // also copies the grids to the device
        c->key_bits = 1;
        while ((1LL << c->key_bits) < c->ncell) c->key_bits++;
        c->sub_bits = std::max(0, std::min(6, (32 - c->key_bits) / 3 * 3));
        c->sscene = dalloc<int>(c, std::max(ns, 1));
        if (ns) h2d(c, c->sscene, M.sscene.data(), ns);
        c->dwoff = dalloc<int>(c, n_scenes + 1);
        h2d(c, c->dwoff, M.woff.data(), n_scenes + 1);
        c->nwmax = 0;
        for (int b = 0; b < n_scenes; b++) c->nwmax = std::max(c->nwmax, M.woff[b + 1] - M.woff[b]);

        // ---- fluid (two sets for the sorted-order ping-pong)
        const bool mx = c->mixed();
        std::vector<double4> ph(nf), vr(nf), acc(nf, make_double4(0, 0, 0, 0));
        std::vector<double2> cn(nf);
        std::vector<double> drho(nf, 0.0), sumw0(nf, 0.0);
        std::vector<int> orig(nf);
        for (int i = 0; i < nf; i++) {
            const double* p = s->fluid_pos + 3 * i;
            ph[i] = make_double4(p[0], p[1], p[2], s->fluid_h[i]);
            double vx = 0, vy = 0, vz = 0;
            if (s->fluid_vel) {
                vx = s->fluid_vel[3 * i];
                vy = s->fluid_vel[3 * i + 1];
                vz = s->fluid_vel[3 * i + 2];
            }
            vr[i] = make_double4(vx, vy, vz, s->fluid_rho0[i]);
            cn[i] = make_double2(s->fluid_c[i], s->fluid_nu[i]);
            if (s->fluid_drho) drho[i] = s->fluid_drho[i];
            orig[i] = i;
        }
        for (int k = 0; k < 2; k++) {
            FluidSet& f = c->fset[k];
            f.ph = dalloc<double4>(c, nf);
            f.vr = dalloc<double4>(c, nf);
            f.drho[0] = dalloc<double>(c, nf);
            f.drho[1] = dalloc<double>(c, nf);
            f.cn = dalloc<double2>(c, nf);
            f.alpha = dalloc<double>(c, nf);
            f.sumw0 = dalloc<double>(c, nf);
            f.acc = dalloc<double4>(c, nf);
            f.orig = dalloc<int>(c, nf);
            f.scene = dalloc<int>(c, std::max(nf, 1));
            f.tid = dalloc<int>(c, nf);
            const size_t rec = 2 * (mx ? sizeof(float4) : sizeof(double4));
            // the pair loops address the record array with 32-bit byte offsets
            if ((size_t)nf * rec > 0x7fffffffull)
                throw ApiError{DSPH_E_INVALID, "too many fluid particles for one context"};
            for (int b = 0; b < 2; b++) f.rec[b] = dalloc<char>(c, rec * std::max(nf, 1));
            // fixed-point span of the mixed records: the domains of all
            // scenes plus one such extent on each side of every axis
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (const GridP& gg : c->grids) {
                lo[0] = std::min(lo[0], gg.minx);
                lo[1] = std::min(lo[1], gg.miny);
                lo[2] = std::min(lo[2], gg.minz);
                hi[0] = std::max(hi[0], gg.maxx);
                hi[1] = std::max(hi[1], gg.maxy);
                hi[2] = std::max(hi[2], gg.maxz);
            }
            // one unit on every axis (the pair loops work in position
            // units), set by the longest extent; at least 2e-10 m, so two
            // records at different fixed-point positions are farther apart
            // than the 1e-10 m below which a pair counts for nothing
            double Lmax = 0.0;
            for (int a = 0; a < 3; a++) Lmax = std::max(Lmax, hi[a] - lo[a]);
            c->fix_lmax = Lmax;
            const double unit = std::max(3.0 * Lmax / 4294967296.0, 2e-10);
            for (int a = 0; a < 3; a++) {
                const double L = hi[a] - lo[a];
                f.qo[a] = lo[a] - L;
                f.qi[a] = 1.0 / unit;
                f.qr[a] = (float)unit;
            }
        }
        {
            FluidSet& f = c->fset[0];
            h2d(c, f.ph, ph.data(), nf);
            h2d(c, f.vr, vr.data(), nf);
            h2d(c, f.drho[0], drho.data(), nf);
            h2d(c, f.drho[1], drho.data(), nf);
            h2d(c, f.cn, cn.data(), nf);
            h2d(c, f.sumw0, sumw0.data(), nf);
            std::vector<double> alpha(nf, 0.0);
            if (s->fluid_alpha) std::copy(s->fluid_alpha, s->fluid_alpha + nf, alpha.begin());
            h2d(c, f.alpha, alpha.data(), nf);
            h2d(c, f.acc, acc.data(), nf);
            h2d(c, f.orig, orig.data(), nf);
            if (nf) h2d(c, f.scene, M.fscene.data(), nf);
            // the fluid property table: one entry per distinct (h, c, rho0,
            // nu), with the particle mass m = (h / 1.3)^3 rho0
            std::map<std::array<double, 4>, int> uniq;
            std::vector<int> tid(nf);
            for (int i = 0; i < nf; i++) {
                const std::array<double, 4> key = {s->fluid_h[i], s->fluid_c[i], s->fluid_rho0[i],
                                                   s->fluid_nu[i]};
                tid[i] = uniq.emplace(key, (int)uniq.size()).first->second;
            }
            const size_t ntab = std::max<size_t>(uniq.size(), 1);
            c->fluid_uniform = uniq.size() == 1;
            if (c->fluid_uniform) {
                const auto& k = uniq.begin()->first;
                const double h13 = k[0] / 1.3;
                c->ufluid.h = k[0];
                c->ufluid.c = k[1];
                c->ufluid.m = h13 * h13 * h13 * k[2];
                c->ufluid.rho0 = k[2];
                c->ufluid.nu = k[3];
            }
            std::vector<double> t4(4 * ntab, 0.0), tn(ntab, 0.0);
            for (const auto& kv : uniq) {
                const double h = kv.first[0], h13 = h / 1.3;
                const int k = kv.second;
                t4[4 * k] = h;
                t4[4 * k + 1] = kv.first[1];
                t4[4 * k + 2] = h13 * h13 * h13 * kv.first[2];
                t4[4 * k + 3] = kv.first[2];
                tn[k] = kv.first[3];
            }
            if (mx) {
                std::vector<float> f4(t4.begin(), t4.end()), fn(tn.begin(), tn.end());
                float* d4 = dalloc<float>(c, 4 * ntab);
                float* dn = dalloc<float>(c, ntab);
                h2d(c, d4, f4.data(), f4.size());
                h2d(c, dn, fn.data(), fn.size());
                HIPCHK(hipStreamSynchronize(c->st));
                for (FluidSet& g : c->fset) {
                    g.tab = d4;
                    g.tabnu = dn;
                }
            } else {
                double* d4 = dalloc<double>(c, 4 * ntab);
                double* dn = dalloc<double>(c, ntab);
                h2d(c, d4, t4.data(), t4.size());
                h2d(c, dn, tn.data(), tn.size());
                for (FluidSet& g : c->fset) {
                    g.tab = d4;
                    g.tabnu = dn;
                }
            }
            h2d(c, f.tid, tid.data(), nf);
            HIPCHK(hipStreamSynchronize(c->st));
        }
        c->fslot = dalloc<int>(c, nf);
        h2d(c, c->fslot, orig.data(), nf);
        c->fmember = dalloc<int>(c, nf);
        c->fkey = dalloc<unsigned>(c, nf);
        c->frank = dalloc<unsigned>(c, nf);
        c->fkey_sorted = dalloc<unsigned>(c, nf);
        {
            // each scene's cells ranked within its own range of indices
            std::vector<unsigned> rank(c->ncell), inv(c->ncell);
            for (const GridP& gg : c->grids) {
                std::vector<unsigned> r1, i1;
                morton_cell_ranks(gg, r1, i1);
                for (int h = 0; h < gg.ncell; h++) {
                    rank[gg.cell0 + h] = gg.cell0 + r1[h];
                    inv[gg.cell0 + h] = gg.cell0 + i1[h];
                }
            }
            // cells whose reference stencil (27 linear offsets) reaches the
            // spare cell: their particles' lists follow the stencil walk
            std::vector<unsigned char> sp(c->ncell, 0);
            for (const GridP& gg : c->grids) {
                const int64_t spare = gg.ncell - 1, gxy = (int64_t)gg.gx * gg.gy;
                for (int64_t h = 0; h < spare; h++) {
                    const int64_t x = h % gg.gx, y = (h / gg.gx) % gg.gy, z = h / gxy;
                    bool hit = false;
                    for (int zz = -1; zz <= 1; zz++)
                        for (int yy = -1; yy <= 1; yy++)
                            for (int xx = -1; xx <= 1; xx++)
                                hit = hit || (z + zz) * gxy + (y + yy) * gg.gx + (x + xx) == spare;
                    sp[gg.cell0 + h] = hit;
                }
                sp[gg.cell0 + spare] = 1;
            }
            c->cspare = dalloc<unsigned char>(c, c->ncell);
            h2d(c, c->cspare, sp.data(), c->ncell);
            c->near_out = dalloc<int>(c, c->ncell);
            fill_int(c->st, c->near_out, c->ncell, 0);
            c->crank = dalloc<unsigned>(c, c->ncell);
            c->cinv = dalloc<unsigned>(c, c->ncell);
            h2d(c, c->crank, rank.data(), c->ncell);
            h2d(c, c->cinv, inv.data(), c->ncell);
        }
        c->fperm = dalloc<int>(c, nf);
        c->fcs = dalloc<int>(c, c->ncell);
        c->fce = dalloc<int>(c, c->ncell);
        build_chunks(c);
        c->hslot = dalloc<unsigned>(c, nf);
        c->fxb = dalloc<double4>(c, nf);
        h2d(c, c->fxb, ph.data(), nf);
        c->fsp = dalloc<double4>(c, nf);
        h2d(c, c->fsp, ph.data(), nf);
        c->dispmax = dalloc<unsigned long long>(c, std::max(1, blocks_for(std::max(nf, 1))));
        c->tsum = dalloc<int>(c, std::max(1, blocks_for(std::max(nf, 1))));
        c->ctl = dalloc<unsigned long long>(c, CTL_WORDS);
        HIPCHK(hipMemsetAsync(c->ctl, 0, CTL_WORDS * sizeof(unsigned long long), c->st));
        c->scs = dalloc<int>(c, c->ncell);
        c->sce = dalloc<int>(c, c->ncell);

        // ---- solids
        std::vector<double4> spr(ns), svel(ns), sw(ns), snrm(ns), zero4(std::max({ns, nc, 1}), make_double4(0, 0, 0, 0));
        std::vector<double> seff(ns);
        std::vector<int> smat(ns), sclus(ns), sclump(ns), sdum(ns), active, live;
        for (int i = 0; i < ns; i++) {
            const double* p = s->solid_pos + 3 * i;
            spr[i] = make_double4(p[0], p[1], p[2], s->solid_radius[i]);
            const double* v = s->solid_vel ? s->solid_vel + 3 * i : nullptr;
            svel[i] = make_double4(v ? v[0] : 0, v ? v[1] : 0, v ? v[2] : 0, s->solid_inv_mass[i]);
            const double* w = s->solid_angvel ? s->solid_angvel + 3 * i : nullptr;
            sw[i] = make_double4(w ? w[0] : 0, w ? w[1] : 0, w ? w[2] : 0, 0);
            snrm[i] = make_double4(0, 0, 0, 0);
            seff[i] = s->solid_eff_radius[i];
            smat[i] = s->solid_material ? s->solid_material[i] : 0;
            sclus[i] = s->solid_cluster ? s->solid_cluster[i] : -1;
            sclump[i] = s->solid_clump ? s->solid_clump[i] : -1;
            sdum[i] = s->solid_dummy ? s->solid_dummy[i] : 0;
            if (sdum[i] != 1) active.push_back(i);
            if (sdum[i] != 1 || svel[i].w > 0.0 || svel[i].x != 0.0 || svel[i].y != 0.0 || svel[i].z != 0.0 ||
                sw[i].x != 0.0 || sw[i].y != 0.0 || sw[i].z != 0.0 || sclump[i] >= 0)
                live.push_back(i);
        }
        Solids& so = c->so;
        so.pr = dalloc<double4>(c, ns);
        so.vel = dalloc<double4>(c, ns);
        so.acc = dalloc<double4>(c, ns);
        so.w = dalloc<double4>(c, ns);
        so.tq = dalloc<double4>(c, ns);
        so.nrm = dalloc<double4>(c, ns);
        so.eff = dalloc<double>(c, ns);
        so.mat = dalloc<int>(c, ns);
        so.clus = dalloc<int>(c, ns);
        so.clump = dalloc<int>(c, ns);
        so.dummy = dalloc<int>(c, ns);
        so.active = dalloc<int>(c, active.size());
        so.nactive = (int)active.size();
        so.live = dalloc<int>(c, live.size());
        so.nlive = (int)live.size();
        so.mrec = c->mixed() ? dalloc<uint4>(c, 2 * (size_t)ns) : nullptr;
        h2d(c, so.pr, spr.data(), ns);
        h2d(c, so.vel, svel.data(), ns);
        h2d(c, so.acc, zero4.data(), ns);
        h2d(c, so.w, sw.data(), ns);
        h2d(c, so.tq, zero4.data(), ns);
        h2d(c, so.nrm, snrm.data(), ns);
        h2d(c, so.eff, seff.data(), ns);
        h2d(c, so.mat, smat.data(), ns);
        h2d(c, so.clus, sclus.data(), ns);
        h2d(c, so.clump, sclump.data(), ns);
        h2d(c, so.dummy, sdum.data(), ns);
        h2d(c, so.active, active.data(), active.size());
        h2d(c, so.live, live.data(), live.size());
        if (s->solid_normal) {
            for (int i = 0; i < ns; i++)
                snrm[i] = make_double4(s->solid_normal[3 * i], s->solid_normal[3 * i + 1], s->solid_normal[3 * i + 2], 0);
            h2d(c, so.nrm, snrm.data(), ns);
        }
        c->skey = dalloc<unsigned>(c, ns);
        c->skey_valid = false;
        {
            // static solids outside the domain (their hashes never change)
            std::vector<char> is_live(ns, 0);
            for (int i : live) is_live[i] = 1;
            c->static_out = 0;
            for (int i = 0; i < ns; i++) {
                if (is_live[i]) continue;
                const GridP& g = c->grids[c->h_sscene[i]];
                if (host_point_hash(g, spr[i].x, spr[i].y, spr[i].z) == g.ncell - 1) c->static_out++;
            }
        }
        c->skey_sorted = dalloc<unsigned>(c, ns);
        c->sperm = dalloc<int>(c, ns);
        c->sxb = dalloc<double4>(c, ns);
        h2d(c, c->sxb, spr.data(), ns);
        c->akey = dalloc<unsigned>(c, active.size());
        c->akey_sorted = dalloc<unsigned>(c, active.size());
        c->aperm = dalloc<int>(c, active.size());
        c->rank_acc = dalloc<int>(c, 3 * std::max<size_t>(active.size(), 1));
        fill_int(c->st, c->rank_acc, 3 * (int64_t)std::max<size_t>(active.size(), 1), 0);
        c->acs = dalloc<int>(c, c->ncell);
        c->ace = dalloc<int>(c, c->ncell);
        fill_int(c->st, c->acs, c->ncell, -1);
        fill_int(c->st, c->ace, c->ncell, -1);
        fill_int(c->st, (int*)c->akey_sorted, (int64_t)active.size(), 0);
        {
            // the static solids' grid, sorted once by their reference hash
            // (a stable counting sort: ascending solid index within a cell)
            c->all_solid_sort = false;
            for (int i = 0; i < ns; i++)
                if (sdum[i] == 1 && (svel[i].w > 0.0 || svel[i].x != 0.0 || svel[i].y != 0.0 ||
                                     svel[i].z != 0.0 || sw[i].x != 0.0 || sw[i].y != 0.0 || sw[i].z != 0.0 ||
                                     sclump[i] >= 0))
                    c->all_solid_sort = true;
            std::vector<int> key(ns, -1), cnt((size_t)c->ncell + 1, 0), perm;
            for (int i = 0; i < ns; i++) {
                if (sdum[i] != 1) continue;
                const GridP& g = c->grids[c->h_sscene[i]];
                key[i] = g.cell0 + host_point_hash(g, spr[i].x, spr[i].y, spr[i].z);
                cnt[key[i] + 1]++;
            }
            for (int h = 0; h < c->ncell; h++) cnt[h + 1] += cnt[h];
            std::vector<int> cs(c->ncell, -1), ce(c->ncell, -1), at(cnt.begin(), cnt.end() - 1);
            perm.resize(cnt[c->ncell]);
            for (int i = 0; i < ns; i++)
                if (key[i] >= 0) perm[at[key[i]]++] = i;
            for (int h = 0; h < c->ncell; h++)
                if (cnt[h + 1] > cnt[h]) {
                    cs[h] = cnt[h];
                    ce[h] = cnt[h + 1];
                }
            c->dcs = dalloc<int>(c, c->ncell);
            c->dce = dalloc<int>(c, c->ncell);
            c->dperm = dalloc<int>(c, std::max<size_t>(perm.size(), 1));
            h2d(c, c->dcs, cs.data(), c->ncell);
            h2d(c, c->dce, ce.data(), c->ncell);
            h2d(c, c->dperm, perm.data(), perm.size());
            HIPCHK(hipStreamSynchronize(c->st));
        }

        // ---- clumps
        std::vector<double4> cpos(nc), cvel(nc), cw(nc), cq(nc);
        std::vector<double> cinv(6 * (size_t)nc);
        std::vector<int> cst(nc), cen(nc);
        for (int k = 0; k < nc; k++) {
            const double* p = s->clump_centroid + 3 * k;
            cpos[k] = make_double4(p[0], p[1], p[2], 0);
            const double* v = s->clump_vel ? s->clump_vel + 3 * k : nullptr;
            cvel[k] = make_double4(v ? v[0] : 0, v ? v[1] : 0, v ? v[2] : 0, s->clump_inv_mass[k]);
            const double* w = s->clump_angvel ? s->clump_angvel + 3 * k : nullptr;
            cw[k] = make_double4(w ? w[0] : 0, w ? w[1] : 0, w ? w[2] : 0, 0);
            if (s->clump_orientation) {
                const double* q = s->clump_orientation + 4 * k;
                cq[k] = make_double4(q[0], q[1], q[2], q[3]);
            } else {
                cq[k] = make_double4(1, 0, 0, 0);
            }
            for (int m = 0; m < 6; m++) cinv[6 * k + m] = s->clump_inv_inertia[6 * k + m];
            cst[k] = s->clump_pebble_start[k];
            cen[k] = s->clump_pebble_end[k];
        }
        Clumps& cl = c->cl;
        cl.pos = dalloc<double4>(c, nc);
        cl.vel = dalloc<double4>(c, nc);
        cl.acc = dalloc<double4>(c, nc);
        cl.w = dalloc<double4>(c, nc);
        cl.tq = dalloc<double4>(c, nc);
        cl.q = dalloc<double4>(c, nc);
        cl.invI = dalloc<double>(c, 6 * (size_t)nc);
        cl.start = dalloc<int>(c, nc);
        cl.end = dalloc<int>(c, nc);
        h2d(c, cl.pos, cpos.data(), nc);
        h2d(c, cl.vel, cvel.data(), nc);
        h2d(c, cl.acc, zero4.data(), nc);
        h2d(c, cl.w, cw.data(), nc);
        h2d(c, cl.tq, zero4.data(), nc);
        h2d(c, cl.q, cq.data(), nc);
        h2d(c, cl.invI, cinv.data(), 6 * (size_t)nc);
        h2d(c, cl.start, cst.data(), nc);
        h2d(c, cl.end, cen.data(), nc);

        // ---- infinite walls and their per-solid spring history
        std::vector<double4> wp(nw), wnv(nw);
        std::vector<int> wm(nw);
        for (int k = 0; k < nw; k++) {
            wp[k] = make_double4(s->wall_point[3 * k], s->wall_point[3 * k + 1], s->wall_point[3 * k + 2], 0);
            wnv[k] = make_double4(s->wall_normal[3 * k], s->wall_normal[3 * k + 1], s->wall_normal[3 * k + 2], 0);
            wm[k] = s->wall_material ? s->wall_material[k] : 0;
        }
        c->wpt = dalloc<double4>(c, nw);
        c->wn = dalloc<double4>(c, nw);
        c->wmat = dalloc<int>(c, nw);
        h2d(c, c->wpt, wp.data(), nw);
        h2d(c, c->wn, wnv.data(), nw);
        h2d(c, c->wmat, wm.data(), nw);
        const size_t nh = 9 * (size_t)ns * c->nwmax;
        c->whist = dalloc<double>(c, nh);
        if (nh) HIPCHK(hipMemsetAsync(c->whist, 0, nh * sizeof(double), c->st));

        // ---- cylinder walls (axes normalised) and their spring history
        {
            const int ny = c->ny = s->n_cylinder;
            std::vector<double4> yp(ny), ya(ny), yv(ny);
            std::vector<int> ym(ny);
            for (int k = 0; k < ny; k++) {
                const double* p = s->cylinder_point + 3 * k;
                const double* ax = s->cylinder_axis + 3 * k;
                const double* w = s->cylinder_angvel ? s->cylinder_angvel + 3 * k : nullptr;
                const double l = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
                yp[k] = make_double4(p[0], p[1], p[2], s->cylinder_radius[k]);
                ya[k] = make_double4(ax[0] / l, ax[1] / l, ax[2] / l, 0);
                yv[k] = make_double4(w ? w[0] : 0, w ? w[1] : 0, w ? w[2] : 0, 0);
                ym[k] = s->cylinder_material ? s->cylinder_material[k] : 0;
            }
            c->ypt = dalloc<double4>(c, ny);
            c->yax = dalloc<double4>(c, ny);
            c->yw = dalloc<double4>(c, ny);
            c->ymat = dalloc<int>(c, ny);
            h2d(c, c->ypt, yp.data(), ny);
            h2d(c, c->yax, ya.data(), ny);
            h2d(c, c->yw, yv.data(), ny);
            h2d(c, c->ymat, ym.data(), ny);
            c->yoff = M.yoff;
            c->dyoff = dalloc<int>(c, n_scenes + 1);
            h2d(c, c->dyoff, M.yoff.data(), n_scenes + 1);
            c->nymax = 0;
            for (int b = 0; b < n_scenes; b++) c->nymax = std::max(c->nymax, M.yoff[b + 1] - M.yoff[b]);
            const size_t nyh = 9 * (size_t)ns * c->nymax;
            c->yhist = dalloc<double>(c, nyh);
            if (nyh) HIPCHK(hipMemsetAsync(c->yhist, 0, nyh * sizeof(double), c->st));
        }

        // ---- contact tables
        c->nmat = nmat;
        c->ncomb = nmat * (nmat + 1) / 2;
        const int ncb = c->ncomb;
        std::vector<double> hz(8 * ncb, 0.0), li(11 * ncb, 0.0), bo(6 * ncb, 0.0);
        for (int k = 0; k < ncb; k++) {
            hz[2 * ncb + k] = 1.0;  // restitution default (hostStructs.h:305)
            bo[0 * ncb + k] = 1.0;  // bond multiplier default (hostStructs.h:346)
            bo[2 * ncb + k] = 2.0;  // kn/ks default
        }
        if (s->hertz) std::copy(s->hertz, s->hertz + 8 * ncb, hz.begin());
        if (s->linear) std::copy(s->linear, s->linear + 11 * ncb, li.begin());
        if (s->bonded) std::copy(s->bonded, s->bonded + 6 * ncb, bo.begin());
        double* dhz = dalloc<double>(c, hz.size());
        double* dli = dalloc<double>(c, li.size());
        double* dbo = dalloc<double>(c, bo.size());
        h2d(c, dhz, hz.data(), hz.size());
        h2d(c, dli, li.data(), li.size());
        h2d(c, dbo, bo.data(), bo.size());
        c->tb = ContactTables{nmat, ncb, dhz, dli, dbo};

        // ---- lists
        Lists& L = c->L;
        L.cnt = dalloc<int>(c, nf);
        L.fscnt = dalloc<int>(c, nf);
        L.fsoff = dalloc<int>(c, (size_t)nf + 1);
        fill_int(c->st, L.fsoff, (int64_t)nf + 1, 0);
        fill_int(c->st, L.cnt, nf, 0);
        alloc_ff(c, nf > 0 ? 128 : 1);
        alloc_sup(c, nf > 0 ? 160 : 1, nf > 0 ? 96 : 1);
        L.sup_cnt = dalloc<int>(c, nf);
        L.fsup_cnt = dalloc<int>(c, nf);
        L.fsm = dalloc<unsigned long long>(c, 2 * (size_t)std::max(nf, 1));
        fill_int(c->st, L.sup_cnt, nf, 0);
        fill_int(c->st, L.fsup_cnt, nf, 0);
        L.grow = dalloc<int>(c, nf);
        {
            std::vector<int> self(nf);
            for (int i = 0; i < nf; i++) self[i] = i;
            h2d(c, L.grow, self.data(), nf);
            HIPCHK(hipStreamSynchronize(c->st));
        }
        alloc_tiles(c);
        alloc_fs(c, std::max<int64_t>(4096, 8LL * nf));
        // solids without pairs keep count 0 and offset 0 throughout
        L.fsr_cnt = dalloc<int>(c, ns);
        L.fsr_off = dalloc<int>(c, (size_t)ns + 1);
        fill_int(c->st, L.fsr_cnt, ns, 0);
        fill_int(c->st, L.fsr_off, (int64_t)ns + 1, 0);
        L.sscnt = dalloc<int>(c, ns);
        L.ssoff = dalloc<int>(c, (size_t)ns + 1);
        L.hoff = dalloc<int>(c, (size_t)ns + 1);
        L.hcnt = dalloc<int>(c, ns);
        L.ssr_cnt = dalloc<int>(c, ns);
        L.ssr_off = dalloc<int>(c, (size_t)ns + 1);
        L.rev_tmp = dalloc<int>(c, ns);
        L.rev_cur = dalloc<int>(c, ns);
        L.rev_tmp2 = dalloc<int>(c, ns);
        L.rev_cur2 = dalloc<int>(c, ns);
        fill_int(c->st, L.ssoff, (int64_t)ns + 1, 0);
        fill_int(c->st, L.sscnt, ns, 0);  // dummies are never searched: their count stays 0
        fill_int(c->st, L.hoff, (int64_t)ns + 1, 0);
        fill_int(c->st, L.hcnt, ns, 0);
        fill_int(c->st, L.ssr_cnt, ns, 0);
        fill_int(c->st, L.ssr_off, (int64_t)ns + 1, 0);
        fill_int(c->st, L.rev_tmp, ns, 0);  // zero between reverse builds
        fill_int(c->st, L.rev_tmp2, ns, 0);
        alloc_ss(c, std::max<int64_t>(4096, 6LL * (int64_t)active.size() + 4096), false);

        // ---- scratch
        const int64_t nmax = std::max<int64_t>(std::max(nf, ns), 1);
        c->radix_tmp = dalloc<unsigned>(c, radix_tmp_size(nmax));
        int64_t sc = std::max<int64_t>(scan_tmp_size(std::max<int64_t>(nmax, c->ncell) + 1),
                                       scan_tmp_size(256 * ((nmax + 1023) / 1024)));
        c->scan_tmp = dalloc<int>(c, sc + 64);
        c->status = dalloc<unsigned long long>(c, ST_WORDS);
        HIPCHK(hipMemsetAsync(c->status, 0, ST_WORDS * sizeof(unsigned long long), c->st));
        {
            FluidArgs fa = fluid_args(c);
            fluid_records(c->st, fa, 0, mx);
            fluid_records(c->st, fa, 1, mx);
            solid_records(c->st, c->so, c->ns, c->fset[0]);
        }
        force_rebuild(c);
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipGetLastError());
        c->uploaded = true;
        if (n_scenes == 1) c->host = std::make_unique<Merged>(std::move(M));
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_set_capacity(dsph_ctx* c, int32_t ff, int64_t fs, int64_t ss) {
    if (!c || !c->uploaded) return DSPH_E_STATE;
    try {
        HIPCHK(hipStreamSynchronize(c->st));
        if (ff > 0) alloc_ff(c, ff);
        if (fs > 0) alloc_fs(c, fs);
        if (ss > 0) alloc_ss(c, ss, true);
        invalidate_graphs(c);
        force_rebuild(c);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

namespace {
// addBondData, initial branch (dataBase.cpp:270-308), on the device: bonds
// are the contacts of the current list in one cluster that touch
// physically, in the contact list's order (flags, scan, fill)
void initial_bonds(dsph_ctx* c) {
    const int64_t cap = std::max<int64_t>(c->L.ss_cap, 1);
    int* flag = dalloc<int>(c, (size_t)cap);
    int* off = dalloc<int>(c, (size_t)cap + 1);
    int* tmp = dalloc<int>(c, (size_t)scan_tmp_size(cap) + 64);
    const DemArgs da = dem_args(c);
    fill_int(c->st, flag, cap, 0);
    dem_bond_flags(c->st, da, flag);
    exclusive_scan(c->st, flag, off, cap, tmp);
    int nb = 0;
    d2h(c, &nb, off + cap, 1);
    HIPCHK(hipStreamSynchronize(c->st));
    alloc_bonds(c, nb);
    dem_bond_fill(c->st, dem_args(c), off, c->B);
    HIPCHK(hipStreamSynchronize(c->st));
    dfree(c, flag);
    dfree(c, off);
    dfree(c, tmp);
}
}  // namespace

int dsph_initialize(dsph_ctx* c, int compute_normals) {
    if (!c || !c->uploaded) return DSPH_E_STATE;
    try {
        search_checked(c, 0, 1);
        initial_bonds(c);
        if (compute_normals && c->ns > 0) {
            // the solid grid of the initial positions (the gated search sorts
            // only on fluid steps)
            radix_sort_iota(c->st, c->skey, c->ns, c->key_bits, c->skey_sorted, c->sperm, c->radix_tmp);
            cell_spans(c->st, c->skey_sorted, c->ns, c->scs, c->sce, c->ncell, nullptr);
            solid_normals(c->st, grids_view(c), c->so, c->ns, c->scs, c->sce, c->sperm);
            solid_records(c->st, c->so, c->ns, c->fset[0]);
        }
        HIPCHK(hipStreamSynchronize(c->st));
        invalidate_graphs(c);
        c->initialized = true;
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

#define API_GUARD(c)                                   \
    if (!c) return DSPH_E_INVALID;                     \
    if (!c->uploaded) return DSPH_E_STATE;

int dsph_neighbor_search(dsph_ctx* c, int32_t i_step, int32_t gap) {
    API_GUARD(c);
    if (gap < 1) return DSPH_E_INVALID;
    try {
        search_checked(c, i_step, gap);  // grows overflowing lists (one status read)
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

namespace {
// The dt the phases compute with: the caller's, or with per-scene time
// steps set (the caller then passes 0) 1, so every phase hands its kernels
// the factor of each scene's own step
static double phase_dt(const dsph_ctx* c, double dt) {
    if (!c->sdt) return dt;
    if (dt != 0.0)
        throw ApiError{DSPH_E_INVALID, "per-scene time steps are set: pass dt = 0 (or clear them)"};
    return 1.0;
}
}  // namespace

int dsph_set_scene_time_steps(dsph_ctx* c, const double* dt) {
    API_GUARD(c);
    if (!c->uploaded) return DSPH_E_STATE;
    try {
        HIPCHK(hipStreamSynchronize(c->st));
        if (!dt) {
            dfree(c, c->sdt);
            c->sdt = nullptr;
        } else {
            require(c->nscene > 1, "per-scene time steps need a batch of scenes (dsph_upload_batch)");
            for (int k = 0; k < c->nscene; k++)
                require(std::isfinite(dt[k]) && dt[k] > 0.0, "scene time steps must be positive and finite");
            if (!c->sdt) c->sdt = dalloc<double>(c, c->nscene);
            h2d(c, c->sdt, dt, c->nscene);
            HIPCHK(hipStreamSynchronize(c->st));
        }
        invalidate_graphs(c);
        c->graph_dt = -1.0;
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_solid_before_contact(dsph_ctx* c, double dt) {
    API_GUARD(c);
    try {
        dt = phase_dt(c, dt);
        solid_before(c, dt);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_fluid_integrate(dsph_ctx* c, double dt, int32_t i_step, int32_t gap) {
    API_GUARD(c);
    if (gap < 1) return DSPH_E_INVALID;
    try {
        dt = phase_dt(c, dt);
        fluid_integrate(c, dt, i_step, gap);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_contact_forces(dsph_ctx* c, double dt) {
    API_GUARD(c);
    try {
        dt = phase_dt(c, dt);
        contacts(c, dt);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_solid_after_contact(dsph_ctx* c, double dt) {
    API_GUARD(c);
    try {
        dt = phase_dt(c, dt);
        solid_after(c, dt);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_step(dsph_ctx* c, int32_t i_step, double dt, int32_t gap, int check_capacity) {
    API_GUARD(c);
    if (!c->initialized) return DSPH_E_STATE;
    if (gap < 1) return DSPH_E_INVALID;
    try {
        dt = phase_dt(c, dt);
        if (check_capacity)
            search_checked(c, i_step, gap);
        else
            neighbor_search(c, i_step, gap, true);
        solid_before(c, dt);
        fluid_integrate(c, dt, i_step, gap);
        contacts(c, dt);
        solid_after(c, dt);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_step_n(dsph_ctx* c, int32_t i0, int32_t n, double dt, int32_t gap) {
    API_GUARD(c);
    if (!c->initialized) return DSPH_E_STATE;
    if (gap < 1 || n < 0) return DSPH_E_INVALID;
    try {
        dt = phase_dt(c, dt);
        const bool timing = c->timing;
        if (dt != c->graph_dt) {
            invalidate_graphs(c);
            c->graph_dt = dt;
        }
        // the state before the run, to roll back to if a list overflows
        // inside it (graph-replayed searches cannot grow anything)
        const HostState before = snapshot_device(c);
        for (int k = 0; k < n; k++) {
            const int is = i0 + k;
            if (timing) {  // events cannot live inside a captured graph
                step_body(c, is, dt, gap);
                continue;
            }
            const uint64_t key = step_key(c, is, gap);
            auto it = c->graphs.find(key);
            if (it == c->graphs.end()) {
                StepGraph sg;
                HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
                try {
                    step_body(c, is, dt, gap);
                } catch (...) {
                    hipGraph_t tmp;
                    (void)hipStreamEndCapture(c->st, &tmp);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(c->st, &sg.graph));
                HIPCHK(hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0));
                sg.dcur_after = c->dcur;
                c->graphs[key] = sg;
                HIPCHK(hipGraphLaunch(sg.exec, c->st));
            } else {
                HIPCHK(hipGraphLaunch(it->second.exec, c->st));
                c->dcur = it->second.dcur_after;
            }
        }
        unsigned long long w[ST_WORDS];
        read_status(c, w);
        // a NaN/inf position met by the last search: reported, not rolled
        // back (the steps before it would only run into it again)
        check_finite(w);
        if (w[ST_FF_OVERFLOW] | w[ST_FS_OVERFLOW] | w[ST_SS_OVERFLOW] | w[ST_SUP_OVERFLOW]) {
            // some search of the run overflowed a list and every later
            // kernel read cut-short rows: back to the state before the run
            // (with its status words, so the flags are clear again), then
            // the same steps eagerly, each search checked: a search that
            // overflows grows its lists and runs again, as in dsph_step
            restore_device(c, before);
            for (int k = 0; k < n; k++) {
                const int is = i0 + k;
                search_checked(c, is, gap);
                solid_before(c, dt);
                fluid_integrate(c, dt, is, gap);
                contacts(c, dt);
                solid_after(c, dt);
            }
            HIPCHK(hipStreamSynchronize(c->st));
        }
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_synchronize(dsph_ctx* c) {
    if (!c) return DSPH_E_INVALID;
    return hipStreamSynchronize(c->st) == hipSuccess ? DSPH_OK : DSPH_E_HIP;
}

// ------------------------------------------------------------ checkpoint
namespace {
// the checkpoint blob: this header, n_arrays byte counts (int64), then the
// arrays of state_arrays() back to back
struct StateBlobHeader {
    char magic[8];
    int32_t version;
    int32_t prec;
    int32_t nf, ns, nc, nw, nscene, nbond, ncell, ncomb, nactive, nwmax;
    int32_t K, supK, fsupK, reorder_period;
    int64_t fs_cap, ss_cap;
    int32_t dcur, skey_valid, tune_bs, tune_bmem, tune_group, nymax;
    uint32_t flags;
    int32_t tune_tile;
    double skin_h, gv[3];
    int64_t n_arrays, payload;
};
constexpr char kStateMagic[8] = {'D', 'S', 'P', 'H', 'S', 'T', 'A', 'T'};
constexpr int32_t kStateVersion = 4;

static StateBlobHeader state_header(dsph_ctx* c, const std::vector<StateArr>& v) {
    StateBlobHeader h{};
    std::memcpy(h.magic, kStateMagic, 8);
    h.version = kStateVersion;
    h.prec = c->prec;
    h.nf = c->nf;
    h.ns = c->ns;
    h.nc = c->nc;
    h.nw = c->nw;
    h.nscene = c->nscene;
    h.nbond = c->B.n;
    h.ncell = c->ncell;
    h.ncomb = c->ncomb;
    h.nactive = c->so.nactive;
    h.nwmax = c->nwmax;
    h.nymax = c->nymax;
    h.K = c->L.K;
    h.supK = c->L.supK;
    h.fsupK = c->L.fsupK;
    h.reorder_period = c->reorder_period;
    h.fs_cap = c->L.fs_cap;
    h.ss_cap = c->L.ss_cap;
    h.dcur = c->dcur;
    h.skey_valid = c->skey_valid ? 1 : 0;
    h.tune_bs = c->tune_bs;
    h.tune_bmem = c->tune_bmem;
    h.tune_group = c->tune_group;
    h.tune_tile = c->tune_tile;
    h.flags = c->flags;
    h.skin_h = c->skin_h;
    for (int k = 0; k < 3; k++) h.gv[k] = c->gv[k];
    h.n_arrays = (int64_t)v.size();
    int64_t tot = 0;
    for (const StateArr& a : v) tot += (int64_t)a.bytes;
    h.payload = tot;
    return h;
}

static int64_t blob_size(const StateBlobHeader& h) {
    return (int64_t)sizeof(StateBlobHeader) + 8 * h.n_arrays + h.payload;
}
}  // namespace

int dsph_save_state(dsph_ctx* c, void* buffer, int64_t cap, int64_t* size_out) {
    API_GUARD(c);
    if (!c->initialized) return DSPH_E_STATE;
    try {
        const std::vector<StateArr> v = state_arrays(c);
        const StateBlobHeader h = state_header(c, v);
        const int64_t need = blob_size(h);
        if (size_out) *size_out = need;
        if (!buffer) return DSPH_OK;
        if (cap < need) throw ApiError{DSPH_E_INVALID, "state buffer too small: " + std::to_string(need) + " bytes needed"};
        char* out = (char*)buffer;
        std::memcpy(out, &h, sizeof(h));
        int64_t* sizes = (int64_t*)(out + sizeof(h));
        char* data = out + sizeof(h) + 8 * h.n_arrays;
        for (size_t k = 0; k < v.size(); k++) {
            sizes[k] = (int64_t)v[k].bytes;
            HIPCHK(hipMemcpyAsync(data, v[k].p, v[k].bytes, hipMemcpyDeviceToHost, c->st));
            data += v[k].bytes;
        }
        HIPCHK(hipStreamSynchronize(c->st));
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_load_state(dsph_ctx* c, const void* buffer, int64_t size) {
    API_GUARD(c);
    if (!buffer || size < (int64_t)sizeof(StateBlobHeader)) return DSPH_E_INVALID;
    try {
        StateBlobHeader h;
        std::memcpy(&h, buffer, sizeof(h));
        require(std::memcmp(h.magic, kStateMagic, 8) == 0 && h.version == kStateVersion,
                "not a state blob of this library version");
        require(size >= blob_size(h), "state blob truncated");
        require(h.prec == c->prec, "state blob of the other precision");
        require(h.nf == c->nf && h.ns == c->ns && h.nc == c->nc && h.nw == c->nw && h.nscene == c->nscene &&
                    h.ncell == c->ncell && h.ncomb == c->ncomb && h.nactive == c->so.nactive &&
                    h.nwmax == c->nwmax && h.nymax == c->nymax,
                "state blob of another scene (particle, wall, cell or material counts differ)");
        require(h.tune_bs >= 64 && h.tune_bs <= 512 && h.tune_bs % 64 == 0 && h.tune_bmem >= 1 &&
                    h.tune_bmem <= 64 && (h.tune_group == 1 || h.tune_group == 2 || h.tune_group == 4 ||
                                          h.tune_group == 8) && h.K >= 1 && h.supK >= 1 && h.fsupK >= 1 &&
                    h.fs_cap >= 1 && h.ss_cap >= 1 && h.nbond >= 0 && h.skin_h >= 0.0 &&
                    (h.tune_tile == 0 || h.tune_tile == 1),
                "state blob header out of range");
        HIPCHK(hipStreamSynchronize(c->st));
        invalidate_graphs(c);
        // what the saved run's results depend on besides its arrays
        c->flags = h.flags;
        for (int k = 0; k < 3; k++) c->gv[k] = h.gv[k];
        c->reorder_period = h.reorder_period;
        c->tune_bs = h.tune_bs;
        c->tune_group = h.tune_group;
        if (c->tune_bmem != h.tune_bmem) {
            c->tune_bmem = h.tune_bmem;
            build_chunks(c);
        }
        c->skin_h = h.skin_h;
        superset_stencil(c);
        c->tune_tile = h.tune_tile;
        if ((c->T.n > 0) != tiles_on(c)) alloc_tiles(c);
        // its capacities
        if (c->L.K != h.K) alloc_ff(c, h.K);
        if (c->L.supK != h.supK || c->L.fsupK != h.fsupK) alloc_sup(c, h.supK, h.fsupK);
        if (c->L.fs_cap != h.fs_cap) alloc_fs(c, h.fs_cap);
        if (c->L.ss_cap != h.ss_cap) alloc_ss(c, h.ss_cap, false);
        if (c->B.n != h.nbond || !c->B.i) alloc_bonds(c, h.nbond);
        const std::vector<StateArr> v = state_arrays(c);
        require((int64_t)v.size() == h.n_arrays, "state blob does not match this context's arrays");
        const char* in = (const char*)buffer;
        const int64_t* sizes = (const int64_t*)(in + sizeof(h));
        for (size_t k = 0; k < v.size(); k++)
            require(sizes[k] == (int64_t)v[k].bytes, "state blob does not match this context's arrays");
        const char* data = in + sizeof(h) + 8 * h.n_arrays;
        for (size_t k = 0; k < v.size(); k++) {
            HIPCHK(hipMemcpyAsync(v[k].p, data, v[k].bytes, hipMemcpyHostToDevice, c->st));
            data += v[k].bytes;
        }
        c->dcur = h.dcur;
        c->skey_valid = h.skey_valid != 0;
        // the solids' mixed records follow from the loaded state
        solid_records(c->st, c->so, c->ns, c->fset[0]);
        HIPCHK(hipStreamSynchronize(c->st));
        c->initialized = true;
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_set_contact_models(dsph_ctx* c, const double* hertz, const double* linear,
                            const double* bonded) {
    API_GUARD(c);
    if (!c->uploaded) return DSPH_E_STATE;
    try {
        const int ncb = c->tb.ncomb;
        if (hertz) h2d(c, const_cast<double*>(c->tb.hertz), hertz, 8 * (size_t)ncb);
        if (linear) h2d(c, const_cast<double*>(c->tb.linear), linear, 11 * (size_t)ncb);
        if (bonded) {
            h2d(c, const_cast<double*>(c->tb.bonded), bonded, 6 * (size_t)ncb);
        }
        HIPCHK(hipStreamSynchronize(c->st));
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_set_clump_velocity(dsph_ctx* c, int32_t idx, const double v[3]) {
    API_GUARD(c);
    if (!c->uploaded) return DSPH_E_STATE;
    if (idx < 0 || idx >= c->nc || !v) return DSPH_E_INVALID;
    try {
        double4 cur;
        d2h(c, &cur, c->cl.vel + idx, 1);
        HIPCHK(hipStreamSynchronize(c->st));
        cur.x = v[0];
        cur.y = v[1];
        cur.z = v[2];  // .w keeps the inverse mass
        h2d(c, c->cl.vel + idx, &cur, 1);
        HIPCHK(hipStreamSynchronize(c->st));
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_add_external_force(dsph_ctx* c, int32_t idx, const double f[3]) {
    API_GUARD(c);
    if (idx < 0 || idx >= c->ns || !f) return DSPH_E_INVALID;
    try {
        dem_external_force(c->st, dem_args(c), idx, f[0], f[1], f[2]);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_add_global_damping(dsph_ctx* c, int32_t idx, double cd) {
    API_GUARD(c);
    if (idx < 0 || idx >= c->ns) return DSPH_E_INVALID;
    try {
        dem_global_damping(c->st, dem_args(c), idx, cd);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

// ------------------------------------------------------------------ frames
namespace {
// byte layout of a frame's staging block: the arrays of the species asked
// for, each 256-byte aligned, the pair count, then the pair arrays (n pairs
// packed consecutively in room for pcap); the part before the pairs is
// copied by dsph_frame_begin, the 32 n bytes of pairs by dsph_frame_write
struct FrameLayout {
    size_t fpos, fh, fvel, frho, fp, spos, srad, sclus, sclump, svel, sw, snrm, head, pairs, bytes;
};
static FrameLayout frame_layout(int64_t nf, int64_t ns, int64_t pcap, uint32_t species) {
    FrameLayout L{};
    size_t at = 0;
    auto take = [&](size_t n) {
        const size_t o = at;
        at += (n + 255) / 256 * 256;
        return o;
    };
    const size_t f = (species & DSPH_FRAME_FLUID) ? (size_t)nf : 0;
    const size_t so = (species & DSPH_FRAME_SOLID) ? (size_t)ns : 0;
    const size_t p = (species & DSPH_FRAME_FLUID_SOLID) ? (size_t)pcap : 0;
    L.fpos = take(12 * f);
    L.fh = take(4 * f);
    L.fvel = take(12 * f);
    L.frho = take(4 * f);
    L.fp = take(4 * f);
    L.spos = take(12 * so);
    L.srad = take(4 * so);
    L.sclus = take(4 * so);
    L.sclump = take(4 * so);
    L.svel = take(12 * so);
    L.sw = take(12 * so);
    L.snrm = take(12 * so);
    L.head = take(8);
    L.pairs = take(32 * p);
    L.bytes = at;
    return L;
}

// One binary appended VTU file (raw encoding, UInt64 block headers) of n
// vertex cells: the FieldData TIME and STEP, the points, the cells, and the
// point arrays, in the reference's element order (Points, Cells, PointData).
struct VtuArray {
    const char* name;
    const char* type;  // "Float32", "Int32"
    int comps;
    const void* data;
    size_t bytes;
};
static bool write_vtu(const std::string& path, double time, int32_t step, int64_t n, const float* points,
               const std::vector<VtuArray>& arrays, const char* scalars) {
    std::vector<int32_t> conn((size_t)n), offs((size_t)n);
    std::vector<uint8_t> types((size_t)n, 1);  // VTK_VERTEX
    for (int64_t i = 0; i < n; i++) {
        conn[i] = (int32_t)i;
        offs[i] = (int32_t)(i + 1);
    }
    const float tf = (float)time;
    struct Block {
        const void* p;
        size_t bytes;
    };
    std::vector<Block> blocks;
    size_t off = 0;
    auto next = [&](const void* p, size_t b) {
        const size_t o = off;
        blocks.push_back({p, b});
        off += 8 + b;
        return o;
    };
    std::string x;
    x += "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
         "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n  <UnstructuredGrid>\n";
    x += "    <FieldData>\n";
    x += "      <DataArray type=\"Float32\" Name=\"TIME\" NumberOfTuples=\"1\" format=\"appended\" offset=\"" +
         std::to_string(next(&tf, 4)) + "\"/>\n";
    x += "      <DataArray type=\"Int32\" Name=\"STEP\" NumberOfTuples=\"1\" format=\"appended\" offset=\"" +
         std::to_string(next(&step, 4)) + "\"/>\n";
    x += "    </FieldData>\n";
    x += "    <Piece NumberOfPoints=\"" + std::to_string(n) + "\" NumberOfCells=\"" + std::to_string(n) + "\">\n";
    x += "      <Points>\n        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"appended\" "
         "offset=\"" + std::to_string(next(points, 12 * (size_t)n)) + "\"/>\n      </Points>\n";
    x += "      <Cells>\n";
    x += "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"appended\" offset=\"" +
         std::to_string(next(conn.data(), 4 * (size_t)n)) + "\"/>\n";
    x += "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"appended\" offset=\"" +
         std::to_string(next(offs.data(), 4 * (size_t)n)) + "\"/>\n";
    x += "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" +
         std::to_string(next(types.data(), (size_t)n)) + "\"/>\n";
    x += "      </Cells>\n";
    x += std::string("      <PointData Scalars=\"") + scalars + "\">\n";
    for (const VtuArray& a : arrays) {
        x += std::string("        <DataArray type=\"") + a.type + "\" Name=\"" + a.name + "\"";
        if (a.comps > 1) x += " NumberOfComponents=\"" + std::to_string(a.comps) + "\"";
        x += " format=\"appended\" offset=\"" + std::to_string(next(a.data, a.bytes)) + "\"/>\n";
    }
    x += "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    bool ok = std::fwrite(x.data(), 1, x.size(), fp) == x.size();
    for (const Block& b : blocks) {
        const uint64_t h = b.bytes;
        ok = ok && std::fwrite(&h, 8, 1, fp) == 1;
        if (b.bytes) ok = ok && std::fwrite(b.p, 1, b.bytes, fp) == b.bytes;
    }
    const char tail[] = "\n  </AppendedData>\n</VTKFile>\n";
    ok = ok && std::fwrite(tail, 1, sizeof(tail) - 1, fp) == sizeof(tail) - 1;
    return std::fclose(fp) == 0 && ok;
}
}  // namespace

int dsph_frame_begin(dsph_ctx* c, int32_t slot, int32_t i_step, double time, uint32_t species) {
    API_GUARD(c);
    if (slot < 0 || slot >= DSPH_FRAME_SLOTS || species == 0 || (species & ~(uint32_t)DSPH_FRAME_ALL))
        return DSPH_E_INVALID;
    try {
        auto& F = c->frames[slot];
        const int64_t nf = c->nf, ns = c->ns;
        const int64_t pcap = (nf > 0 && ns > 0) ? c->L.fs_cap : 0;
        const FrameLayout L = frame_layout(nf, ns, pcap, species);
        // this slot's previous copy has landed before its blocks are reused
        if (F.copied) HIPCHK(hipEventSynchronize(F.copied));
        if (!c->cst) HIPCHK(hipStreamCreateWithFlags(&c->cst, hipStreamNonBlocking));
        if (!F.packed) HIPCHK(hipEventCreateWithFlags(&F.packed, hipEventDisableTiming));
        if (!F.copied) HIPCHK(hipEventCreateWithFlags(&F.copied, hipEventDisableTiming));
        if (F.dev_cap < L.bytes) {
            dfree(c, F.dev);
            F.dev = nullptr;
            F.dev_cap = 0;
            F.dev = dalloc<char>(c, L.bytes);
            F.dev_cap = L.bytes;
        }
        if (F.cap < L.bytes) {
            if (F.host) HIPCHK(hipHostFree(F.host));
            F.host = nullptr;
            F.cap = 0;
            if (hipHostMalloc(&F.host, L.bytes, hipHostMallocDefault) != hipSuccess)
                throw ApiError{DSPH_E_NOMEM, "pinned host memory for a frame"};
            F.cap = L.bytes;
        }
        char* d = (char*)F.dev;
        FrameOut o;
        o.fpos = (float*)(d + L.fpos);
        o.fh = (float*)(d + L.fh);
        o.fvel = (float*)(d + L.fvel);
        o.frho = (float*)(d + L.frho);
        o.fp = (float*)(d + L.fp);
        o.spos = (float*)(d + L.spos);
        o.srad = (float*)(d + L.srad);
        o.sclus = (int*)(d + L.sclus);
        o.sclump = (int*)(d + L.sclump);
        o.svel = (float*)(d + L.svel);
        o.sw = (float*)(d + L.sw);
        o.snrm = (float*)(d + L.snrm);
        o.head = (long long*)(d + L.head);
        o.pairs = (float*)(d + L.pairs);
        // the records of buffer dcur hold the fluid's current dRho; the
        // packing is on the context stream, its copy on the copy stream, so
        // the steps that follow run beside the copy
        {
            FamTimer t(c, F_FPACK);
            frame_pack(c->st, species, c->fset[0], (int)nf, c->dcur, c->so, (int)ns, c->L, pcap, o);
        }
        HIPCHK(hipEventRecord(F.packed, c->st));
        HIPCHK(hipStreamWaitEvent(c->cst, F.packed, 0));
        {
            FamTimer t(c, F_FCOPY, c->cst);
            HIPCHK(hipMemcpyAsync(F.host, F.dev, L.pairs, hipMemcpyDeviceToHost, c->cst));
        }
        HIPCHK(hipEventRecord(F.copied, c->cst));
        F.species = species;
        F.step = i_step;
        F.time = time;
        F.nf = nf;
        F.ns = ns;
        F.pcap = pcap;
        F.pending = true;
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_frame_write(dsph_ctx* c, int32_t slot, const char* dir, int32_t frame) {
    if (!c || !dir || slot < 0 || slot >= DSPH_FRAME_SLOTS || frame < 0) return DSPH_E_INVALID;
    auto& F = c->frames[slot];
    if (!F.pending) return DSPH_E_STATE;
    if (hipEventSynchronize(F.copied) != hipSuccess) return DSPH_E_HIP;
    const FrameLayout L = frame_layout(F.nf, F.ns, F.pcap, F.species);
    char* const h = (char*)F.host;
    long long np = 0;
    if (F.species & DSPH_FRAME_FLUID_SOLID) {
        std::memcpy(&np, h + L.head, 8);
        if (np < 0) np = 0;  // the list had overflowed: no pairs
        if (np > 0) {
            // the pairs the device packed, and no more (the slot's staging
            // block is not packed again before this call returns)
            if (hipMemcpyAsync(h + L.pairs, (const char*)F.dev + L.pairs, 32 * (size_t)np, hipMemcpyDeviceToHost,
                               c->cst) != hipSuccess ||
                hipEventRecord(F.copied, c->cst) != hipSuccess || hipEventSynchronize(F.copied) != hipSuccess)
                return DSPH_E_HIP;
        }
    }
    char num[16];
    std::snprintf(num, sizeof(num), "%04d", (int)frame);
    const std::string base = std::string(dir) + "/";
    bool ok = true;
    if (F.species & DSPH_FRAME_FLUID) {
        const size_t n = (size_t)F.nf;
        ok = ok && write_vtu(base + "fluid_" + num + ".vtu", F.time, F.step, F.nf, (const float*)(h + L.fpos),
                             {{"smooth length", "Float32", 1, h + L.fh, 4 * n},
                              {"velocity", "Float32", 3, h + L.fvel, 12 * n},
                              {"density", "Float32", 1, h + L.frho, 4 * n},
                              {"pressure", "Float32", 1, h + L.fp, 4 * n}},
                             "smooth length");
    }
    if (F.species & DSPH_FRAME_SOLID) {
        const size_t n = (size_t)F.ns;
        ok = ok && write_vtu(base + "solid_" + num + ".vtu", F.time, F.step, F.ns, (const float*)(h + L.spos),
                             {{"radius", "Float32", 1, h + L.srad, 4 * n},
                              {"clusterID", "Int32", 1, h + L.sclus, 4 * n},
                              {"clumpID", "Int32", 1, h + L.sclump, 4 * n},
                              {"velocity", "Float32", 3, h + L.svel, 12 * n},
                              {"angularVelocity", "Float32", 3, h + L.sw, 12 * n},
                              {"normal", "Float32", 3, h + L.snrm, 12 * n}},
                             "radius");
    }
    if (F.species & DSPH_FRAME_FLUID_SOLID) {
        const size_t n = (size_t)np;
        const char* mid = h + L.pairs;
        const char* force = mid + 12 * n;
        const char* fid = force + 12 * n;
        const char* sid = fid + 4 * n;
        ok = ok && write_vtu(base + "fluid2Solid_" + num + ".vtu", F.time, F.step, np, (const float*)mid,
                             {{"force", "Float32", 3, force, 12 * n},
                              {"fluidID", "Int32", 1, fid, 4 * n},
                              {"solidID", "Int32", 1, sid, 4 * n}},
                             "force");
    }
    F.pending = false;
    return ok ? DSPH_OK : DSPH_E_IO;
}

// ------------------------------------------------------- mid-run additions
extern "C++" {
namespace {
// the bond arrays on the host
struct HostBonds {
    std::vector<int> i, j, on;
    std::vector<double4> nrm, pt, Fs, Tb;
    std::vector<double> Fn, Tt;
    void get(dsph_ctx* c) {
        const Bonds& B = c->B;
        const size_t n = (size_t)B.n;
        i.resize(n); j.resize(n); on.resize(n); nrm.resize(n); pt.resize(n); Fs.resize(n); Tb.resize(n);
        Fn.resize(n); Tt.resize(n);
        d2h(c, i.data(), B.i, n); d2h(c, j.data(), B.j, n); d2h(c, on.data(), B.on, n);
        d2h(c, nrm.data(), B.nrm, n); d2h(c, pt.data(), B.pt, n); d2h(c, Fs.data(), B.Fs, n);
        d2h(c, Tb.data(), B.Tb, n); d2h(c, Fn.data(), B.Fn, n); d2h(c, Tt.data(), B.Tt, n);
        HIPCHK(hipStreamSynchronize(c->st));
    }
    void append(const HostBonds& o, size_t k) {
        i.push_back(o.i[k]); j.push_back(o.j[k]); on.push_back(o.on[k]); nrm.push_back(o.nrm[k]);
        pt.push_back(o.pt[k]); Fs.push_back(o.Fs[k]); Tb.push_back(o.Tb[k]); Fn.push_back(o.Fn[k]);
        Tt.push_back(o.Tt[k]);
    }
    void put(dsph_ctx* c) const {
        const size_t n = i.size();
        alloc_bonds(c, (int)n);
        const Bonds& B = c->B;
        h2d(c, B.i, i.data(), n); h2d(c, B.j, j.data(), n); h2d(c, B.on, on.data(), n);
        h2d(c, B.nrm, nrm.data(), n); h2d(c, B.pt, pt.data(), n); h2d(c, B.Fs, Fs.data(), n);
        h2d(c, B.Tb, Tb.data(), n); h2d(c, B.Fn, Fn.data(), n); h2d(c, B.Tt, Tt.data(), n);
        HIPCHK(hipStreamSynchronize(c->st));
    }
};
template <class T>
void cat(std::vector<T>& v, const std::vector<T>& a, const T* b, size_t nb, size_t width, T dflt) {
    v = a;
    put(v, b, nb, width, dflt);
}
// xyz of caller-order arrays [3n] into the .xyz of device quads (w kept)
void put_xyz(dsph_ctx* c, double4* d, const std::vector<double>& v, size_t n) {
    std::vector<double4> q(n);
    d2h(c, q.data(), d, n);
    HIPCHK(hipStreamSynchronize(c->st));
    for (size_t k = 0; k < n; k++) {
        q[k].x = v[3 * k];
        q[k].y = v[3 * k + 1];
        q[k].z = v[3 * k + 2];
    }
    h2d(c, d, q.data(), n);
}
}  // namespace
}  // extern "C++"

int dsph_append_scene(dsph_ctx* c, const dsph_scene* e) {
    API_GUARD(c);
    if (!e) return DSPH_E_INVALID;
    if (!c->initialized || !c->host) return DSPH_E_STATE;
    try {
        require(e->n_wall == 0 && e->n_cylinder == 0, "walls cannot be added after the start");
        {
            // checked as a part of the running scene: its domain and
            // materials (extra's own are ignored)
            dsph_scene ev = *e;
            const dsph_scene& run = c->host->flat;
            std::memcpy(ev.domain_origin, run.domain_origin, sizeof(ev.domain_origin));
            std::memcpy(ev.domain_size, run.domain_size, sizeof(ev.domain_size));
            ev.n_material = run.n_material;
            ev.hertz = ev.linear = ev.bonded = nullptr;
            validate_scene(&ev);
        }
        const int nf0 = c->nf, ns0 = c->ns, nc0 = c->nc;
        const int nf1 = e->n_fluid, ns1 = e->n_solid, nc1 = e->n_clump;
        for (int k = 0; k < nc1; k++)
            require(e->clump_pebble_start[k] >= 0 && e->clump_pebble_end[k] <= ns1 &&
                        e->clump_pebble_start[k] <= e->clump_pebble_end[k],
                    "an added clump's pebbles must be added solids");
        HIPCHK(hipStreamSynchronize(c->st));
        if (c->cst) HIPCHK(hipStreamSynchronize(c->cst));
        // 1. the running state (caller order), the contact list with its
        // springs, the wall histories, the bonds and the contact tables
        std::vector<double> fpos(3 * nf0), fvel(3 * nf0), facc(3 * nf0), fdrho(nf0), fsw(nf0);
        std::vector<double> spos(3 * ns0), svel(3 * ns0), sacc(3 * ns0), sw(3 * ns0), stq(3 * ns0),
            snrm(3 * ns0);
        std::vector<double> cpos(3 * nc0), cvel(3 * nc0), cacc(3 * nc0), cw(3 * nc0), ctq(3 * nc0),
            cq(4 * nc0);
        auto ok = [](int rc) {
            if (rc != DSPH_OK) throw ApiError{rc, "state read-back failed"};
        };
        if (nf0) ok(dsph_download_fluid(c, fpos.data(), fvel.data(), facc.data(), fdrho.data(), fsw.data()));
        if (ns0)
            ok(dsph_download_solid(c, spos.data(), svel.data(), sacc.data(), sw.data(), stq.data(), snrm.data()));
        if (nc0)
            ok(dsph_download_clump(c, cpos.data(), cvel.data(), cacc.data(), cw.data(), ctq.data(), cq.data()));
        int tss = 0;
        if (ns0) d2h(c, &tss, c->L.ssoff + ns0, 1);
        HIPCHK(hipStreamSynchronize(c->st));
        const int64_t nss = std::min<int64_t>(tss, c->L.ss_cap);
        std::vector<int> ssi(nss), ssj(nss);
        std::vector<double> ssS(9 * nss);
        d2h(c, ssi.data(), c->L.ss_i, nss);
        d2h(c, ssj.data(), c->L.ss_j, nss);
        d2h(c, ssS.data(), c->L.ssS, 9 * nss);
        const size_t nwh = 9 * (size_t)ns0 * c->nwmax, nyh = 9 * (size_t)ns0 * c->nymax;
        std::vector<double> wh(nwh), yh(nyh);
        d2h(c, wh.data(), c->whist, nwh);
        d2h(c, yh.data(), c->yhist, nyh);
        const size_t ncb = (size_t)c->tb.ncomb;
        std::vector<double> hz(8 * ncb), ln(11 * ncb), bd(6 * ncb);
        d2h(c, hz.data(), c->tb.hertz, 8 * ncb);
        d2h(c, ln.data(), c->tb.linear, 11 * ncb);
        d2h(c, bd.data(), c->tb.bonded, 6 * ncb);
        HostBonds old;
        old.get(c);
        // 2. the scene again: the existing particles' static arrays and
        // running state, then the added ones
        const Merged& H = *c->host;
        std::vector<double> Fpos, Fvel, Fh, Frho0, Fdrho, Fc, Fnu, Falpha;
        cat(Fpos, fpos, e->fluid_pos, nf1, 3, 0.0);
        cat(Fvel, fvel, e->fluid_vel, nf1, 3, 0.0);
        cat(Fh, H.fh, e->fluid_h, nf1, 1, 0.0);
        cat(Frho0, H.frho0, e->fluid_rho0, nf1, 1, 0.0);
        cat(Fdrho, fdrho, e->fluid_drho, nf1, 1, 0.0);
        cat(Fc, H.fc, e->fluid_c, nf1, 1, 0.0);
        cat(Fnu, H.fnu, e->fluid_nu, nf1, 1, 0.0);
        cat(Falpha, H.falpha, e->fluid_alpha, nf1, 1, 0.0);
        std::vector<double> Spos, Svel, Sw, Snrm, Srad, Seff, Sinvm;
        std::vector<int> Smat, Sclus, Sclump, Sdum;
        cat(Spos, spos, e->solid_pos, ns1, 3, 0.0);
        cat(Svel, svel, e->solid_vel, ns1, 3, 0.0);
        cat(Sw, sw, e->solid_angvel, ns1, 3, 0.0);
        cat(Snrm, snrm, e->solid_normal, ns1, 3, 0.0);
        cat(Srad, H.srad, e->solid_radius, ns1, 1, 0.0);
        cat(Seff, H.seff, e->solid_eff_radius, ns1, 1, 0.0);
        cat(Sinvm, H.sinvm, e->solid_inv_mass, ns1, 1, 0.0);
        cat(Smat, H.smat, e->solid_material, ns1, 1, 0);
        cat(Sclus, H.sclus, e->solid_cluster, ns1, 1, -1);
        cat(Sdum, H.sdum, e->solid_dummy, ns1, 1, 0);
        Sclump = H.sclump;
        for (int i = 0; i < ns1; i++) {
            const int k = e->solid_clump ? e->solid_clump[i] : -1;
            Sclump.push_back(k >= 0 ? k + nc0 : -1);
        }
        std::vector<double> Cpos, Cvel, Cw, Cq, CinvI, Cinvm;
        std::vector<int> Cstart = H.cstart, Cend = H.cend;
        cat(Cpos, cpos, e->clump_centroid, nc1, 3, 0.0);
        cat(Cvel, cvel, e->clump_vel, nc1, 3, 0.0);
        cat(Cw, cw, e->clump_angvel, nc1, 3, 0.0);
        Cq = cq;
        for (int k = 0; k < nc1; k++)
            for (int m = 0; m < 4; m++)
                Cq.push_back(e->clump_orientation ? e->clump_orientation[4 * k + m] : (m == 0 ? 1.0 : 0.0));
        cat(CinvI, H.cinvI, e->clump_inv_inertia, nc1, 6, 0.0);
        cat(Cinvm, H.cinvm, e->clump_inv_mass, nc1, 1, 0.0);
        for (int k = 0; k < nc1; k++) {
            Cstart.push_back(e->clump_pebble_start[k] + ns0);
            Cend.push_back(e->clump_pebble_end[k] + ns0);
        }
        dsph_scene S = H.flat;
        S.gravity[0] = c->gv[0];
        S.gravity[1] = c->gv[1];
        S.gravity[2] = c->gv[2];
        S.n_fluid = nf0 + nf1;
        S.fluid_pos = Fpos.data(); S.fluid_vel = Fvel.data(); S.fluid_h = Fh.data();
        S.fluid_rho0 = Frho0.data(); S.fluid_drho = Fdrho.data(); S.fluid_c = Fc.data();
        S.fluid_nu = Fnu.data(); S.fluid_alpha = Falpha.data();
        S.n_solid = ns0 + ns1;
        S.solid_pos = Spos.data(); S.solid_vel = Svel.data(); S.solid_angvel = Sw.data();
        S.solid_normal = Snrm.data(); S.solid_radius = Srad.data(); S.solid_eff_radius = Seff.data();
        S.solid_inv_mass = Sinvm.data(); S.solid_material = Smat.data(); S.solid_cluster = Sclus.data();
        S.solid_clump = Sclump.data(); S.solid_dummy = Sdum.data();
        S.n_clump = nc0 + nc1;
        S.clump_centroid = Cpos.data(); S.clump_vel = Cvel.data(); S.clump_angvel = Cw.data();
        S.clump_orientation = Cq.data(); S.clump_inv_inertia = CinvI.data(); S.clump_inv_mass = Cinvm.data();
        S.clump_pebble_start = Cstart.data(); S.clump_pebble_end = Cend.data();
        S.hertz = hz.data();
        S.linear = ln.data();
        S.bonded = bd.data();
        const int rc = dsph_upload_batch(c, &S, 1);
        if (rc != DSPH_OK) return rc;
        // 3. the contact list as it was, for the search to carry its springs
        // over by (i, j): a CSR by pointed solid (the new solids have none)
        if (nss > 0) {
            if (nss > c->L.ss_cap) alloc_ss(c, nss * 3 / 2 + 1024, false);
            std::vector<int64_t> ord(nss);
            for (int64_t k = 0; k < nss; k++) ord[k] = k;
            std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return ssi[a] < ssi[b]; });
            std::vector<int> cnt(c->ns, 0), off(c->ns + 1, 0), pi(nss), pj(nss);
            std::vector<double> pS(9 * nss);
            for (int64_t k = 0; k < nss; k++) {
                const int64_t o = ord[k];
                pi[k] = ssi[o];
                pj[k] = ssj[o];
                for (int m = 0; m < 9; m++) pS[9 * k + m] = ssS[9 * o + m];
                cnt[pi[k]]++;
            }
            for (int i = 0; i < c->ns; i++) off[i + 1] = off[i] + cnt[i];
            h2d(c, c->L.sscnt, cnt.data(), c->ns);
            h2d(c, c->L.ssoff, off.data(), c->ns + 1);
            h2d(c, c->L.ss_i, pi.data(), nss);
            h2d(c, c->L.ss_j, pj.data(), nss);
            h2d(c, c->L.ssS, pS.data(), 9 * nss);
            HIPCHK(hipStreamSynchronize(c->st));
        }
        if (nwh) h2d(c, c->whist, wh.data(), nwh);
        if (nyh) h2d(c, c->yhist, yh.data(), nyh);
        search_checked(c, 0, 1);
        // 4. bonds: the old ones as they were; with solids added, the
        // contacts that addBondData's later branch (dataBase.cpp:309-343)
        // bonds: pointed after the last bonded solid, touching, one cluster
        if (ns1 > 0) {
            int last = -1;
            for (int i : old.i) last = std::max(last, i);
            initial_bonds(c);
            HostBonds fresh;
            fresh.get(c);
            for (size_t k = 0; k < fresh.i.size(); k++)
                if (fresh.i[k] > last) old.append(fresh, k);
        }
        old.put(c);
        // 5. state the scene does not carry: accelerations, torques, sumW0
        // (the existing particles keep theirs, a new particle starts from
        // its kernel sum now)
        if (c->nf > 0) {
            const int nf = c->nf;
            FluidSet& f = c->fset[0];
            density_reinit(c->st, fluid_args(c), 1, c->mixed());
            std::vector<int> orig(nf);
            std::vector<double> w(nf);
            std::vector<double4> a(nf);
            d2h(c, orig.data(), f.orig, nf);
            d2h(c, w.data(), f.sumw0, nf);
            d2h(c, a.data(), f.acc, nf);
            HIPCHK(hipStreamSynchronize(c->st));
            for (int sl = 0; sl < nf; sl++) {
                const int i = orig[sl];
                if (i >= nf0) continue;
                w[sl] = fsw[i];
                a[sl].x = facc[3 * i];
                a[sl].y = facc[3 * i + 1];
                a[sl].z = facc[3 * i + 2];
            }
            h2d(c, f.sumw0, w.data(), nf);
            h2d(c, f.acc, a.data(), nf);
        }
        if (ns0) {
            std::vector<double> A(sacc), T(stq);
            A.resize(3 * (size_t)c->ns, 0.0);
            T.resize(3 * (size_t)c->ns, 0.0);
            put_xyz(c, c->so.acc, A, c->ns);
            put_xyz(c, c->so.tq, T, c->ns);
        }
        if (nc0) {
            std::vector<double> A(cacc), T(ctq);
            A.resize(3 * (size_t)c->nc, 0.0);
            T.resize(3 * (size_t)c->nc, 0.0);
            put_xyz(c, c->cl.acc, A, c->nc);
            put_xyz(c, c->cl.tq, T, c->nc);
        }
        HIPCHK(hipStreamSynchronize(c->st));
        invalidate_graphs(c);
        c->initialized = true;
        return DSPH_OK;
    } catch (const ApiError& ex) {
        return fail(c, ex);
    }
}

// ------------------------------------------------------------- read-back
int dsph_download_fluid(dsph_ctx* c, double* pos, double* vel, double* acc, double* drho,
                        double* sumw0) {
    API_GUARD(c);
    try {
        const int n = c->nf;
        const FluidSet& f = c->fset[0];
        std::vector<double4> ph(n), vr(n), ac(n);
        std::vector<double> dr(n), sw(n);
        std::vector<int> orig(n);
        d2h(c, ph.data(), f.ph, n);
        d2h(c, vr.data(), f.vr, n);
        d2h(c, ac.data(), f.acc, n);
        d2h(c, dr.data(), f.drho[c->dcur], n);
        d2h(c, sw.data(), f.sumw0, n);
        d2h(c, orig.data(), f.orig, n);
        HIPCHK(hipStreamSynchronize(c->st));
        for (int s = 0; s < n; s++) {
            const int i = orig[s];
            if (pos) { pos[3 * i] = ph[s].x; pos[3 * i + 1] = ph[s].y; pos[3 * i + 2] = ph[s].z; }
            if (vel) { vel[3 * i] = vr[s].x; vel[3 * i + 1] = vr[s].y; vel[3 * i + 2] = vr[s].z; }
            if (acc) { acc[3 * i] = ac[s].x; acc[3 * i + 1] = ac[s].y; acc[3 * i + 2] = ac[s].z; }
            if (drho) drho[i] = dr[s];
            if (sumw0) sumw0[i] = sw[s];
        }
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

static void put4(double* dst, const std::vector<double4>& v) {
    if (!dst) return;
    for (size_t i = 0; i < v.size(); i++) {
        dst[3 * i] = v[i].x;
        dst[3 * i + 1] = v[i].y;
        dst[3 * i + 2] = v[i].z;
    }
}

int dsph_download_solid(dsph_ctx* c, double* pos, double* vel, double* acc, double* angvel,
                        double* torque, double* normal) {
    API_GUARD(c);
    try {
        const int n = c->ns;
        std::vector<double4> a(n), b(n), d(n), e(n), f(n), g(n);
        d2h(c, a.data(), c->so.pr, n);
        d2h(c, b.data(), c->so.vel, n);
        d2h(c, d.data(), c->so.acc, n);
        d2h(c, e.data(), c->so.w, n);
        d2h(c, f.data(), c->so.tq, n);
        d2h(c, g.data(), c->so.nrm, n);
        HIPCHK(hipStreamSynchronize(c->st));
        put4(pos, a);
        put4(vel, b);
        put4(acc, d);
        put4(angvel, e);
        put4(torque, f);
        put4(normal, g);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_download_clump(dsph_ctx* c, double* centroid, double* vel, double* acc, double* angvel,
                        double* torque, double* orientation) {
    API_GUARD(c);
    try {
        const int n = c->nc;
        std::vector<double4> a(n), b(n), d(n), e(n), f(n), q(n);
        d2h(c, a.data(), c->cl.pos, n);
        d2h(c, b.data(), c->cl.vel, n);
        d2h(c, d.data(), c->cl.acc, n);
        d2h(c, e.data(), c->cl.w, n);
        d2h(c, f.data(), c->cl.tq, n);
        d2h(c, q.data(), c->cl.q, n);
        HIPCHK(hipStreamSynchronize(c->st));
        put4(centroid, a);
        put4(vel, b);
        put4(acc, d);
        put4(angvel, e);
        put4(torque, f);
        if (orientation)
            for (int k = 0; k < n; k++) {
                orientation[4 * k] = q[k].x;
                orientation[4 * k + 1] = q[k].y;
                orientation[4 * k + 2] = q[k].z;
                orientation[4 * k + 3] = q[k].w;
            }
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}


int dsph_download_fluid_solid_force(dsph_ctx* c, double* force, int64_t cap, int64_t* n_out) {
    API_GUARD(c);
    try {
        std::vector<int> fi, sa;
        std::vector<int64_t> ids = fs_reference_order(c, fi, sa, nullptr);
        if (n_out) *n_out = (int64_t)ids.size();
        if (!force) return DSPH_OK;
        if (cap < (int64_t)ids.size()) return DSPH_E_INVALID;
        int64_t tot = ids.size();
        std::vector<double4> F(tot);
        d2h(c, F.data(), c->L.fsF, tot);
        HIPCHK(hipStreamSynchronize(c->st));
        for (size_t k = 0; k < ids.size(); k++) {
            const double4 f = F[ids[k]];
            force[3 * k] = f.x;
            force[3 * k + 1] = f.y;
            force[3 * k + 2] = f.z;
        }
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_download_contacts(dsph_ctx* c, double* force, double* torque, double* sliding,
                           double* rolling, double* torsion, int64_t cap, int64_t* n_out) {
    API_GUARD(c);
    try {
        int t = 0;
        d2h(c, &t, c->L.ssoff + c->ns, 1);
        HIPCHK(hipStreamSynchronize(c->st));
        int64_t n = std::min<int64_t>(t, c->L.ss_cap);
        if (n_out) *n_out = n;
        if ((force || torque || sliding || rolling || torsion) && cap < n) return DSPH_E_INVALID;
        std::vector<double4> F(n), T(n);
        std::vector<double> S(9 * n);
        d2h(c, F.data(), c->L.ssF, n);
        d2h(c, T.data(), c->L.ssT, n);
        d2h(c, S.data(), c->L.ssS, 9 * n);
        HIPCHK(hipStreamSynchronize(c->st));
        put4(force, F);
        put4(torque, T);
        for (int64_t k = 0; k < n; k++)
            for (int m = 0; m < 3; m++) {
                if (sliding) sliding[3 * k + m] = S[9 * k + m];
                if (rolling) rolling[3 * k + m] = S[9 * k + 3 + m];
                if (torsion) torsion[3 * k + m] = S[9 * k + 6 + m];
            }
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_download_bonds(dsph_ctx* c, double* fn, double* tt, double* shear, double* bending,
                        int32_t* on, int64_t cap, int64_t* n_out) {
    API_GUARD(c);
    try {
        const int n = c->B.n;
        if (n_out) *n_out = n;
        if ((fn || tt || shear || bending || on) && cap < n) return DSPH_E_INVALID;
        std::vector<double> a(n), b(n);
        std::vector<double4> s(n), t(n);
        std::vector<int> o(n);
        d2h(c, a.data(), c->B.Fn, n);
        d2h(c, b.data(), c->B.Tt, n);
        d2h(c, s.data(), c->B.Fs, n);
        d2h(c, t.data(), c->B.Tb, n);
        d2h(c, o.data(), c->B.on, n);
        HIPCHK(hipStreamSynchronize(c->st));
        for (int k = 0; k < n; k++) {
            if (fn) fn[k] = a[k];
            if (tt) tt[k] = b[k];
            if (on) on[k] = o[k];
        }
        put4(shear, s);
        put4(bending, t);
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_download_list(dsph_ctx* c, int kind, int32_t* pointed, int32_t* pointing, int64_t cap,
                       int64_t* n_out, int32_t* count) {
    API_GUARD(c);
    try {
        std::vector<int> A, Bv, cnt;
        if (kind == DSPH_LIST_FLUID_FLUID) {
            const int nf = c->nf, KS = std::max(c->L.K, c->L.supK);
            const FluidHost fh = fluid_host(c);
            std::vector<int> nbr((size_t)KS * nf), cn(nf);
            if (nf > 0) {
                // exact rows from the supersets (and the walked rows) into a
                // scratch ELL of the superset's capacity
                int *rows = nullptr, *rcnt = nullptr;
                HIPCHK(hipMalloc(&rows, sizeof(int) * (size_t)KS * nf));
                HIPCHK(hipMalloc(&rcnt, sizeof(int) * (size_t)nf));
                fluid_exact_rows(c->st, search_args(c), rows, rcnt);
                d2h(c, nbr.data(), rows, (size_t)KS * nf);
                d2h(c, cn.data(), rcnt, nf);
                HIPCHK(hipStreamSynchronize(c->st));
                HIPCHK(hipFree(rows));
                HIPCHK(hipFree(rcnt));
            }
            cnt.assign(nf, 0);
            std::vector<RowEnt> row;
            for (int i = 0; i < nf; i++) {
                const int s2 = fh.slot_of[i];
                if (cn[s2] < 0) throw ApiError{DSPH_E_CAPACITY, "fluid-fluid list overflowed"};
                int H[27];
                const GridP& g = c->grids[c->h_fscene[i]];
                stencil_hashes(g, fh.ph[s2].x, fh.ph[s2].y, fh.ph[s2].z, H);
                row.clear();
                for (int k = 0; k < cn[s2]; k++) {
                    const int jo = fh.orig[nbr[(size_t)k * nf + s2]];
                    if (jo > i)  // the reference keeps the j > i half
                        row.push_back({(int)fh.fkey[jo], jo, 0, 0});
                }
                order_row(g, H, row);
                for (auto& e : row) {
                    A.push_back(i);
                    Bv.push_back(e.idx);
                }
                cnt[i] = (int)row.size();
            }
        } else if (kind == DSPH_LIST_FLUID_SOLID) {
            fs_reference_order(c, A, Bv, &cnt);
        } else if (kind == DSPH_LIST_SOLID_SOLID) {
            int t = 0;
            d2h(c, &t, c->L.ssoff + c->ns, 1);
            HIPCHK(hipStreamSynchronize(c->st));
            if (t > c->L.ss_cap) throw ApiError{DSPH_E_CAPACITY, "solid-solid list overflowed"};
            A.resize(t);
            Bv.resize(t);
            cnt.resize(c->ns);
            d2h(c, A.data(), c->L.ss_i, t);
            d2h(c, Bv.data(), c->L.ss_j, t);
            d2h(c, cnt.data(), c->L.sscnt, c->ns);
            HIPCHK(hipStreamSynchronize(c->st));
        } else if (kind == DSPH_LIST_BOND) {
            A.resize(c->B.n);
            Bv.resize(c->B.n);
            d2h(c, A.data(), c->B.i, c->B.n);
            d2h(c, Bv.data(), c->B.j, c->B.n);
            HIPCHK(hipStreamSynchronize(c->st));
        } else {
            return DSPH_E_INVALID;
        }
        const int64_t n = (int64_t)A.size();
        if (n_out) *n_out = n;
        if ((pointed || pointing) && cap < n) return DSPH_E_INVALID;
        for (int64_t k = 0; k < n; k++) {
            if (pointed) pointed[k] = A[k];
            if (pointing) pointing[k] = Bv[k];
        }
        if (count)
            for (size_t i = 0; i < cnt.size(); i++) count[i] = cnt[i];
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_batch_info(dsph_ctx* c, int32_t* n_scenes, int32_t* fluid_offset, int32_t* solid_offset,
                    int32_t* clump_offset, int32_t* wall_offset, int32_t* cell_offset) {
    API_GUARD(c);
    const int k = c->nscene;
    if (n_scenes) *n_scenes = k;
    for (int b = 0; b <= k; b++) {
        if (fluid_offset) fluid_offset[b] = c->foff[b];
        if (solid_offset) solid_offset[b] = c->soff[b];
        if (clump_offset) clump_offset[b] = c->coff[b];
        if (wall_offset) wall_offset[b] = c->woff[b];
        if (cell_offset) cell_offset[b] = b < k ? c->grids[b].cell0 : c->ncell;
    }
    return DSPH_OK;
}

int dsph_grid_info(dsph_ctx* c, int32_t* gs, double* cs, int32_t* n_cells) {
    API_GUARD(c);
    const GridP& g = c->grids[0];  // a batch: the first scene (dsph_batch_info)
    if (gs) { gs[0] = g.gx; gs[1] = g.gy; gs[2] = g.gz; }
    if (cs) { cs[0] = g.csx; cs[1] = g.csy; cs[2] = g.csz; }
    if (n_cells) *n_cells = g.ncell;
    return DSPH_OK;
}

int dsph_download_grid(dsph_ctx* c, int species, int32_t* hash, int32_t* perm, int32_t* start,
                       int32_t* end) {
    API_GUARD(c);
    if (species != 0 && species != 1) return DSPH_E_INVALID;
    try {
        // the device sorts only when a build or an exact walk needs it (and
        // the fluid by Morton cell rank); the reference's arrays of the last
        // search (stable sort by hash) are rebuilt from its hashes
        const int n = species == 0 ? c->nf : c->ns;
        std::vector<unsigned> key(n);
        d2h(c, key.data(), species == 0 ? c->fkey : c->skey, n);
        HIPCHK(hipStreamSynchronize(c->st));
        std::vector<int> cnt((size_t)c->ncell + 1, 0), p(n);
        for (int i = 0; i < n; i++) cnt[key[i] + 1]++;
        for (int h = 0; h < c->ncell; h++) cnt[h + 1] += cnt[h];
        for (int i = 0; i < n; i++) p[cnt[key[i]]++] = i;
        if (hash)
            for (int i = 0; i < n; i++) hash[i] = (int32_t)key[i];
        if (perm)
            for (int i = 0; i < n; i++) perm[i] = p[i];
        if (start) std::fill(start, start + c->ncell, -1);
        if (end) std::fill(end, end + c->ncell, -1);
        for (int t = 0; t < n; t++) {
            const unsigned k = key[p[t]];
            if (start && (t == 0 || key[p[t - 1]] != k)) start[k] = t;
            if (end && (t == n - 1 || key[p[t + 1]] != k)) end[k] = t + 1;
        }
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_status(dsph_ctx* c, int64_t words[DSPH_STATUS_WORDS]) {
    API_GUARD(c);
    try {
        unsigned long long w[ST_WORDS];
        read_status(c, w);
        for (int k = 0; k < ST_WORDS; k++) words[k] = (int64_t)w[k];
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_superset_stats(dsph_ctx* c, int64_t out[4]) {
    API_GUARD(c);
    if (!out) return DSPH_E_INVALID;
    if (!c->initialized) return DSPH_E_STATE;
    try {
        const int nf = c->nf;
        std::vector<int> cnt(nf), grow(nf), sup(nf);
        if (nf > 0) {
            HIPCHK(hipMemcpyAsync(cnt.data(), c->L.cnt, nf * sizeof(int), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipMemcpyAsync(grow.data(), c->L.grow, nf * sizeof(int), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipMemcpyAsync(sup.data(), c->L.sup_cnt, nf * sizeof(int), hipMemcpyDeviceToHost, c->st));
        }
        HIPCHK(hipStreamSynchronize(c->st));
        int64_t walked = 0, stored = 0, longest = 0;
        for (int s = 0; s < nf; s++) {
            int64_t m;
            if (cnt[s] >= 0) {  // walked on the reference stencil this step
                m = std::min(cnt[s], c->L.K);
                stored += m;
            } else {
                const int r = row_slot(grow[s]);
                m = std::min(sup[r], c->L.supK);
                if (r == s) stored += m;  // the group's row, counted once
            }
            walked += m;
            longest = std::max(longest, m);
        }
        out[0] = nf;
        out[1] = walked;
        out[2] = stored;
        out[3] = longest;
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_tile_stats(dsph_ctx* c, int64_t out[4]) {
    API_GUARD(c);
    if (!out) return DSPH_E_INVALID;
    if (!c->initialized) return DSPH_E_STATE;
    try {
        std::vector<int> u(c->T.n);
        if (c->T.n > 0) d2h(c, u.data(), c->T.ucnt, c->T.n);
        HIPCHK(hipStreamSynchronize(c->st));
        int64_t staged = 0, total = 0, largest = 0;
        for (int m : u) {
            if (m < 0) continue;
            staged++;
            total += m;
            largest = std::max<int64_t>(largest, m);
        }
        out[0] = c->T.n;
        out[1] = staged;
        out[2] = total;
        out[3] = largest;
        return DSPH_OK;
    } catch (const ApiError& e) {
        return fail(c, e);
    }
}

int dsph_set_timing(dsph_ctx* c, int enabled) {
    if (!c) return DSPH_E_INVALID;
    c->timing = enabled != 0;
    for (int f = 0; f < F_NFAM; f++) c->ev_used[f] = 0;
    return DSPH_OK;
}

int dsph_kernel_times(dsph_ctx* c, char* names, int64_t names_cap, double* ms, int32_t cap) {
    if (!c) return DSPH_E_INVALID;
    if (hipStreamSynchronize(c->st) != hipSuccess) return DSPH_E_HIP;
    if (c->cst && hipStreamSynchronize(c->cst) != hipSuccess) return DSPH_E_HIP;
    std::string all;
    int m = 0;
    for (int f = 0; f < F_NFAM && m < cap; f++) {
        double tot = 0;
        for (size_t k = 0; k < c->ev_used[f]; k++) {
            float t = 0;
            if (hipEventElapsedTime(&t, c->ev[f][k].first, c->ev[f][k].second) == hipSuccess) tot += t;
        }
        // per family: total ms and launch count
        ms[2 * m] = tot;
        ms[2 * m + 1] = (double)c->ev_used[f];
        all += kFamNames[f];
        all.push_back('\0');
        m++;
    }
    if (names && names_cap > 0) {
        int64_t n = std::min<int64_t>((int64_t)all.size(), names_cap);
        memcpy(names, all.data(), n);
    }
    for (int f = 0; f < F_NFAM; f++) c->ev_used[f] = 0;
    return m;
}

}  // extern "C"
