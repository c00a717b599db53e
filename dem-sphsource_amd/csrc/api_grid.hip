// api_grid.hip — the hashed grid of one particle set on caller device arrays:
// dsph_hash_sort, the reference's updateGridCellStartEnd for one species
// (core/neighborSearch.cu:89-113) with calculatePointHash (:19-36), the
// stable key/value sort (sortKeyValuePairs, :3-9) and the cell ranges
// (findStartAndEnd, :46-87), for callers that hold their own particles.
//
// The work buffer (sorted keys and the radix sort's ping-pong and histogram
// buffers) is the context's own, grown on demand.  dsph_hash_sort_noalloc
// takes it from a fixed device array instead and allocates nothing.
#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/dsph.h"
#include "dsph_internal.h"

using namespace dsph;

// the members of dsph_ctx this file uses are reached through these accessors
// (api.hip owns the context layout)
namespace dsph {
int ctx_device(dsph_ctx* c);
hipStream_t ctx_stream(dsph_ctx* c);
void ctx_set_error(dsph_ctx* c, const char* msg);
// the context's hash-sort work buffer of at least `words` 32-bit words
unsigned* ctx_work(dsph_ctx* c, int64_t words);
}  // namespace dsph

namespace {

// the spatialGrid of dataBase::setSpatialGrids (dataBase.cpp:415-432) from
// its bounds and cell counts: cell size L / G per axis, a spare cell last
bool make_grid(const double origin[3], const double size[3], const int32_t cells[3], GridP& g) {
    for (int k = 0; k < 3; k++)
        if (!(size[k] > 0.0) || cells[k] < 1) return false;
    const long long n = (long long)cells[0] * cells[1] * cells[2] + 1;
    if (n >= (1LL << 30)) return false;
    g = GridP{};
    g.minx = origin[0];
    g.miny = origin[1];
    g.minz = origin[2];
    g.maxx = origin[0] + size[0];
    g.maxy = origin[1] + size[1];
    g.maxz = origin[2] + size[2];
    g.gx = cells[0];
    g.gy = cells[1];
    g.gz = cells[2];
    g.csx = size[0] / double(g.gx);
    g.csy = size[1] / double(g.gy);
    g.csz = size[2] / double(g.gz);
    g.ncell = (int)n;
    g.cell0 = 0;
    return true;
}

int key_bits_for(int ncell) {
    int b = 1;
    while ((1LL << b) < ncell) b++;
    return b;
}

// words of work buffer a set of n points needs: sorted keys, then the radix
// sort's buffers
int64_t work_words(int64_t n) { return n + radix_tmp_size(n); }

// hashes into `hash`, the stable permutation into `perm`, the cell ranges
void hash_sort_on(hipStream_t st, const GridP& g, const double* pos, int n, unsigned* work, int32_t* hash,
                  int32_t* perm, int32_t* cell_start, int32_t* cell_end) {
    hash_points(st, g, pos, n, (unsigned*)hash);
    unsigned* sorted = work;
    radix_sort_iota(st, (const unsigned*)hash, n, key_bits_for(g.ncell), sorted, perm, work + n);
    cell_spans(st, sorted, n, cell_start, cell_end, g.ncell, nullptr);
}

int check_args(dsph_ctx* c, const double* pos, int32_t n, const double* origin, const double* size,
               const int32_t* cells, int32_t* hash, int32_t* perm, int32_t* cs, int32_t* ce, GridP& g) {
    if (!c) return DSPH_E_INVALID;
    if (n < 0 || !origin || !size || !cells || !cs || !ce || (n > 0 && (!pos || !hash || !perm))) {
        ctx_set_error(c, "dsph_hash_sort: missing array or negative count");
        return DSPH_E_INVALID;
    }
    if (!make_grid(origin, size, cells, g)) {
        ctx_set_error(c, "dsph_hash_sort: grid needs a positive size and at least one cell per axis "
                         "(and fewer than 2^30 cells)");
        return DSPH_E_INVALID;
    }
    return DSPH_OK;
}

// ----------------------------------------------- allocation-free variant
// The scratch of dsph_hash_sort_noalloc: the sorted keys and the radix
// sort's buffers for up to kNoAllocMaxPoints points, a file-scope device
// array that every GPU the library runs on holds (its own copy) until the
// process ends.  Nothing in it needs to start at zero: each call writes the
// parts it reads (every histogram entry, every scan block sum).  Calls that
// overlap on one GPU share it and race.
constexpr int32_t kNoAllocMaxPoints = 1 << 20;
constexpr int64_t kScratchWords = 6 * (int64_t(1) << 20);  // 24 MiB
static __device__ __align__(256) unsigned scratch[kScratchWords];

// the address of `scratch` on each device, looked up once per device with
// that device current; host code passes these addresses, never the name
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned* g_scratch_ptr[kMaxDevices] = {};

unsigned* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace

namespace dsph {
// the scratch's address on `device`, which must be current; called by
// dsph_create for its device and, failing that, by the entry point
int noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return DSPH_E_NODEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return DSPH_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DSPH_E_HIP;
    g_scratch_ptr[device] = (unsigned*)p;
    return DSPH_OK;
}
}  // namespace dsph

extern "C" int dsph_hash_sort_noalloc(dsph_ctx* c, const double* pos, int32_t n, const double origin[3],
                                      const double size[3], const int32_t cells[3], int32_t* hash,
                                      int32_t* perm, int32_t* cell_start, int32_t* cell_end) {
    GridP g;
    const int rc = check_args(c, pos, n, origin, size, cells, hash, perm, cell_start, cell_end, g);
    if (rc != DSPH_OK) return rc;
    if (n > kNoAllocMaxPoints || work_words(std::max(n, 1)) > kScratchWords) {
        ctx_set_error(c, "dsph_hash_sort_noalloc: at most 1048576 points (use dsph_hash_sort)");
        return DSPH_E_INVALID;
    }
    const int dev = ctx_device(c);
    if (hipSetDevice(dev) != hipSuccess) return DSPH_E_HIP;
    unsigned* work = scratch_on(dev);
    if (!work) {
        if (noalloc_scratch_setup(dev) != DSPH_OK) {
            ctx_set_error(c, "dsph_hash_sort_noalloc: no address for the device scratch");
            return DSPH_E_HIP;
        }
        work = scratch_on(dev);
    }
    hash_sort_on(ctx_stream(c), g, pos, n, work, hash, perm, cell_start, cell_end);
    if (hipGetLastError() != hipSuccess) {
        ctx_set_error(c, "dsph_hash_sort_noalloc: kernel launch failed");
        return DSPH_E_HIP;
    }
    return DSPH_OK;
}

extern "C" int dsph_hash_sort(dsph_ctx* c, const double* pos, int32_t n, const double origin[3],
                              const double size[3], const int32_t cells[3], int32_t* hash, int32_t* perm,
                              int32_t* cell_start, int32_t* cell_end) {
    GridP g;
    const int rc = check_args(c, pos, n, origin, size, cells, hash, perm, cell_start, cell_end, g);
    if (rc != DSPH_OK) return rc;
    if (hipSetDevice(ctx_device(c)) != hipSuccess) return DSPH_E_HIP;
    unsigned* work = ctx_work(c, work_words(std::max(n, 1)));
    if (!work) return DSPH_E_NOMEM;
    hash_sort_on(ctx_stream(c), g, pos, n, work, hash, perm, cell_start, cell_end);
    if (hipGetLastError() != hipSuccess) {
        ctx_set_error(c, "dsph_hash_sort: kernel launch failed");
        return DSPH_E_HIP;
    }
    return DSPH_OK;
}
