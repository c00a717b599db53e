// k_sort.hip — device primitives for the hashed grid: exclusive scan, stable
// LSD radix sort of (cell key, index) pairs and race-free cell start/end.
//
// Replaces thrust::sort_by_key / thrust::inclusive_scan and the
// setInitialIndices / setHashAux / findStartAndEnd kernels of the reference
// (core/neighborSearch.cu:3-69).  The sort is stable, so the permutation is
// the one a stable sort of the keys over the caller's index order gives —
// the reference's objectHash::index, bit for bit.
//
// Radix pass (8-bit digit, 1024-key tiles of 256 threads x 4 striped rounds):
//   hist    — per-tile digit histogram in LDS, stored digit-major;
//   scan    — exclusive scan of the digit-major histogram table = the output
//             offset of every (digit, tile);
//   scatter — each round ranks its 256 keys stably: wave64 ballots on the 8
//             digit bits give the peer mask (same digit) of every lane, the
//             rank inside the wave is popc(peers & lanes_below), per-wave
//             digit counts in LDS give the offset of the lower waves.
#include "dsph_internal.h"

namespace dsph {

namespace {

constexpr int kBS = 256;        // threads per block
constexpr int kTile = 1024;     // keys per radix tile / scan block

__global__ void k_fill_int(int* __restrict__ p, int64_t n, int v, Gate gate) {
    if (gate && *gate == 0) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// two arrays of n ints set to v in one launch
__global__ void k_fill_int2(int* __restrict__ p, int* __restrict__ q, int64_t n, int v, Gate gate) {
    if (gate && *gate == 0) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        p[i] = v;
        q[i] = v;
    }
}

__global__ void k_copy_int(int* __restrict__ d, const int* __restrict__ s, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

// inclusive scan of one value per thread across a workgroup (lds_wave: one
// int per wavefront)
__device__ __forceinline__ int block_inclusive_scan(int v, int* lds_wave) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    if (lane == 63) lds_wave[w] = v;
    __syncthreads();
    int add = 0;
    for (int k = 0; k < w; k++) add += lds_wave[k];
    __syncthreads();
    return v + add;
}

__global__ void __launch_bounds__(kBS) k_scan_reduce(const int* __restrict__ in, int64_t n,
                                                     int* __restrict__ bsum, Gate gate) {
    if (gate && *gate == 0) return;
    __shared__ int lw[4];
    int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x * 4;
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (base + k < n) s += in[base + k];
    int inc = block_inclusive_scan(s, lw);
    if (threadIdx.x == kBS - 1) bsum[blockIdx.x] = inc;
}

// exclusive scan of a tile plus the tile offset; the last block writes the
// grand total to out[n]
__global__ void __launch_bounds__(kBS) k_scan_final(const int* __restrict__ in, int64_t n,
                                                    const int* __restrict__ boff,
                                                    int* __restrict__ out, Gate gate) {
    if (gate && *gate == 0) return;
    __shared__ int lw[4];
    int64_t base = (int64_t)blockIdx.x * kTile + threadIdx.x * 4;
    int v[4];
    int s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    int inc = block_inclusive_scan(s, lw);
    int run = (boff ? boff[blockIdx.x] : 0) + inc - s;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBS - 1) out[n] = run;
}

// Exclusive scan of up to kScanOneMax ints in one workgroup; out[n] = the
// total.  Each wavefront takes a contiguous segment (a multiple of 256
// entries): it sums the segment with coalesced loads, the segment sums are
// scanned across the workgroup, and the wavefront then scans its segment
// 256 entries at a time (four consecutive entries per lane, a wavefront
// scan of the lanes' sums, a running carry).  One launch instead of the
// three of the tiled scan (a gated launch costs the same whether it works
// or exits).
constexpr int kScanOneBS = 1024;
constexpr int64_t kScanOneMax = 262144;
__global__ void __launch_bounds__(kScanOneBS) k_scan_one(const int* __restrict__ in, int64_t n,
                                                         int* __restrict__ out, Gate gate) {
    if (gate && *gate == 0) return;
    constexpr int W = kScanOneBS / 64;
    __shared__ int wsum[W];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t seg = ((n + W - 1) / W + 255) / 256 * 256;
    const int64_t b = (int64_t)w * seg;
    const int64_t e = b + seg < n ? b + seg : n;
    int sum = 0;
    for (int64_t i = b + lane; i < e; i += 64) sum += in[i];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if (lane == 0) wsum[w] = sum;
    __syncthreads();
    int carry = 0, total = 0;
    for (int k = 0; k < W; k++) {
        carry += k < w ? wsum[k] : 0;
        total += wsum[k];
    }
    for (int64_t i0 = b; i0 < e; i0 += 256) {
        const int64_t i = i0 + 4 * lane;
        int v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = i + k < e ? in[i + k] : 0;
        const int t = v[0] + v[1] + v[2] + v[3];
        int inc = t;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(inc, d, 64);
            if (lane >= d) inc += u;
        }
        int run = carry + inc - t;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (i + k < e) out[i + k] = run;
            run += v[k];
        }
        carry += __shfl(inc, 63, 64);
    }
    if (threadIdx.x == 0) out[n] = total;
}

__global__ void __launch_bounds__(kBS) k_rs_hist(const unsigned* __restrict__ kin, int64_t n,
                                                 int shift, int* __restrict__ hist, int ntiles,
                                                 Gate gate) {
    if (gate && *gate == 0) return;
    __shared__ int h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int64_t idx = base + r * kBS + t;
        if (idx < n) atomicAdd(&h[(kin[idx] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[(int64_t)t * ntiles + blockIdx.x] = h[t];
}

__global__ void __launch_bounds__(kBS) k_rs_scatter(const unsigned* __restrict__ kin,
                                                    const int* __restrict__ vin, int64_t n,
                                                    int shift, const int* __restrict__ off,
                                                    int ntiles, unsigned* __restrict__ kout,
                                                    int* __restrict__ vout, Gate gate) {
    if (gate && *gate == 0) return;
    __shared__ int run[256];
    __shared__ int wcnt[4][256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    run[t] = off[(int64_t)t * ntiles + blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < 4; r++) {
        const int64_t idx = base + r * kBS + t;
        const bool valid = idx < n;
        const unsigned key = valid ? kin[idx] : 0u;
        const unsigned d = (key >> shift) & 255u;
        wcnt[0][t] = 0; wcnt[1][t] = 0; wcnt[2][t] = 0; wcnt[3][t] = 0;
        __syncthreads();
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = run[d] + rank;
            for (int k = 0; k < w; k++) pos += wcnt[k][d];
            kout[pos] = key;
            vout[pos] = vin ? vin[idx] : (int)idx;
        }
        __syncthreads();
        run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();
    }
}

// findStartAndEnd without the end[] race of the reference (last key -> n)
__global__ void k_spans(const unsigned* __restrict__ sorted, int64_t n, int* __restrict__ start,
                        int* __restrict__ end, const unsigned* __restrict__ inv, Gate gate,
                        int shift) {
    if (gate && *gate == 0) return;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned k = sorted[i] >> shift;
    const unsigned cell = inv ? inv[k] : k;
    if (i == 0 || (sorted[i - 1] >> shift) != k) start[cell] = (int)i;
    if (i == n - 1 || (sorted[i + 1] >> shift) != k) end[cell] = (int)(i + 1);
}

// Exclusive scan over the listed indices only, in one workgroup of
// kKeyBS threads walking the list in tiles: keys ascending, every unlisted
// index has count 0.  off[keys[r]] = sum of counts of keys[0..r),
// off[n_all] = the total.  With cnt_out the counts are moved: read from cnt,
// written to cnt_out and cleared in cnt (so cnt is all zero again for the
// next count); with cursor, cursor[key] = off[key].
constexpr int kKeyBS = 1024;
__global__ void __launch_bounds__(kKeyBS) k_scan_keys(const int* __restrict__ keys, int nkeys,
                                                      int* __restrict__ cnt, int* __restrict__ cnt_out,
                                                      int* __restrict__ off, int* __restrict__ cursor,
                                                      int n_all) {
    __shared__ int lds[kKeyBS / 64];
    __shared__ int total;
    int carry = 0;
    for (int base = 0; base < nkeys; base += kKeyBS) {
        const int r = base + (int)threadIdx.x;
        int key = -1, v = 0;
        if (r < nkeys) {
            key = keys[r];
            v = cnt[key];
        }
        const int inc = block_inclusive_scan(v, lds);
        const int excl = carry + inc - v;
        if (r < nkeys) {
            off[key] = excl;
            if (cnt_out) {
                cnt_out[key] = v;
                cnt[key] = 0;
            }
            if (cursor) cursor[key] = excl;
        }
        // every thread learns the tile total from the last thread's sum
        if (threadIdx.x == kKeyBS - 1) total = inc;
        __syncthreads();
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) off[n_all] = carry;
}

}  // namespace

void fill_int(hipStream_t s, int* p, int64_t n, int v, Gate gate) {
    if (n <= 0) return;
    int b = blocks_for(n);
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_fill_int, dim3(b), dim3(kBS), 0, s, p, n, v, gate);
}

void copy_int(hipStream_t s, int* d, const int* src, int64_t n) {
    if (n <= 0) return;
    int b = blocks_for(n);
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_copy_int, dim3(b), dim3(kBS), 0, s, d, src, n);
}

int64_t scan_tmp_size(int64_t n) {
    int64_t tot = 0;
    while (n > kTile) {
        int64_t nb = (n + kTile - 1) / kTile;
        tot += 2 * nb + 2;
        n = nb;
    }
    return tot + 2;
}

void scan_keys(hipStream_t s, const int* keys, int nkeys, int* cnt, int* cnt_out, int* off, int* cursor,
               int n_all) {
    hipLaunchKernelGGL(k_scan_keys, dim3(1), dim3(kKeyBS), 0, s, keys, nkeys, cnt, cnt_out, off, cursor,
                       n_all);
}

void exclusive_scan(hipStream_t s, const int* in, int* out, int64_t n, int* tmp, Gate gate) {
    if (n <= 0) {
        fill_int(s, out, 1, 0, gate);
        return;
    }
    int64_t nb = (n + kTile - 1) / kTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_final, dim3(1), dim3(kBS), 0, s, in, n, (const int*)nullptr, out, gate);
        return;
    }
    int* bsum = tmp;
    int* boff = tmp + nb + 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kBS), 0, s, in, n, bsum, gate);
    exclusive_scan(s, bsum, boff, nb, tmp + 2 * nb + 2, gate);
    hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(kBS), 0, s, in, n, (const int*)boff, out,
                       gate);
}

int64_t radix_tmp_size(int64_t n) {
    int64_t nt = (n + kTile - 1) / kTile;
    int64_t hist = 256 * nt;
    // two key + two value ping-pong buffers, histogram, scanned histogram, scan tmp
    return 4 * n + hist + (hist + 1) + scan_tmp_size(hist) + 16;
}

void radix_sort_iota(hipStream_t s, const unsigned* keys_in, int64_t n, int bits,
                     unsigned* keys_out, int* perm_out, unsigned* tmp, Gate gate) {
    if (n <= 0) return;
    const int64_t nt = (n + kTile - 1) / kTile;
    unsigned* kb[2] = {tmp, tmp + n};
    int* vb[2] = {(int*)(tmp + 2 * n), (int*)(tmp + 3 * n)};
    int* hist = (int*)(tmp + 4 * n);
    int* hoff = hist + 256 * nt;
    int* stmp = hoff + 256 * nt + 1;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    for (int p = 0; p < passes; p++) {
        const unsigned* kin = p == 0 ? keys_in : kb[(p - 1) & 1];
        const int* vin = p == 0 ? nullptr : vb[(p - 1) & 1];
        unsigned* kout = p == passes - 1 ? keys_out : kb[p & 1];
        int* vout = p == passes - 1 ? perm_out : vb[p & 1];
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nt), dim3(kBS), 0, s, kin, n, 8 * p, hist, (int)nt,
                           gate);
        if (256 * nt <= kScanOneMax)
            hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanOneBS), 0, s, (const int*)hist, 256 * nt, hoff,
                               gate);
        else
            exclusive_scan(s, hist, hoff, 256 * nt, stmp, gate);
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nt), dim3(kBS), 0, s, kin, vin, n, 8 * p,
                           (const int*)hoff, (int)nt, kout, vout, gate);
    }
}

void cell_spans(hipStream_t s, const unsigned* sorted, int64_t n, int* start, int* end, int ncell,
                const unsigned* inv, Gate gate, int shift) {
    if (ncell > 0) {
        int b = blocks_for(ncell);
        if (b > 4096) b = 4096;
        hipLaunchKernelGGL(k_fill_int2, dim3(b), dim3(kBS), 0, s, start, end, (int64_t)ncell, -1, gate);
    }
    if (n <= 0) return;
    hipLaunchKernelGGL(k_spans, dim3(blocks_for(n)), dim3(kBS), 0, s, sorted, n, start, end, inv, gate,
                       shift);
}

}  // namespace dsph
