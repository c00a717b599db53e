// k_frame.hip — the step loop's VTU frames, packed on the device.
//
// The reference's output (solverBase::outputFluidVTU, outputSolidVTU,
// outputFluid2SolidVTU, solverBase.cpp:69-333) copies the whole device state
// into its host mirrors and prints every value.  A frame here is packed by
// three kernels into one staging block of FP32 arrays (Int32 for indices) in
// the caller's order, which the host copies off the compute stream: 36 B per
// fluid particle, 60 B per solid and 32 B per fluid–solid pair, against
// 64–128 B of FP64 state per particle.
#include "dsph_internal.h"

namespace dsph {

namespace {

constexpr int kBS = 256;

// fluid: slot s goes to its caller index (the reference's order)
__global__ void __launch_bounds__(kBS) k_frame_fluid(FluidSet f, int nf, int cur, FrameOut o) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nf) return;
    const int i = f.orig[s];
    const double4 p = f.ph[s], v = f.vr[s];
    const double d = f.drho[cur][s];
    const double c = f.cn[s].x;
    o.fpos[3 * (int64_t)i] = (float)p.x;
    o.fpos[3 * (int64_t)i + 1] = (float)p.y;
    o.fpos[3 * (int64_t)i + 2] = (float)p.z;
    o.fh[i] = (float)p.w;
    o.fvel[3 * (int64_t)i] = (float)v.x;
    o.fvel[3 * (int64_t)i + 1] = (float)v.y;
    o.fvel[3 * (int64_t)i + 2] = (float)v.z;
    o.frho[i] = (float)(v.w + d);    // rho0 + dRho
    o.fp[i] = (float)(d * (c * c));  // dRho c^2
}

__global__ void __launch_bounds__(kBS) k_frame_solid(Solids so, int ns, FrameOut o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const double4 p = so.pr[i], v = so.vel[i], w = so.w[i], n = so.nrm[i];
    const int64_t k = 3 * (int64_t)i;
    o.spos[k] = (float)p.x;
    o.spos[k + 1] = (float)p.y;
    o.spos[k + 2] = (float)p.z;
    o.srad[i] = (float)p.w;
    o.sclus[i] = so.clus[i];
    o.sclump[i] = so.clump[i];
    o.svel[k] = (float)v.x;
    o.svel[k + 1] = (float)v.y;
    o.svel[k + 2] = (float)v.z;
    o.sw[k] = (float)w.x;
    o.sw[k + 1] = (float)w.y;
    o.sw[k + 2] = (float)w.z;
    o.snrm[k] = (float)n.x;
    o.snrm[k + 1] = (float)n.y;
    o.snrm[k + 2] = (float)n.z;
}

// fluid–solid pairs of the current list, packed as consecutive arrays of
// the n pairs (midpoint[3n], force[3n], fluid index[n], solid index[n]) so
// that the host copies exactly 32 n bytes; at most cap pairs fit (head[0] =
// n, or -1 when the list overflowed its storage)
__global__ void __launch_bounds__(kBS) k_frame_pairs(FluidSet f, Solids so, Lists L, int nf, int64_t cap,
                                                     FrameOut o) {
    const int64_t tot = nf > 0 && cap > 0 ? L.fsoff[nf] : 0;
    const int64_t n = tot <= cap ? tot : -1;
    if (blockIdx.x == 0 && threadIdx.x == 0) o.head[0] = n;
    float* const mid = o.pairs;
    float* const force = mid + 3 * (n > 0 ? n : 0);
    int* const fid = (int*)(force + 3 * (n > 0 ? n : 0));
    int* const sid = fid + (n > 0 ? n : 0);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int s = L.fs_own[k], a = L.fs_a[k];
        const double4 p = f.ph[s], q = so.pr[a];
        const double4 F = L.fsF[k];
        mid[3 * k] = (float)(0.5 * (p.x + q.x));
        mid[3 * k + 1] = (float)(0.5 * (p.y + q.y));
        mid[3 * k + 2] = (float)(0.5 * (p.z + q.z));
        force[3 * k] = (float)F.x;
        force[3 * k + 1] = (float)F.y;
        force[3 * k + 2] = (float)F.z;
        fid[k] = f.orig[s];
        sid[k] = a;
    }
}

}  // namespace

void frame_pack(hipStream_t s, unsigned species, const FluidSet& f, int nf, int cur, const Solids& so, int ns,
                const Lists& L, int64_t pair_cap, const FrameOut& o) {
    if ((species & DSPH_FRAME_FLUID) && nf > 0)
        hipLaunchKernelGGL(k_frame_fluid, dim3(blocks_for(nf)), dim3(kBS), 0, s, f, nf, cur, o);
    if ((species & DSPH_FRAME_SOLID) && ns > 0)
        hipLaunchKernelGGL(k_frame_solid, dim3(blocks_for(ns)), dim3(kBS), 0, s, so, ns, o);
    if (species & DSPH_FRAME_FLUID_SOLID) {
        const int64_t g = pair_cap > 0 ? (pair_cap + kBS - 1) / kBS : 1;
        hipLaunchKernelGGL(k_frame_pairs, dim3((unsigned)(g < 4096 ? g : 4096)), dim3(kBS), 0, s, f, so, L, nf,
                           pair_cap, o);
    }
}

}  // namespace dsph
