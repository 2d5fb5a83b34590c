// MI355X (gfx950) kernels of the multiphase-field time step.
//
// One fused pass advances every cell by one step: it rebuilds the cell's neighbour-merged phase
// list with Laplacians (and gradients) from the 3^d neighbourhood (PhaseField::CalculateDerivativesSR,
// src/PhaseField.cpp:566-615, flags of SetFlagsSR :953-969), evaluates the interface properties
// (InterfaceProperties::SetSR, src/InterfaceProperties.cpp:603-762), the double-obstacle pair
// increments (DoubleObstacle::CalculatePhaseFieldIncrementsSR, src/DoubleObstacle.cpp:502-571), the
// optional constant driving force (src/UserDrivingForce.cpp:152-175, src/DrivingForce.cpp:257-274,
// 497-602), the sum-to-one limiter (PhaseField::NormalizeIncrementsSR, src/PhaseField.cpp:1351-1567),
// the merge (MergeIncrementsSR :1268-1299) and the finalize clamp (NodePF::finalize,
// include/Containers/NodePF.h:214-255), and writes the next state. Periodic halos are wrapped
// indices (BoundaryConditions::SetX/Y/Z, include/BoundaryConditions.h:694-712).
//
// Work per cell is dominated by the merged-list length n (2-4 in grain growth). The main kernel
// merges in registers with capacity 4 and compiles the pair/limiter/finalize part separately for
// n = 2, 3, 4, so a wavefront only executes the sizes its cells have. Cells with longer lists go
// to a work list for a second kernel (n = 5, 6 in registers) and from there to a third (up to 16).
// The neighbour lists are read with batched loads (all count bytes of the box, then the first two
// entries of each plane of nine neighbours), so a wavefront waits on a few memory round trips per
// cell rather than one per neighbour entry.
//
// Per-grain volume sums (PhaseField::CalculateGrainsVolume, src/PhaseField.cpp:812-951) are
// reduced per wavefront in a fixed lane order and added across wavefronts as 128-bit fixed point
// integers, so they do not depend on scheduling.
//
// Everything is built with -ffp-contract=off: the FP64 path performs the reference's operations
// in the reference's order, so active-phase sets match bit for bit.

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "mpf_device.h"
#include "mpf_lists.h"
#include "mpf_step_ops.h"

namespace mpf {

namespace {

// PhaseField::FinalizeInitialization: NodePF::finalize on flag-2 cells (more than one entry after
// a raw read) plus the volume sums; derivatives and flags are rebuilt by the next step kernel.
template <typename TV>
__global__ __launch_bounds__(kBlock) void finalize_init_kernel(const StepArgs A)
{
    constexpr int K = kWideSlots;
    const int64_t N = A.ncells;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool valid = c < N;
    const int64_t cc = valid ? c : 0;
    const uint8_t code = valid ? A.cnt_in[cc] : 0;
    const int n0 = ListCode<TV>::count(code);
    const int imp0 = ListCode<TV>::implicit(code);
    CellList<K, double, false> L;
    bool bad_volume = false;
    double others = 0.0;
    for (int t = 0; t < K; ++t) {
        L.eid[t] = 0u; L.phi[t] = 0.0; L.keep[t] = false;
        if (t < n0) {
            L.eid[t] = A.id_in[static_cast<int64_t>(t) * N + cc];
            if (t != imp0) {
                L.phi[t] = static_cast<double>(load_val<TV>(A.val_in, static_cast<int64_t>(t) * N + cc));
                others += L.phi[t];
            }
            L.keep[t] = true;
        }
    }
    if (imp0 >= 0 && imp0 < n0) L.phi[imp0] = 1.0 - others;
    if (valid && n0 > 1) {
        double total = 0.0;
        int ns = 0;
        for (int t = 0; t < n0; ++t) {
            double v = L.phi[t];
            bool erase = false;
            if (v >= 1.0 - DBL_EPSILON) v = 1.0;
            if (v <= 0.0 + DBL_EPSILON) { v = 0.0; erase = true; }
            total += v;
            L.phi[t] = v;
            L.keep[t] = !erase;
            ns += erase ? 0 : 1;
        }
        if (ns == 0) atomicOr(&A.status[kStatusEmpty], 1);
        else if (ns > 1) {
            const double norm = 1.0 / total;
            for (int t = 0; t < n0; ++t) if (L.keep[t]) L.phi[t] *= norm;
        } else {
            for (int t = 0; t < n0; ++t) if (L.keep[t]) L.phi[t] = 1.0;
        }
    }
    // the list was validated against the capacity on upload
    if (valid) store_cell<K, TV, double, false>(A, cc, L);
    for (int t = 0; t < K; ++t) wave_volume_add<true>(L.keep[t], L.eid[t], L.phi[t], A.vol_acc, bad_volume);
    if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
}

// Correctly rounded V * 2^-116 of the 4-limb sum.
__device__ double limbs_to_double(const unsigned long long l[kLimbs])
{
    // w = l0 + l1*2^32 + l2*2^64 + l3*2^96 as three 64-bit words
    unsigned long long w0 = 0, w1 = 0, w2 = 0;
    auto add_at = [&](unsigned long long v, int shift) {  // adds v * 2^shift, shift in {0,32,64,96}
        unsigned long long a0 = 0, a1 = 0, a2 = 0;
        if (shift == 0) { a0 = v; }
        else if (shift == 32) { a0 = v << 32; a1 = v >> 32; }
        else if (shift == 64) { a1 = v; }
        else { a1 = v << 32; a2 = v >> 32; }
        unsigned long long s0 = w0 + a0;
        unsigned long long c0 = s0 < w0;
        unsigned long long s1 = w1 + a1;
        unsigned long long c1 = s1 < w1;
        unsigned long long s1b = s1 + c0;
        c1 += s1b < s1;
        w0 = s0; w1 = s1b; w2 = w2 + a2 + c1;
    };
    add_at(l[0], 0); add_at(l[1], 32); add_at(l[2], 64); add_at(l[3], 96);
    if (w0 == 0 && w1 == 0 && w2 == 0) return 0.0;
    // 64-bit window from the top set bit; everything below folds into a sticky bit
    int top;
    unsigned long long hi, lo, rest;
    if (w2) { top = 2; hi = w2; lo = w1; rest = w0; }
    else if (w1) { top = 1; hi = w1; lo = w0; rest = 0; }
    else { top = 0; hi = w0; lo = 0; rest = 0; }
    const int lz = __clzll(static_cast<long long>(hi));
    unsigned long long m = lz ? ((hi << lz) | (lo >> (64 - lz))) : hi;
    unsigned long long below = lz ? (lo << lz) : lo;
    if (below | rest) m |= 1ull;  // round-to-odd sticky: the conversion below then rounds exactly
    const int exp2 = 64 * top + (64 - lz) - 64 - 116;  // value = m * 2^exp2
    return ldexp(static_cast<double>(m), exp2);
}

// PhaseField::CalculateGrainsVolume tail (src/PhaseField.cpp:893-951): MAXVolume, VolumeRatio,
// Exist/Stage; clears the accumulators and the deferred-cell counter for the next step.
__global__ void grains_kernel(const GrainArgs* __restrict__ batch, const uint32_t* __restrict__ first, int32_t ndom)
{
    const int dom = ndom > 1 ? batch_domain(first, ndom, blockIdx.x) : 0;
    const GrainArgs& G = batch[dom];
    const int g = static_cast<int>(blockIdx.x - first[dom]) * blockDim.x + threadIdx.x;
    if (g == 0) {
        G.stats[0] += G.work_count[0];
        G.stats[1] += G.work_count[1] + G.work_count[2];
        G.work_count[0] = 0u;
        G.work_count[1] = 0u;
        G.work_count[2] = 0u;
        G.work_count[3] = 0u;
        // this step's IP.Set maxima become the reported ones; the next step starts from 0
        G.obs[2] = G.obs[0];
        G.obs[3] = G.obs[1];
        G.obs[0] = 0ull;
        G.obs[1] = 0ull;
    }
    if (g >= G.ngrains) return;
    unsigned long long l[kLimbs];
#pragma unroll
    for (int i = 0; i < kLimbs; ++i) {
        l[i] = G.vol_acc[static_cast<int64_t>(g) * kLimbs + i];
        G.vol_acc[static_cast<int64_t>(g) * kLimbs + i] = 0ull;
    }
    double vol = limbs_to_double(l);
    double maxv = smax(vol, G.maxvolume[g]);
    double vr = maxv / G.refvolume[g];
    uint8_t exist = G.exist[g];
    int stage = G.stage[g];
    if (exist && vol <= 0.0) {
        exist = 0; stage = 2; vol = 0.0; maxv = 0.0; vr = 1.0;
    } else if (vol > 0.0) {
        exist = 1;
        if (stage == 0) stage = 1;
        if (vr > 1.0) { stage = 2; vr = 1.0; }
    }
    G.volume[g] = vol;
    G.maxvolume[g] = maxv;
    G.vr[g] = vr;
    G.exist[g] = exist;
    G.stage[g] = stage;
}

// Streaming copy (16-byte vectors, grid-stride): the bandwidth reference of the roofline.
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

int launch_copy(const void* src, void* dst, int64_t bytes, void* stream)
{
    const int64_t n = bytes / 16;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 256 * 64) blocks = 256 * 64;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
    return static_cast<int>(hipGetLastError());
}

// Two variants per precision: the plain isotropic step, and the full one (orientation-dependent
// energies and constant driving force; with ISO tables or no driving-force pairs it reduces to
// the plain arithmetic at run time). Each is compiled in its own translation unit
// (mpf_kernels_variant.hip).
int launch_step(const StepArgs& one, const StepArgs* args, const uint32_t* first, int ndom, unsigned blocks, bool fp32,
                bool full, bool uni, void* stream)
{
    if (fp32) return full ? v32_full::launch_step(one, args, first, ndom, blocks, uni, stream)
                          : v32_plain::launch_step(one, args, first, ndom, blocks, uni, stream);
    return full ? v64_full::launch_step(one, args, first, ndom, blocks, uni, stream)
                : v64_plain::launch_step(one, args, first, ndom, blocks, uni, stream);
}

int launch_deferred(const StepArgs* args, int ndom, int per_domain, bool fp32, bool full, void* stream)
{
    if (fp32) return full ? v32_full::launch_deferred(args, ndom, per_domain, stream)
                          : v32_plain::launch_deferred(args, ndom, per_domain, stream);
    return full ? v64_full::launch_deferred(args, ndom, per_domain, stream)
                : v64_plain::launch_deferred(args, ndom, per_domain, stream);
}

int launch_finalize_init(const StepArgs& a, bool fp32, void* stream)
{
    const dim3 grid(static_cast<unsigned>((a.ncells + kBlock - 1) / kBlock));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (fp32) hipLaunchKernelGGL((finalize_init_kernel<float>), grid, dim3(kBlock), 0, s, a);
    else      hipLaunchKernelGGL((finalize_init_kernel<double>), grid, dim3(kBlock), 0, s, a);
    return static_cast<int>(hipGetLastError());
}

int launch_grains(const GrainArgs* args, const uint32_t* first, int ndom, unsigned blocks, void* stream)
{
    hipLaunchKernelGGL(grains_kernel, dim3(blocks), dim3(kGrainBlock), 0, static_cast<hipStream_t>(stream), args, first, ndom);
    return static_cast<int>(hipGetLastError());
}

}  // namespace mpf
