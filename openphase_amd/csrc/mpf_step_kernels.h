// Step and deferred-cell kernels of the per-cell path (templates), instantiated per precision and
// variant by mpf_kernels_variant.hip. See mpf_kernels.hip for what the step computes.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "mpf_device.h"
#include "mpf_lists.h"
#include "mpf_step_ops.h"

namespace mpf {

namespace {

// wave_volume_add for entries of exactly 1.0 (single-entry cells, most of the grid) by integer
// popcounts: 1.0 is 2^116 in the fixed-point sums, bit 20 of limb 3. Other values are converted
// to fixed point per lane and summed as integers, so every step kernel forms the same exact sum
// whatever the cell-to-lane mapping. Must be reached by whole wavefronts.
__device__ __forceinline__ void wave_volume_add_split(bool act, uint32_t g, double v,
                                                      unsigned long long* acc, bool& bad)
{
    const int lane = static_cast<int>(threadIdx.x) & (warpSize - 1);
    bool one = act && v == 1.0;
    while (true) {
        const unsigned long long m = __ballot(one);
        if (m == 0ull) break;
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        const uint32_t g0 = static_cast<uint32_t>(__shfl(static_cast<int>(g), leader));
        const bool sel = one && g == g0;
        const unsigned long long cells = static_cast<unsigned long long>(__popcll(__ballot(sel)));
        if (lane == leader) atomicAdd(&acc[static_cast<int64_t>(g0) * kLimbs + 3], cells << 20);
        one = one && !sel;
    }
    wave_volume_add<true>(act && v != 1.0, g, v, acc, bad);
}

// One block = kBlock consecutive cells. Phase 1 classifies every cell from the count bytes of its
// box: bulk cells (one entry, no multi-entry neighbour) keep their list and are copied through;
// wide cells are compacted in thread order into LDS. Phase 2 runs full wavefronts of wide cells
// through merge, pair increments, limiter, merge and finalize, so bulk lanes no longer ride along
// with interface lanes. Cells needing more than kMainSlots entries go to the mid kernel's list.
constexpr int kStepWaves = kBlock / 64;

template <typename TV, typename TC, bool GRAD, bool DG, bool UNI>
__device__ __forceinline__ void step_block(const StepArgs& A, int64_t base, uint32_t (&s_codes)[7][kBlock],
                                           uint16_t (&s_list)[kBlock], uint32_t (&s_wide)[kStepWaves])
{
    constexpr int K = kMainSlots;
    constexpr int kWaves = kStepWaves;
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int64_t c = base + tid;
    const bool valid = c < A.ncells;
    bool bad_volume = false;

    // ---- phase 1: classify; copy bulk cells
    bool wide = false, single = false, defer = false;
    uint32_t g = 0u;
    double v = 0.0;
    if (valid) {
        const uint8_t code = A.cnt_in[c];
        const int n0 = ListCode<TV>::count(code);
        const Box B = make_box(A, c);
        uint32_t codes[7];
        if (load_codes<TV>(A, B, code, codes)) {
            if (n0 > K) {
                defer = true;
            } else {
                wide = true;
#pragma unroll
                for (int w = 0; w < 7; ++w) s_codes[w][tid] = codes[w];
            }
        } else if (n0 == 1) {
            // NodePF::finalize keeps a lone entry at 1; an FP32 list stores it implicitly
            single = true;
            g = A.id_in[c];
            A.id_out[c] = static_cast<uint16_t>(g);
            if (ListCode<TV>::kImplicit) {
                v = 1.0;
                A.cnt_out[c] = 1;
            } else {
                v = static_cast<double>(load_val<TV>(A.val_in, c));
                static_cast<TV*>(A.val_out)[c] = static_cast<TV>(v);
                A.cnt_out[c] = code;
            }
        } else {
            A.cnt_out[c] = 0;  // empty node (reported when a step empties it)
        }
    }
    if (defer) {
        const unsigned slot = atomicAdd(&A.work_count[0], 1u);
        A.work_list[slot] = static_cast<uint32_t>(c);
    }
    wave_volume_add_split(single, g, v, A.vol_acc, bad_volume);

    // ---- compaction of the wide cells in thread order (deterministic)
    const unsigned long long wm = __ballot(wide);
    if (lane == 0) s_wide[wave] = static_cast<uint32_t>(__popcll(wm));
    __syncthreads();
    int offset = __popcll(wm & ((1ull << lane) - 1ull)), nwide = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const int n = static_cast<int>(s_wide[w]);
        offset += w < wave ? n : 0;
        nwide += n;
    }
    if (wide) s_list[offset] = static_cast<uint16_t>(tid);
    __syncthreads();
    if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
    bad_volume = false;
    if (wave * 64 >= nwide) return;  // whole wavefront; no barrier follows

    // ---- phase 2: wide cells
    CellList<K, TC, GRAD> L;
#pragma unroll
    for (int t = 0; t < K; ++t) { L.keep[t] = false; L.eid[t] = 0u; L.phi[t] = 0.0; }
    L.emax = 0.0;
    L.mmax = 0.0;
    if (tid < nwide) {
        const int local = s_list[tid];
        const int64_t cw = base + local;
        uint32_t codes[7];
#pragma unroll
        for (int w = 0; w < 7; ++w) codes[w] = s_codes[w][local];
        bool empty = false;
        load_own<K, TV, TC, GRAD>(A, cw, box_code(codes, 13), L);
        int r = merge_neighbours<K, TV, TC, GRAD, true>(A, make_box(A, cw), codes, L);
        if (r == kCellWide) {
            // merged lists of 2-4 entries here; longer ones (few cells) go to the mid kernel
            // UNI: the equal-energy form only; cells next to a grain below its reference volume
            // (VolumeRatio < 1) go to the mid kernel
            constexpr int PATH = UNI ? 1 : 2;
            bool done = false;
            switch (L.ne) {
            case 2: done = compute_cell<2, true, K, TC, GRAD, DG, PATH>(A, L, empty, cw); break;
            case 3: done = compute_cell<3, true, K, TC, GRAD, DG, PATH>(A, L, empty, cw); break;
            case 4: done = compute_cell<4, true, K, TC, GRAD, DG, PATH>(A, L, empty, cw); break;
            default: break;
            }
            if (!done) r = kCellDeferred;
        }
        if (r == kCellDeferred) {
#pragma unroll
            for (int t = 0; t < K; ++t) L.keep[t] = false;
            const unsigned slot = atomicAdd(&A.work_count[0], 1u);
            A.work_list[slot] = static_cast<uint32_t>(cw);
        } else if (!store_cell<K, TV, TC, GRAD>(A, cw, L)) {
            report_overflow<K, TC, GRAD>(A, cw, L);
        }
        if (empty) atomicOr(&A.status[kStatusEmpty], 1);
    }
    note_maxima(A.obs, L.emax, L.mmax);
#pragma unroll
    for (int t = 0; t < K; ++t) wave_volume_add_split(L.keep[t], L.eid[t], L.phi[t], A.vol_acc, bad_volume);
    if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
}

// One domain (ndom <= 1): arguments by value in the kernel-argument segment, which the hardware
// preloads into scalar registers. Several domains in one launch (ensembles): the argument table
// and block prefix sums in device memory; blocks are XCD-ordered over the whole launch first.
// Register budget: 6 wavefronts per SIMD (80 VGPRs, some spilled) for the plain variants, 3
// (168 VGPRs) for the orientation / driving-force ones, which spill heavily below that. Plain
// step kernel at 256^3 (FP64 / FP32): 3.25 / 2.79 ms at 5 wavefronts (87 VGPRs, no spill),
// 2.94 / 2.45 ms at 6, 3.27 / 2.54 ms at 7, 4.18 / 2.87 ms at 8: latency hiding outweighs spills.
template <typename TV, typename TC, bool GRAD, bool DG, bool UNI>
__global__ __launch_bounds__(kBlock, (GRAD || DG) ? 3 : 6) void step_kernel(const StepArgs A0, const StepArgs* __restrict__ batch,
                                                          const uint32_t* __restrict__ first, int32_t ndom)
{
    __shared__ uint32_t s_codes[7][kBlock];
    __shared__ uint16_t s_list[kBlock];
    __shared__ uint32_t s_wide[kStepWaves];
    if (ndom <= 1) {
        const unsigned gb = A0.xcd_order ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
        step_block<TV, TC, GRAD, DG, UNI>(A0, static_cast<int64_t>(gb) * kBlock, s_codes, s_list, s_wide);
    } else {
        const unsigned gb = batch[0].xcd_order ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
        const int dom = batch_domain(first, ndom, gb);
        step_block<TV, TC, GRAD, DG, UNI>(batch[dom], static_cast<int64_t>(gb - first[dom]) * kBlock, s_codes, s_list,
                                          s_wide);
    }
}

// Cells the main kernel deferred, with register-resident lists. LEVEL 0 (list 0: merged lists of
// 5-6 entries, more than 4 own entries, or 2-4 entries next to a grain below its reference
// volume) computes 2-6 entries with capacity 8 and sends 7-8 entries to list 1 and longer lists
// (capacity 8 exceeded) to list 2; LEVEL 1 computes 7 and 8 entries, LEVEL 2 (capacity 12) 9 and
// 10; both send anything else to list 3, the 16-entry kernel's. Lists 1 and 2 are independent, so
// one launch runs both levels (mid_tail_kernel) and their few long cells overlap. Splitting the
// sizes keeps the register count of the common level low and the few long lists off the
// private-memory path.
template <typename TV, typename TC, bool GRAD, bool DG, int LEVEL>
__device__ __forceinline__ void mid_level(const StepArgs& A, unsigned lblock, unsigned per)
{
    constexpr int K = LEVEL < 2 ? 8 : 12;
    const unsigned n = A.work_count[LEVEL];
    const uint32_t* list = A.work_list + LEVEL * A.ncells;
    const unsigned stride = per * kWideBlock;
    for (unsigned base = lblock * kWideBlock; base < n; base += stride) {
        const unsigned i = base + threadIdx.x;
        const bool valid = i < n;
        CellList<K, TC, GRAD> L;
        bool empty = false, bad_volume = false;
#pragma unroll
        for (int t = 0; t < K; ++t) { L.keep[t] = false; L.eid[t] = 0u; L.phi[t] = 0.0; }
        L.emax = 0.0;
        L.mmax = 0.0;
        if (valid) {
            const int64_t c = list[i];
            const int r = gather_cell<K, TV, TC, GRAD>(A, c, L);
            int dest = -1;  // work list the cell moves on to
            if (r == kCellWide) {
                if (LEVEL == 0) {
                    switch (L.ne) {
                    case 2: compute_cell<2, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    case 3: compute_cell<3, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    case 4: compute_cell<4, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    case 5: compute_cell<5, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    case 6: compute_cell<6, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    default: dest = 1; break;
                    }
                } else if (LEVEL == 1) {
                    switch (L.ne) {
                    case 7: compute_cell<7, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    case 8: compute_cell<8, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    default: dest = 3; break;
                    }
                } else {
                    switch (L.ne) {
                    case 9: compute_cell<9, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    case 10: compute_cell<10, true, K, TC, GRAD, DG>(A, L, empty, c); break;
                    default: dest = 3; break;
                    }
                }
            } else if (r == kCellDeferred) {
                dest = LEVEL == 0 ? 2 : 3;  // more entries than this level's capacity
            }
            if (dest >= 0) {
#pragma unroll
                for (int t = 0; t < K; ++t) L.keep[t] = false;
                const unsigned slot = atomicAdd(&A.work_count[dest], 1u);
                A.work_list[dest * A.ncells + slot] = static_cast<uint32_t>(c);
            } else if (!store_cell<K, TV, TC, GRAD>(A, c, L)) {
                report_overflow<K, TC, GRAD>(A, c, L);
            }
            if (empty) atomicOr(&A.status[kStatusEmpty], 1);
        }
        note_maxima(A.obs, L.emax, L.mmax);
#pragma unroll
        for (int t = 0; t < K; ++t) wave_volume_add<true>(L.keep[t], L.eid[t], L.phi[t], A.vol_acc, bad_volume);
        if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
    }
}

// gridDim.x / ndom blocks per domain, each domain's list walked grid-stride by its blocks. The
// plain variants run at 3 wavefronts per SIMD (168 VGPRs, no spill; level 0 at 256^3: 0.60 ms vs
// 0.71 ms at the unconstrained 219 VGPRs and 2 wavefronts).
template <typename TV, typename TC, bool GRAD, bool DG, int LEVEL>
__global__ __launch_bounds__(kWideBlock, (GRAD || DG) ? 1 : 3) void mid_kernel(const StepArgs* __restrict__ batch, int32_t ndom)
{
    const unsigned per = gridDim.x / static_cast<unsigned>(ndom);
    const unsigned dom = blockIdx.x / per, lblock = blockIdx.x - dom * per;
    mid_level<TV, TC, GRAD, DG, LEVEL>(batch[dom], lblock, per);
}

// Levels 1 and 2 in one launch: the first half of each domain's blocks walks list 1, the second
// half list 2.
template <typename TV, typename TC, bool GRAD, bool DG>
__global__ __launch_bounds__(kWideBlock) void mid_tail_kernel(const StepArgs* __restrict__ batch, int32_t ndom)
{
    const unsigned per = gridDim.x / static_cast<unsigned>(ndom);
    const unsigned dom = blockIdx.x / per, lblock = blockIdx.x - dom * per, half = per / 2;
    if (lblock < half) mid_level<TV, TC, GRAD, DG, 1>(batch[dom], lblock, half);
    else mid_level<TV, TC, GRAD, DG, 2>(batch[dom], lblock - half, per - half);
}

// Cells whose merged list is longer than 8 entries (up to 16, held in private memory).
template <typename TV, typename TC, bool GRAD, bool DG>
__global__ __launch_bounds__(kWideBlock) void wide_kernel(const StepArgs* __restrict__ batch, int32_t ndom)
{
    constexpr int K = kWideSlots;
    const unsigned per = gridDim.x / static_cast<unsigned>(ndom);
    const unsigned dom = blockIdx.x / per, lblock = blockIdx.x - dom * per;
    const StepArgs& A = batch[dom];
    const unsigned n = A.work_count[3];
    const unsigned stride = per * kWideBlock;
    for (unsigned base = lblock * kWideBlock; base < n; base += stride) {
        const unsigned i = base + threadIdx.x;
        const bool valid = i < n;
        CellList<K, TC, GRAD> L;
        bool empty = false, bad_volume = false;
        for (int t = 0; t < K; ++t) { L.keep[t] = false; L.eid[t] = 0u; L.phi[t] = 0.0; }
        L.emax = 0.0;
        L.mmax = 0.0;
        if (valid) {
            const int64_t c = A.work_list[3 * A.ncells + i];
            const int r = gather_cell<K, TV, TC, GRAD>(A, c, L);
            if (r == kCellWide) compute_cell<K, false, K, TC, GRAD, DG>(A, L, empty, c);
            if (r == kCellDeferred || !store_cell<K, TV, TC, GRAD>(A, c, L)) report_overflow<K, TC, GRAD>(A, c, L);
            if (empty) atomicOr(&A.status[kStatusEmpty], 1);
        }
        note_maxima(A.obs, L.emax, L.mmax);
        for (int t = 0; t < K; ++t) wave_volume_add<true>(L.keep[t], L.eid[t], L.phi[t], A.vol_acc, bad_volume);
        if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
    }
}

}  // namespace

}  // namespace mpf
