// Quantities derived from the state on demand: per-cell phase fractions
// (PhaseField::CalculateFractions, src/PhaseField.cpp:411-423), and the interface energy of the
// double-obstacle functional summed over the grid:
// DoubleObstacle::PointEnergy (src/DoubleObstacle.cpp:903-923) per cell, added up by
// DoubleObstacle::Energy / AverageEnergyDensity (:953-967, :989-1005).
//
// Per wide-interface cell of the current state, over the pairs alpha < beta of its
// neighbour-merged list (zero-valued appended entries included), 4 E (phi_a phi_b - Eta^2/Pi^2
// grad_a . grad_b) / Eta. The gradients are those of the current lists
// (PhaseField::CalculateDerivativesSR, src/PhaseField.cpp:588-603); E is the energy
// InterfaceProperties::Properties holds for the pair, i.e. what the last IP.Set evaluated on the
// state it saw (`ip_*`, the step-start state when IP.Set ran before the last step): 0 for a pair
// absent from that cell's list then, or for a cell that was not a wide interface
// (NodeIP::get_energy, include/Containers/NodeIP.h:379-388).
//
// One thread per cell; each block sums its cells in a fixed tree order, and one block adds the
// partial sums in a fixed order, so the result is the same from run to run (it differs from the
// reference's sequential sum only in rounding).
#include <hip/hip_runtime.h>

#include "mpf_device.h"
#include "mpf_lists.h"
#include "mpf_step_ops.h"

namespace mpf {

namespace {

constexpr int kEnergyBlock = 256;
constexpr int K = kWideSlots;   // merged entries per cell (the step's own limit)

__device__ __forceinline__ double block_sum(double v, double* lds)
{
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int s = kEnergyBlock / 2; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) lds[threadIdx.x] = lds[threadIdx.x] + lds[threadIdx.x + s];
        __syncthreads();
    }
    return lds[0];
}

template <typename TV>
__global__ __launch_bounds__(kEnergyBlock) void energy_kernel(const EnergyArgs E)
{
    __shared__ double lds[kEnergyBlock];
    const StepArgs& A = E.cur;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kEnergyBlock + threadIdx.x;
    double e = 0.0;
    if (c < A.ncells && E.ip_cnt != nullptr) {
        CellList<K, double, true> L;
        for (int t = 0; t < K; ++t) L.eid[t] = 0u;
        const int r = gather_cell<K, TV, double, true>(A, c, L);
        if (r == kCellDeferred) {
            atomicOr(E.status, 1);
        } else if (r == kCellWide && L.ne > 1) {
            StepArgs B = A;
            B.cnt_in = E.ip_cnt;
            B.id_in = E.ip_id;
            B.val_in = E.ip_val;
            CellList<K, double, true> P;
            for (int t = 0; t < K; ++t) P.eid[t] = 0u;
            const int rp = gather_cell<K, TV, double, true>(B, c, P);
            if (rp == kCellDeferred) atomicOr(E.status, 1);
            const bool ip_wide = rp == kCellWide;
            const int NP = A.nphases;
            for (int a = 0; a < L.ne; ++a)
                for (int b = a + 1; b < L.ne; ++b) {
                    double en = 0.0;
                    if (ip_wide) {
                        int pa = -1, pb = -1;
                        for (int t = 0; t < P.ne; ++t) {
                            if (P.eid[t] == L.eid[a]) pa = t;
                            if (P.eid[t] == L.eid[b]) pb = t;
                        }
                        if (pa >= 0 && pb >= 0) {
                            if (pa > pb) { const int q = pa; pa = pb; pb = q; }
                            const int k = A.gphase[P.eid[pa]] * NP + A.gphase[P.eid[pb]];
                            en = ip_pair_energy<double>(A, k, P.eid[pa], P.eid[pb], P.phi[pa], P.phi[pb], P.gr[pa], P.gr[pb]);
                        }
                    }
                    const double dot = L.gr[a][0] * L.gr[b][0] + L.gr[a][1] * L.gr[b][1] + L.gr[a][2] * L.gr[b][2];
                    e += 4.0 * en * (L.phi[a] * L.phi[b] - E.pref * dot) / E.eta;
                }
        }
    }
    const double s = block_sum(e, lds);
    if (threadIdx.x == 0) E.partial[blockIdx.x] = s;
}

// Adds the per-block sums in a fixed order: thread t sums blocks t, t + 256, ... in order, then the
// tree of block_sum.
__global__ __launch_bounds__(kEnergyBlock) void energy_total_kernel(const double* __restrict__ partial, int64_t n,
                                                                    double* __restrict__ out)
{
    __shared__ double lds[kEnergyBlock];
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kEnergyBlock) v += partial[i];
    const double s = block_sum(v, lds);
    if (threadIdx.x == 0) *out = s;
}

// One thread per cell: for every thermodynamic phase the sum of its grains' values in list order
// (the FP32 implicit entry as 1 - the others), out[phase * ncells + cell].
template <typename TV>
__global__ __launch_bounds__(kEnergyBlock) void fractions_kernel(const StepArgs A, double* __restrict__ out)
{
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kEnergyBlock + threadIdx.x;
    if (c >= A.ncells) return;
    CellList<kMaxSlots, double, false> L;
    load_own<kMaxSlots, TV, double, false>(A, c, A.cnt_in[c], L);
    for (int n = 0; n < A.nphases; ++n) {
        double f = 0.0;
        for (int t = 0; t < L.n0; ++t)
            if (A.gphase[L.eid[t]] == n) f += L.phi[t];
        out[static_cast<int64_t>(n) * A.ncells + c] = f;
    }
}

}  // namespace

int launch_fractions(const StepArgs& a, bool fp32, double* out, void* stream)
{
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nb = static_cast<unsigned>((a.ncells + kEnergyBlock - 1) / kEnergyBlock);
    if (fp32)
        hipLaunchKernelGGL(fractions_kernel<float>, dim3(nb), dim3(kEnergyBlock), 0, s, a, out);
    else
        hipLaunchKernelGGL(fractions_kernel<double>, dim3(nb), dim3(kEnergyBlock), 0, s, a, out);
    return static_cast<int>(hipGetLastError());
}

int64_t energy_blocks(int64_t ncells) { return (ncells + kEnergyBlock - 1) / kEnergyBlock; }

int launch_energy(const EnergyArgs& a, bool fp32, double* out, void* stream)
{
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nb = energy_blocks(a.cur.ncells);
    if (fp32)
        hipLaunchKernelGGL(energy_kernel<float>, dim3(static_cast<unsigned>(nb)), dim3(kEnergyBlock), 0, s, a);
    else
        hipLaunchKernelGGL(energy_kernel<double>, dim3(static_cast<unsigned>(nb)), dim3(kEnergyBlock), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return static_cast<int>(e);
    hipLaunchKernelGGL(energy_total_kernel, dim3(1), dim3(kEnergyBlock), 0, s, a.partial, nb, out);
    return static_cast<int>(hipGetLastError());
}

}  // namespace mpf
