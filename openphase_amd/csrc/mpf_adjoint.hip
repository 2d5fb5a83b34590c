// Adjoint (vector-Jacobian product) of one multiphase-field step on MI355X.
//
// Forward step t -> t+1 per interface cell c: the merged list of c (own entries, entries appended
// from neighbours) with Laplacians L_e(c) = sum_k w_k phi_e(c+k) and gradients
// G_e(c) = sum_k g_k phi_e(c+k) (src/PhaseField.cpp:566-615) feeds the cell update: pair energies
// (InterfaceProperties::SetSR, src/InterfaceProperties.cpp:603-762), double-obstacle pair
// increments psi (src/DoubleObstacle.cpp:502-571) with the constant driving force
// (src/DrivingForce.cpp:497-602), the sum-to-one limiter (PhaseField::NormalizeIncrementsSR,
// src/PhaseField.cpp:1351-1567), the merge (:1268-1299) and NodePF::finalize
// (include/Containers/NodePF.h:214-255). Bulk cells keep their single entry.
//
// Given the cotangent of the values of S_{t+1}, three stages return the cotangent of the values
// of S_t and add the cotangents of the pair tables:
//  1. adjoint_local_kernel, one thread per cell: the cell update of a two-entry list (in
//     registers) is recomputed in FP64 and swept backwards by hand - finalize and merge, then the
//     increments psi (pair energies with the cubic anisotropy, triple-junction and driving-force
//     terms). Longer lists go to a work list that adjoint_list_kernel takes in dense wavefronts
//     (the list length C a compile-time constant), sweeping the sum-to-one limiter backwards
//     through its iterations, recomputed from the increments entering it; lists of five or more
//     entries (a few cells per step) go on to a second list and adjoint_long_kernel (five
//     entries in registers, six to eight in private memory), so that their long code holds up
//     only its own wavefronts.
//     Out: the direct part per own entry, the cotangents of each merged entry's Laplacian and
//     gradient, and the table cotangents - per block of cells, in thread order, and per listed
//     cell, folded into its block's sum in cell order by adjoint_fold_kernel.
//  2. adjoint_gather_kernel: the transpose of the stencils as a gather (Laplacian weights
//     symmetric, gradient weights antisymmetric in the offset):
//     lam_t(c, g) = direct(c, g) + sum_k w_k lamL(c+k, g) - sum_k g_k . lamG(c+k, g).
//  3. adjoint_grad_kernel: the per-block table cotangents summed over the blocks in a fixed order
//     and added to the caller's array. Results are the same from run to run.
// Derivatives follow the branches the forward step takes (clamps, erasures and limiter factors as
// evaluated); grain VolumeRatios are held fixed. FP32 states are differentiated at their values
// with FP64 arithmetic; S_{t+1}'s grain lists map the next-state cotangents to the entries.

#include <hip/hip_runtime.h>

#include "mpf_adjoint_cell.h"
#include "mpf_device.h"
#include "mpf_lists.h"

namespace mpf {

namespace {

constexpr int kAdjBlock = 64;      // one wavefront per block
enum AdjError { kAdjCapacity = 3 };  // status word: a cell needed more than kAdjSlots entries

constexpr int kAdjMid = 4;         // list capacity of the work-list kernel's register path
constexpr int kAdjListBlocks = 256;  // blocks of the work-list kernel (grid-stride)
constexpr int kAdjLongBlocks = 64;   // blocks of the second work-list kernel (grid-stride)

// Writes the adjoint outputs of a wide cell: merged entries with the cotangents of their
// Laplacians and gradients, and the direct part of the own entries.
template <int K, bool GRAD>
__device__ __forceinline__ void store_adjoint(const AdjointArgs& J, int64_t c, const CellList<K, double, GRAD>& L,
                                              const double (&lam_phi)[K], const double (&lam_lap)[K],
                                              const double (&lam_gr)[GRAD ? K : 1][3])
{
    const int64_t N = J.s.ncells;
    const int ne = L.ne;
    J.adj_cnt[c] = static_cast<uint8_t>(ne);
#pragma unroll
    for (int e = 0; e < K; ++e) {
        if (e >= ne) continue;
        J.adj_id[static_cast<int64_t>(e) * N + c] = static_cast<uint16_t>(L.eid[e]);
        J.adj_lap[static_cast<int64_t>(e) * N + c] = lam_lap[e];
        if (GRAD)
            for (int d = 0; d < 3; ++d) J.adj_grad[static_cast<int64_t>(3 * e + d) * N + c] = lam_gr[GRAD ? e : 0][d];
        if (e < L.n0) J.lam_own[static_cast<int64_t>(e) * N + c] = lam_phi[e];
    }
}

// Main local kernel: bulk cells and wide cells of two merged entries (lists in registers). Longer
// lists go to the work list (J.defer marks them by cell for the fold in cell order).
template <typename TV, bool GRAD>
__global__ __launch_bounds__(kAdjBlock) void adjoint_local_kernel(const AdjointArgs J)
{
    extern __shared__ double gsh[];  // [kAdjBlock][3 nphases^2] table cotangents per thread
    constexpr int K = 2;
    const StepArgs& A = J.s;
    const bool fp32 = sizeof(TV) == 4;
    const int G = 3 * A.nphases * A.nphases;
    double* gsl = gsh + static_cast<int>(threadIdx.x) * G;
    for (int e = 0; e < G; ++e) gsl[e] = 0.0;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
    if (c < A.ncells) {
        CellList<K, double, GRAD> L;
        const int r = gather_cell<K, TV, double, GRAD>(A, c, L);
        bool defer = r == kCellDeferred;
        if (r == kCellBulk) {
            // a bulk cell keeps its single entry
            J.adj_cnt[c] = 0;
            if (L.n0 == 1) J.lam_own[c] = lam_of(J, fp32, c, L.eid[0], 0);
        } else if (r == kCellWide) {
            double lam_phi[K], lam_lap[K], lam_gr[GRAD ? K : 1][3];
            cell_adjoint<2, GRAD, K>(J, fp32, c, L, lam_phi, lam_lap, lam_gr, gsl);
            store_adjoint<K, GRAD>(J, c, L, lam_phi, lam_lap, lam_gr);
        }
        J.defer[c] = defer ? 1 : 0;
        if (defer) J.list[atomicAdd(J.list_count, 1u)] = static_cast<uint32_t>(c);
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < G) {
        double s = 0.0;
        for (int t = 0; t < kAdjBlock; ++t) s += gsh[t * G + threadIdx.x];
        J.partial[static_cast<int64_t>(blockIdx.x) * G + threadIdx.x] = s;
    }
}

// One listed cell of at most kAdjMid merged entries (lists in registers): cotangents stored, its
// table cotangents to J.cellg[cell]. A longer list goes to the second work list.
template <typename TV, bool GRAD>
__device__ __forceinline__ void list_cell(const AdjointArgs& J, int64_t c, double* gsl)
{
    const StepArgs& A = J.s;
    const bool fp32 = sizeof(TV) == 4;
    const int G = 3 * A.nphases * A.nphases;
    constexpr int K = kAdjMid;
    CellList<K, double, GRAD> L;
    const int r = gather_cell<K, TV, double, GRAD>(A, c, L);
    if (r != kCellWide) {
        J.list_long[atomicAdd(J.long_count, 1u)] = static_cast<uint32_t>(c);
        return;
    }
    for (int e = 0; e < G; ++e) gsl[e] = 0.0;
    double lam_phi[K], lam_lap[K], lam_gr[GRAD ? K : 1][3];
    switch (L.ne) {
        case 2: cell_adjoint<2, GRAD, K>(J, fp32, c, L, lam_phi, lam_lap, lam_gr, gsl); break;
        case 3: cell_adjoint<3, GRAD, K>(J, fp32, c, L, lam_phi, lam_lap, lam_gr, gsl); break;
        default: cell_adjoint<4, GRAD, K>(J, fp32, c, L, lam_phi, lam_lap, lam_gr, gsl); break;
    }
    store_adjoint<K, GRAD>(J, c, L, lam_phi, lam_lap, lam_gr);
    for (int e = 0; e < G; ++e) J.cellg[c * G + e] = gsl[e];
}

// A cell of the second work list: five merged entries in registers (almost all of them), up to
// kAdjSlots in private memory; more is reported in the status word.
template <typename TV, bool GRAD>
__device__ __forceinline__ void long_cell(const AdjointArgs& J, int64_t c, double* gsl)
{
    const StepArgs& A = J.s;
    const bool fp32 = sizeof(TV) == 4;
    const int G = 3 * A.nphases * A.nphases;
    for (int e = 0; e < G; ++e) gsl[e] = 0.0;
    bool done = false;
    {
        constexpr int K = 5;
        CellList<K, double, GRAD> L;
        const int r = gather_cell<K, TV, double, GRAD>(A, c, L);
        if (r == kCellWide && L.ne == 5) {
            double lam_phi[K], lam_lap[K], lam_gr[GRAD ? K : 1][3];
            cell_adjoint<5, GRAD, K>(J, fp32, c, L, lam_phi, lam_lap, lam_gr, gsl);
            store_adjoint<K, GRAD>(J, c, L, lam_phi, lam_lap, lam_gr);
            done = true;
        }
    }
    if (!done) {
        constexpr int K = kAdjSlots;
        CellList<K, double, GRAD> L;
        const int r = gather_cell<K, TV, double, GRAD>(A, c, L);
        if (r != kCellWide) {
            atomicOr(&A.status[kAdjCapacity], 1);
            J.adj_cnt[c] = 0;
        } else {
            double lam_phi[K], lam_lap[K], lam_gr[GRAD ? K : 1][3];
            switch (L.ne) {
                case 6: cell_adjoint<6, GRAD, K>(J, fp32, c, L, lam_phi, lam_lap, lam_gr, gsl); break;
                case 7: cell_adjoint<7, GRAD, K>(J, fp32, c, L, lam_phi, lam_lap, lam_gr, gsl); break;
                default: cell_adjoint<8, GRAD, K>(J, fp32, c, L, lam_phi, lam_lap, lam_gr, gsl); break;
            }
            store_adjoint<K, GRAD>(J, c, L, lam_phi, lam_lap, lam_gr);
        }
    }
    for (int e = 0; e < G; ++e) J.cellg[c * G + e] = gsl[e];
}

// Work-list kernel: the cells of three or four merged entries in dense wavefronts.
template <typename TV, bool GRAD>
__global__ __launch_bounds__(kAdjBlock) void adjoint_list_kernel(const AdjointArgs J)
{
    extern __shared__ double gsh[];
    const int G = 3 * J.s.nphases * J.s.nphases;
    double* gsl = gsh + static_cast<int>(threadIdx.x) * G;
    const unsigned n = *J.list_count;
    for (unsigned i = blockIdx.x * kAdjBlock + threadIdx.x; i < n; i += gridDim.x * kAdjBlock)
        list_cell<TV, GRAD>(J, J.list[i], gsl);
}

// Second work-list kernel: the few cells of five or more merged entries.
template <typename TV, bool GRAD>
__global__ __launch_bounds__(kAdjBlock) void adjoint_long_kernel(const AdjointArgs J)
{
    extern __shared__ double gsh[];
    const int G = 3 * J.s.nphases * J.s.nphases;
    double* gsl = gsh + static_cast<int>(threadIdx.x) * G;
    const unsigned n = *J.long_count;
    for (unsigned i = blockIdx.x * kAdjBlock + threadIdx.x; i < n; i += gridDim.x * kAdjBlock)
        long_cell<TV, GRAD>(J, J.list_long[i], gsl);
}

// One wavefront per main-kernel block: adds the table cotangents of the block's listed cells to
// the block's partial sums, summed over the 64 cells by a fixed shuffle tree.
__global__ __launch_bounds__(64) void adjoint_fold_kernel(const AdjointArgs J)
{
    const int64_t b = blockIdx.x;
    const int G = 3 * J.s.nphases * J.s.nphases;
    const int64_t c = b * kAdjBlock + threadIdx.x;
    const bool listed = c < J.s.ncells && J.defer[c];
    if (!__any(listed)) return;
    for (int e = 0; e < G; ++e) {
        double v = listed ? J.cellg[c * G + e] : 0.0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (threadIdx.x == 0) J.partial[b * G + e] += v;
    }
}

template <typename TV, bool GRAD>
__global__ __launch_bounds__(kAdjBlock) void adjoint_gather_kernel(const AdjointArgs J)
{
    const StepArgs& A = J.s;
    const int64_t N = A.ncells;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
    if (c >= N) return;
    const int n0 = ListCode<TV>::count(A.cnt_in[c]);
    const Box B = make_box(A, c);
    for (int s = 0; s < A.slots; ++s) {
        double acc = 0.0;
        if (s < n0 && s < kAdjSlots) {
            const uint32_t g = A.id_in[static_cast<int64_t>(s) * N + c];
            acc = J.lam_own[static_cast<int64_t>(s) * N + c];
            for (int k = 0; k < 27; ++k) {
                if (!box_active(A, k)) continue;
                const int ii = k / 9 - 1, jj = (k / 3) % 3 - 1, kk = k % 3 - 1;
                const int nnz = (ii != 0) + (jj != 0) + (kk != 0);
                const int64_t nb = B.cell(k);
                const int m = min(static_cast<int>(J.adj_cnt[nb]), kAdjSlots);
                for (int e = 0; e < m; ++e) {
                    if (J.adj_id[static_cast<int64_t>(e) * N + nb] != g) continue;
                    acc += A.wl[nnz] * J.adj_lap[static_cast<int64_t>(e) * N + nb];
                    if (GRAD && nnz) {
                        const double gw = A.wg[nnz];
                        const double gx = ii == 0 ? 0.0 : (ii < 0 ? -gw : gw);
                        const double gy = jj == 0 ? 0.0 : (jj < 0 ? -gw : gw);
                        const double gz = kk == 0 ? 0.0 : (kk < 0 ? -gw : gw);
                        acc -= gx * J.adj_grad[static_cast<int64_t>(3 * e + 0) * N + nb] +
                               gy * J.adj_grad[static_cast<int64_t>(3 * e + 1) * N + nb] +
                               gz * J.adj_grad[static_cast<int64_t>(3 * e + 2) * N + nb];
                    }
                }
            }
        }
        J.lam_prev[static_cast<int64_t>(s) * N + c] = acc;
    }
}

__global__ void adjoint_reset_kernel(unsigned* count)
{
    if (threadIdx.x < 2) count[threadIdx.x] = 0u;
}

// One block per table entry e: grad[e] += sum over blocks of partial[block][e], in block order per
// thread, then a fixed tree.
__global__ __launch_bounds__(256) void adjoint_grad_kernel(const double* __restrict__ partial, int64_t nb, int G,
                                                           double* __restrict__ grad)
{
    __shared__ double lds[256];
    const int e = blockIdx.x;
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) v += partial[b * G + e];
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) lds[threadIdx.x] = lds[threadIdx.x] + lds[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[e] += lds[0];
}

template <typename TV, bool GRAD>
void launch_all(const AdjointArgs& a, hipStream_t s)
{
    const int64_t nb = adjoint_blocks(a.s.ncells);
    const int G = 3 * a.s.nphases * a.s.nphases;
    const dim3 grid(static_cast<unsigned>(nb));
    hipLaunchKernelGGL(adjoint_reset_kernel, dim3(1), dim3(64), 0, s, a.list_count);
    hipLaunchKernelGGL((adjoint_local_kernel<TV, GRAD>), grid, dim3(kAdjBlock), kAdjBlock * G * sizeof(double), s, a);
    hipLaunchKernelGGL((adjoint_list_kernel<TV, GRAD>), dim3(kAdjListBlocks), dim3(kAdjBlock),
                       kAdjBlock * G * sizeof(double), s, a);
    hipLaunchKernelGGL((adjoint_long_kernel<TV, GRAD>), dim3(kAdjLongBlocks), dim3(kAdjBlock),
                       kAdjBlock * G * sizeof(double), s, a);
    hipLaunchKernelGGL(adjoint_fold_kernel, grid, dim3(kAdjBlock), 0, s, a);
    hipLaunchKernelGGL((adjoint_gather_kernel<TV, GRAD>), grid, dim3(kAdjBlock), 0, s, a);
    hipLaunchKernelGGL(adjoint_grad_kernel, dim3(G), dim3(256), 0, s, a.partial, nb, G, a.grad);
}

}  // namespace

int64_t adjoint_blocks(int64_t ncells) { return (ncells + kAdjBlock - 1) / kAdjBlock; }

int launch_adjoint(const AdjointArgs& a, bool fp32, void* stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (fp32) {
        if (a.grad_energy) launch_all<float, true>(a, s);
        else launch_all<float, false>(a, s);
    } else {
        if (a.grad_energy) launch_all<double, true>(a, s);
        else launch_all<double, false>(a, s);
    }
    return static_cast<int>(hipGetLastError());
}

}  // namespace mpf
