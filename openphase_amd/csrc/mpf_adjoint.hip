// Adjoint (vector-Jacobian product) of one FP64 multiphase-field step on MI355X.
//
// Forward step t -> t+1 per interface cell c: the merged list of c (own entries, entries appended
// from neighbours) with Laplacians L_e(c) = sum_k w_k phi_e(c+k) and gradients
// G_e(c) = sum_k g_k phi_e(c+k) (src/PhaseField.cpp:566-615) feeds the cell update of mpf_cell.h
// (pair energies, double-obstacle increments, driving force, limiter, merge, finalize). Bulk
// cells keep their single entry.
//
// Given the cotangent of the values of S_{t+1}, two kernels return the cotangent of the values of
// S_t and accumulate the cotangents of the pair tables:
//  1. adjoint_local_kernel: per interface cell, the cell update is re-evaluated with dual numbers
//     (eight directions per pass) over own values, Laplacians, gradients and the pair tables; the
//     derivatives are contracted with the next-state cotangent. Out: the direct part per own
//     entry and the cotangents of each merged entry's Laplacian and gradient.
//  2. adjoint_gather_kernel: the transpose of the stencils, as a gather (the Laplacian weights
//     are symmetric, the gradient weights antisymmetric in the offset):
//     lam_t(c, g) = direct(c, g) + sum_k w_k lamL(c+k, g) - sum_k g_k . lamG(c+k, g).
// Derivatives follow the branch the forward step takes (clamps, erasures and limiter factors as
// evaluated); grain VolumeRatios are held fixed.

#include <hip/hip_runtime.h>

#include "mpf_device.h"
#include "mpf_lists.h"
#include "mpf_cell.h"

namespace mpf {

namespace {

constexpr int kDirs = 8;          // dual directions per pass
constexpr int kAdjBlock = 64;

enum AdjError { kAdjCapacity = 3 };  // status word: a cell needed more than kAdjSlots entries

template <typename T>
__device__ void fill_tables(const AdjointArgs& J, CellTables<T>& P)
{
    const StepArgs& A = J.s;
    const int np2 = A.nphases * A.nphases;
    P.np = A.nphases;
    for (int k = 0; k < np2; ++k) {
        P.sigma[k] = T(A.sigma[k]);
        P.mu[k] = T(A.mu[k]);
        P.dg[k] = T(A.dg_value[k]);
    }
    P.eps = A.eps;
    P.emodel = A.emodel;
    P.dg_mode = A.dg_mode;
    P.dg_limit = J.dg_limit;
    P.regularization = J.regularization;
    P.prefactor = A.prefactor;
    P.dg_prefactor = A.dg_prefactor;
    P.dt = A.dt;
    P.consider_nucleus = A.consider_nucleus;
    P.tj_model = A.tj_model;
    P.tj_factor = A.tj_factor;
    P.tj_energy = A.tj_energy;
    P.dg_on = J.driving_force;
    P.grad = J.grad_energy;
}

template <typename T, typename L>
__device__ void fill_state(const StepArgs& A, const L& list, CellState<T>& S)
{
    S.ne = list.ne;
    S.n0 = list.n0;
    for (int e = 0; e < kCellMax; ++e) {
        const bool in = e < list.ne;
        const uint32_t g = in ? list.eid[e] : 0u;
        S.ph[e] = in ? A.gphase[g] : 0;
        S.vr[e] = in ? A.gvr[g] : 1.0;
        S.solid[e] = in ? A.gsolid[g] : 0;
        S.rot[e] = A.grot + 9 * static_cast<int64_t>(g);
        S.phi[e] = T(in ? list.phi[e] : 0.0);
        S.lap[e] = T(in ? list.lap[e] : 0.0);
        for (int d = 0; d < 3; ++d) S.gr[e][d] = T(0.0);
        S.keep[e] = false;
    }
}

// Direction index layout: [0, n0) own values, [n0, n0+ne) Laplacians, then 3 ne gradient parts
// (orientation-dependent energies only), then sigma, mu (and driving force) per pair-table entry.
template <bool GRAD>
__global__ __launch_bounds__(kAdjBlock) void adjoint_local_kernel(const AdjointArgs J)
{
    const StepArgs& A = J.s;
    const int64_t N = A.ncells;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
    if (c >= N) return;
    const uint8_t code = A.cnt_in[c];
    const int n0 = code;
    CellList<kAdjSlots, double, GRAD> L;
    if (load_own<kAdjSlots, double, double, GRAD>(A, c, code, L) == kCellDeferred) {
        atomicOr(&A.status[kAdjCapacity], 1);
        J.adj_cnt[c] = 0;
        return;
    }
    const Box B = make_box(A, c);
    uint32_t codes[7];
    if (!load_codes<double>(A, B, code, codes)) {
        // bulk cell: its single entry is copied
        J.adj_cnt[c] = 0;
        if (n0 == 1) J.lam_own[c] = J.lam_next[c];
        return;
    }
    if (merge_neighbours<kAdjSlots, double, double, GRAD, true>(A, B, codes, L) == kCellDeferred) {
        atomicOr(&A.status[kAdjCapacity], 1);
        J.adj_cnt[c] = 0;
        return;
    }
    const int ne = L.ne;

    // forward pass in FP64: kept entries and their slots in S_{t+1}
    double lam_out[kCellMax];
    {
        CellTables<double> P;
        fill_tables(J, P);
        CellState<double> S;
        fill_state(A, L, S);
        if (GRAD)
            for (int e = 0; e < ne; ++e)
                for (int d = 0; d < 3; ++d) S.gr[e][d] = L.gr[GRAD ? e : 0][d];
        cell_update(P, S);
        int w = 0;
        for (int e = 0; e < kCellMax; ++e) {
            lam_out[e] = 0.0;
            if (e < ne && S.keep[e]) {
                if (w < A.slots) lam_out[e] = J.lam_next[static_cast<int64_t>(w) * N + c];
                else atomicOr(&A.status[kAdjCapacity], 1);  // the forward step overflows here
                ++w;
            }
        }
    }

    const int np2 = A.nphases * A.nphases;
    const int nparam = np2 * (J.driving_force ? 3 : 2);
    const int ngrad = GRAD ? 3 * ne : 0;
    const int ndir = n0 + ne + ngrad + nparam;
    double lam_phi[kCellMax], lam_lap[kCellMax], lam_gr[GRAD ? kCellMax : 1][3];
    for (int e = 0; e < kCellMax; ++e) {
        lam_phi[e] = 0.0; lam_lap[e] = 0.0;
        if (GRAD) lam_gr[GRAD ? e : 0][0] = lam_gr[GRAD ? e : 0][1] = lam_gr[GRAD ? e : 0][2] = 0.0;
    }

    for (int base = 0; base < ndir; base += kDirs) {
        CellTables<Dual<kDirs>> P;
        fill_tables(J, P);
        CellState<Dual<kDirs>> S;
        fill_state(A, L, S);
        if (GRAD)
            for (int e = 0; e < ne; ++e)
                for (int d = 0; d < 3; ++d) S.gr[e][d] = Dual<kDirs>(L.gr[GRAD ? e : 0][d]);
        for (int i = 0; i < kDirs; ++i) {
            int dir = base + i;
            if (dir >= ndir) break;
            if (dir < n0) { S.phi[dir].d[i] = 1.0; continue; }
            dir -= n0;
            if (dir < ne) { S.lap[dir].d[i] = 1.0; continue; }
            dir -= ne;
            if (dir < ngrad) { S.gr[dir / 3][dir % 3].d[i] = 1.0; continue; }
            dir -= ngrad;
            if (dir < np2) { P.sigma[dir].d[i] = 1.0; continue; }
            dir -= np2;
            if (dir < np2) { P.mu[dir].d[i] = 1.0; continue; }
            dir -= np2;
            P.dg[dir].d[i] = 1.0;
        }
        cell_update(P, S);
        for (int i = 0; i < kDirs; ++i) {
            int dir = base + i;
            if (dir >= ndir) break;
            double g = 0.0;
            for (int e = 0; e < ne; ++e)
                if (S.keep[e]) g += lam_out[e] * S.phi[e].d[i];
            if (dir < n0) { lam_phi[dir] = g; continue; }
            dir -= n0;
            if (dir < ne) { lam_lap[dir] = g; continue; }
            dir -= ne;
            if (dir < ngrad) { if (GRAD) lam_gr[GRAD ? dir / 3 : 0][dir % 3] = g; continue; }
            dir -= ngrad;
            if (g != 0.0) atomicAdd(&J.grad[dir], g);  // sigma, mu, dg tables in that order
        }
    }

    J.adj_cnt[c] = static_cast<uint8_t>(ne);
    for (int e = 0; e < ne; ++e) {
        J.adj_id[static_cast<int64_t>(e) * N + c] = static_cast<uint16_t>(L.eid[e]);
        J.adj_lap[static_cast<int64_t>(e) * N + c] = lam_lap[e];
        if (GRAD)
            for (int d = 0; d < 3; ++d)
                J.adj_grad[static_cast<int64_t>(3 * e + d) * N + c] = lam_gr[GRAD ? e : 0][d];
    }
    for (int e = 0; e < n0; ++e) J.lam_own[static_cast<int64_t>(e) * N + c] = lam_phi[e];
}

template <bool GRAD>
__global__ __launch_bounds__(kAdjBlock) void adjoint_gather_kernel(const AdjointArgs J)
{
    const StepArgs& A = J.s;
    const int64_t N = A.ncells;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
    if (c >= N) return;
    const int n0 = A.cnt_in[c];
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

}  // namespace

int launch_adjoint(const AdjointArgs& a, void* stream)
{
    const dim3 grid(static_cast<unsigned>((a.s.ncells + kAdjBlock - 1) / kAdjBlock));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (a.grad_energy) {
        hipLaunchKernelGGL((adjoint_local_kernel<true>), grid, dim3(kAdjBlock), 0, s, a);
        hipLaunchKernelGGL((adjoint_gather_kernel<true>), grid, dim3(kAdjBlock), 0, s, a);
    } else {
        hipLaunchKernelGGL((adjoint_local_kernel<false>), grid, dim3(kAdjBlock), 0, s, a);
        hipLaunchKernelGGL((adjoint_gather_kernel<false>), grid, dim3(kAdjBlock), 0, s, a);
    }
    return static_cast<int>(hipGetLastError());
}

}  // namespace mpf
