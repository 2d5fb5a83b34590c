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

namespace mpf {

namespace {

// Driving force of pair (a, b) before the limiter: the value averaged over the interface for this
// step when averaging is on (own entries only: a pair with a zero-valued appended entry weighs 0),
// else the phase-pair value (DrivingForce::SkipAverage).
template <typename TC>
__device__ __forceinline__ TC dg_local(const StepArgs& A, int a, int b, int n0, int64_t cell, int k)
{
    if (A.dg_avg != nullptr && b < n0)
        return static_cast<TC>(A.dg_avg[static_cast<int64_t>(b * (b - 1) / 2 + a) * A.ncells + cell]);
    return static_cast<TC>(A.dg_value[k]);
}

// IP.Set + increments + normalize + merge + finalize of a wide cell whose merged list has C
// entries (EXACT) or at most C entries (ne at run time). Only entries < C are touched.
// PATH selects the increment formula at compile time: 0 both (chosen per cell), 1 only the
// equal-energy form (returns false, doing nothing, for a cell that needs the general one), 2 only
// the general form (exact for every cell: the equal-energy form drops exact zero terms).
template <int C, bool EXACT, int K, typename TC, bool GRAD, bool DG, int PATH = 0>
__device__ __forceinline__ bool compute_cell(const StepArgs& A, CellList<K, TC, GRAD>& L, bool& empty, int64_t cell)
{
    constexpr int U = C <= kFastSlots ? C : 1;
    constexpr int NPC = C * (C - 1) / 2;
    const int ne = EXACT ? C : L.ne;
    const int n0 = L.n0;
    const int NP = A.nphases;

    // ---- per-entry grain data
    int ph[C];
    double vr[C];
#pragma unroll U
    for (int t = 0; t < C; ++t) {
        ph[t] = 0; vr[t] = 0.0;
        if (EXACT || t < ne) { ph[t] = A.gphase[L.eid[t]]; vr[t] = A.gvr[L.eid[t]]; }
    }
    double localn = 0.0;
    if (A.consider_nucleus) {
#pragma unroll U
        for (int t = 0; t < C; ++t) if (EXACT || t < ne) localn += vr[t];
    } else {
        localn = static_cast<double>(ne);
    }
    const double norm_1 = 1.0 / localn;
    bool uni = A.uniform != 0;
#pragma unroll U
    for (int t = 0; t < C; ++t) uni = uni && (!(EXACT || t < ne) || vr[t] == 1.0);
    if (PATH == 1 && !uni) return false;
    if (PATH == 2) uni = false;

    // ---- double-obstacle pair increments
    const TC P = static_cast<TC>(A.prefactor);
    double psi[NPC];
#pragma unroll U
    for (int p = 0; p < NPC; ++p) psi[p] = 0.0;
    if (uni) {
        // equal energies and VolumeRatio 1: scale = 1, cf = 0.5*(1+1)*P = P and every gamma term
        // is (sigma - sigma) * (...) = 0, so d = sigma*(X_a - X_b), X = L + P*phi
        const TC s0 = static_cast<TC>(A.sigma0);
        const TC mf = static_cast<TC>(A.mu0) * static_cast<TC>(norm_1);
        if (ne > 1) {
            L.emax = static_cast<double>(s0);
            L.mmax = A.mu0;
        }
        TC X[C];
#pragma unroll U
        for (int t = 0; t < C; ++t) X[t] = L.lap[t] + P * L.ev(t);
#pragma unroll U
        for (int a = 0; a < C; ++a)
#pragma unroll U
        for (int b = 0; b < C; ++b) {
            if (b <= a || !(EXACT || b < ne)) continue;
            TC d = s0 * (X[a] - X[b]);
            d *= mf;
            if (DG && n0 > 1 && A.dg_mode[ph[a] * NP + ph[b]] == 1) {
                const TC raw = dg_local<TC>(A, a, b, n0, cell, ph[a] * NP + ph[b]);
                const TC weight = sqrt(L.ev(a) * L.ev(b));
                const TC allowed = static_cast<TC>(A.dg_allowed[ph[a] * NP + ph[b]]);
                const TC dgl = allowed * tanh(raw / allowed);
                d += dgl * static_cast<TC>(A.mu0) * weight * static_cast<TC>(A.dg_prefactor);
            }
            psi[pair_index<C>(a, b)] = static_cast<double>(d);
        }
    } else {
        // interface energies of the pairs (ISO: phase-pair table; CUBIC: orientation dependent)
        TC sig[GRAD ? NPC : 1];
        if (GRAD) {
#pragma unroll U
            for (int a = 0; a < C; ++a)
#pragma unroll U
            for (int b = 0; b < C; ++b) {
                if (b <= a || !(EXACT || b < ne)) continue;
                const int k = ph[a] * NP + ph[b];
                TC en = TC(0);
                if (A.emodel[k] == 0) {
                    en = static_cast<TC>(A.sigma[k]);
                } else {
                    const TC E = static_cast<TC>(A.sigma[k]), e1 = static_cast<TC>(A.eps[k]);
                    TC wa = TC(0), wb = TC(0);
                    if (A.gsolid[L.eid[a]]) {
                        wa = fabs(L.ev(a) * (TC(1.0) - L.ev(a)));
                        if (wa > TC(DBL_EPSILON))
                            en += cubic_energy<TC>(A.grot + 9 * static_cast<int64_t>(L.eid[a]), L.gr[GRAD ? a : 0][0],
                                                   L.gr[GRAD ? a : 0][1], L.gr[GRAD ? a : 0][2], E, e1) * wa;
                        else
                            wa = TC(0);
                    }
                    if (A.gsolid[L.eid[b]]) {
                        wb = fabs(L.ev(b) * (TC(1.0) - L.ev(b)));
                        if (wb > TC(DBL_EPSILON))
                            en += cubic_energy<TC>(A.grot + 9 * static_cast<int64_t>(L.eid[b]), L.gr[GRAD ? b : 0][0],
                                                   L.gr[GRAD ? b : 0][1], L.gr[GRAD ? b : 0][2], E, e1) * wb;
                        else
                            wb = TC(0);
                    }
                    if (wa + wb > TC(DBL_EPSILON)) en /= wa + wb;
                    else en = static_cast<TC>(A.max_energy[k]);
                }
                sig[GRAD ? pair_index<C>(a, b) : 0] = en;
            }
        }
        auto sigma_of = [&](int a, int b) -> TC {
            if (GRAD) return a < b ? sig[GRAD ? pair_index<C>(a, b) : 0] : sig[GRAD ? pair_index<C>(b, a) : 0];
            return static_cast<TC>(A.sigma[ph[a] * NP + ph[b]]);
        };
#pragma unroll U
        for (int a = 0; a < C; ++a)
#pragma unroll U
        for (int b = 0; b < C; ++b) {
            if (b <= a || !(EXACT || b < ne)) continue;
            const double scaled = sqrt(vr[a] * vr[b]);
            const TC scale = static_cast<TC>(scaled);
            const TC cf = TC(0.5) * (TC(1.0) + scale) * P;
            TC d = sigma_of(a, b) * ((L.lap[a] + cf * L.ev(a)) - (L.lap[b] + cf * L.ev(b)));
            if (ne > 2) {
#pragma unroll U
                for (int g = 0; g < C; ++g) {
                    if (g == a || g == b || !(EXACT || g < ne)) continue;
                    d += (sigma_of(b, g) - sigma_of(a, g)) * (L.lap[g] + cf * L.ev(g));
                    if (A.tj_model == 0) {
                        if (A.tj_factor != 0.0 || scaled < 1.0) {
                            const TC e3 = static_cast<TC>(A.sigma[ph[a] * NP + ph[b]]) +
                                          static_cast<TC>(A.sigma[ph[b] * NP + ph[g]]) +
                                          static_cast<TC>(A.sigma[ph[a] * NP + ph[g]]);
                            d += static_cast<TC>(smax(A.tj_factor, (1.0 - scaled))) * P * e3 *
                                 (L.ev(a) * L.ev(g) - L.ev(b) * L.ev(g));
                        }
                    } else if (A.tj_model == 1) {
                        d += static_cast<TC>(A.tj_energy) * P * (L.ev(a) * L.ev(g) - L.ev(b) * L.ev(g));
                    }
                }
            }
            const TC mob = static_cast<TC>(A.mu[ph[a] * NP + ph[b]]);
            L.emax = smax(L.emax, static_cast<double>(sigma_of(a, b)));
            L.mmax = smax(L.mmax, static_cast<double>(mob));
            d *= mob * static_cast<TC>(norm_1) * scale;
            // constant driving force on interface cells (flag 2 = more than one own entry)
            if (DG && n0 > 1 && A.dg_mode[ph[a] * NP + ph[b]] == 1) {
                const TC raw = dg_local<TC>(A, a, b, n0, cell, ph[a] * NP + ph[b]);
                const TC weight = sqrt(L.ev(a) * L.ev(b));
                const TC allowed = static_cast<TC>(A.dg_allowed[ph[a] * NP + ph[b]]);
                const TC dgl = allowed * tanh(raw / allowed);
                d += dgl * mob * weight * static_cast<TC>(A.dg_prefactor);
            }
            psi[pair_index<C>(a, b)] = static_cast<double>(d);
        }
    }

    // ---- normalization / limiter (FP64); alive[p]: pair still in the increment list
    double (&phi)[K] = L.phi;
    const double dt = A.dt;
    bool alive[NPC];
#pragma unroll U
    for (int p = 0; p < NPC; ++p) alive[p] = false;
    if (ne == 2) {
        const double va = psi[0] * dt;
        const double old = phi[0];
        alive[0] = true;
        if ((old == 0.0 && va < 0.0) || (old == 1.0 && va > 0.0)) {
            alive[0] = false;
        } else {
            const double nv = old + va;
            double norm = 1.0;
            if (nv < 0.0) norm *= -old / va;
            else if (nv > 1.0) norm *= (1.0 - old) / va;
            if (norm > DBL_EPSILON) psi[0] *= norm;
            else alive[0] = false;
        }
    } else if (ne > 2) {
        // zero the pairs that would push a field at 0 below or at 1 above
#pragma unroll U
        for (int a = 0; a < C; ++a) {
            if (!(EXACT || a < ne)) continue;
            double dpsi = 0.0;
#pragma unroll U
            for (int b = 0; b < C; ++b) {
                if (b == a || !(EXACT || b < ne)) continue;
                dpsi += a < b ? psi[pair_index<C>(a, b)] : psi[pair_index<C>(b, a)] * (-1.0);
            }
            if ((phi[a] == 0.0 && dpsi < 0.0) || (phi[a] == 1.0 && dpsi > 0.0)) {
#pragma unroll U
                for (int b = 0; b < C; ++b) {
                    if (b == a || !(EXACT || b < ne)) continue;
                    psi[a < b ? pair_index<C>(a, b) : pair_index<C>(b, a)] = 0.0;
                }
            }
        }
        bool any = false;
#pragma unroll U
        for (int a = 0; a < C; ++a)
#pragma unroll U
        for (int b = 0; b < C; ++b) {
            if (b <= a || !(EXACT || b < ne)) continue;
            const int p = pair_index<C>(a, b);
            alive[p] = psi[p] != 0.0;
            any = any || alive[p];
        }

        if (any) {
            bool needed = true;
            int iter = 0;
            while (needed) {
                ++iter;
                needed = false;
                double pos[C], neg[C], lim1[C], lim2[C];
#pragma unroll U
                for (int t = 0; t < C; ++t) { pos[t] = 0.0; neg[t] = 0.0; }
#pragma unroll U
                for (int a = 0; a < C; ++a)
#pragma unroll U
                for (int b = 0; b < C; ++b) {
                    if (b <= a || !(EXACT || b < ne)) continue;
                    const int p = pair_index<C>(a, b);
                    if (!alive[p]) continue;
                    const double v = psi[p];
                    if (v < 0.0) { neg[a] += v; pos[b] += -v; }
                    if (v > 0.0) { pos[a] += v; neg[b] += -v; }
                }
#pragma unroll U
                for (int t = 0; t < C; ++t) {
                    lim1[t] = 1.0; lim2[t] = 1.0;
                    if (!(EXACT || t < ne)) continue;
                    const double npf = phi[t] + (pos[t] + neg[t]) * dt;
                    if (npf < 0.0) lim2[t] = smin(lim2[t], -(phi[t] + pos[t] * dt) / (neg[t] * dt));
                    if (npf > 1.0) lim1[t] = smin(lim1[t], (1.0 - (phi[t] + neg[t] * dt)) / (pos[t] * dt));
                }
#pragma unroll U
                for (int a = 0; a < C; ++a)
#pragma unroll U
                for (int b = 0; b < C; ++b) {
                    if (b <= a || !(EXACT || b < ne)) continue;
                    const int p = pair_index<C>(a, b);
                    if (!alive[p]) continue;
                    if (psi[p] < 0.0) {
                        const double t = smin(lim2[a], lim1[b]);
                        psi[p] *= t;
                        if (t < 1.0) needed = true;
                    }
                    if (psi[p] > 0.0) {
                        const double t = smin(lim1[a], lim2[b]);
                        psi[p] *= t;
                        if (t < 1.0) needed = true;
                    }
                }
                if (iter > 24) needed = false;
            }
#pragma unroll U
            for (int p = 0; p < NPC; ++p)
                if (alive[p] && fabs(psi[p] * dt) < DBL_EPSILON) alive[p] = false;

            // plausibility check of the limited increments (warning count)
            double tv[C];
#pragma unroll U
            for (int t = 0; t < C; ++t) tv[t] = phi[t];
#pragma unroll U
            for (int a = 0; a < C; ++a)
#pragma unroll U
            for (int b = 0; b < C; ++b) {
                if (b <= a || !(EXACT || b < ne)) continue;
                const int p = pair_index<C>(a, b);
                if (!alive[p] || psi[p] == 0.0) continue;
                tv[a] += psi[p] * dt;
                tv[b] += -psi[p] * dt;
            }
            unsigned long long warn = 0;
#pragma unroll U
            for (int t = 0; t < C; ++t)
                if ((EXACT || t < ne) && (tv[t] < -FLT_EPSILON || tv[t] > 1.0 + FLT_EPSILON)) ++warn;
            if (warn) atomicAdd(A.warnings, warn);
        }
    }

    // ---- merge
#pragma unroll U
    for (int a = 0; a < C; ++a)
#pragma unroll U
    for (int b = 0; b < C; ++b) {
        if (b <= a || !(EXACT || b < ne)) continue;
        const int p = pair_index<C>(a, b);
        if (!alive[p]) continue;
        const double value = 1.0 * psi[p] * dt;
        if (fabs(value) >= DBL_EPSILON) {
            phi[a] += value;
            phi[b] += -value;
        }
    }

    // ---- finalize: clamp, erase, renormalize (NodePF::finalize)
    double total = 0.0;
    int ns = 0;
#pragma unroll U
    for (int t = 0; t < C; ++t) {
        L.keep[t] = false;
        if (!(EXACT || t < ne)) continue;
        double v = phi[t];
        bool erase = false;
        if (v >= 1.0 - DBL_EPSILON) v = 1.0;
        if (v <= 0.0 + DBL_EPSILON) { v = 0.0; erase = true; }
        total += v;
        phi[t] = v;
        L.keep[t] = !erase;
        ns += erase ? 0 : 1;
    }
    empty = ns == 0;
    if (ns > 1) {
        const double norm = 1.0 / total;
#pragma unroll U
        for (int t = 0; t < C; ++t) if (L.keep[t]) phi[t] *= norm;
    } else if (ns == 1) {
#pragma unroll U
        for (int t = 0; t < C; ++t) if (L.keep[t]) phi[t] = 1.0;
    }
#pragma unroll
    for (int t = C; t < K; ++t) L.keep[t] = false;
    return true;
}

// Writes the finalized list of cell cc (compacted, order preserved); rounds phi to the storage
// type. Returns false (nothing written) when the list exceeds the stored capacity.
template <int K, typename TV, typename TC, bool GRAD>
__device__ __forceinline__ bool store_cell(const StepArgs& A, int64_t cc, CellList<K, TC, GRAD>& L)
{
    constexpr int U = K <= kFastSlots ? K : 1;
    constexpr bool IMPL = ListCode<TV>::kImplicit;
    const int64_t N = A.ncells;
    int ns = 0;
    int imp_t = -1;     // implicit entry (first largest value) by list position
    double vmax = -1.0;
#pragma unroll U
    for (int t = 0; t < K; ++t)
        if (L.keep[t]) {
            ns += 1;
            if (IMPL && L.phi[t] > vmax) { vmax = L.phi[t]; imp_t = t; }
        }
    if (ns > A.slots) return false;
    int w = 0, imp_w = 0;
    double others = 0.0;
#pragma unroll U
    for (int t = 0; t < K; ++t)
        if (L.keep[t]) {
            A.id_out[static_cast<int64_t>(w) * N + cc] = static_cast<uint16_t>(L.eid[t]);
            const TV sv = static_cast<TV>(L.phi[t]);
            if (IMPL && t == imp_t) {
                imp_w = w;
                if (ns > 1) static_cast<TV*>(A.val_out)[static_cast<int64_t>(w) * N + cc] = sv;
            } else {
                L.phi[t] = static_cast<double>(sv);
                others += L.phi[t];
                static_cast<TV*>(A.val_out)[static_cast<int64_t>(w) * N + cc] = sv;
            }
            ++w;
        }
    if (IMPL) {
        // the value the next step will recover (same summation order: slot order)
#pragma unroll U
        for (int t = 0; t < K; ++t)
            if (t == imp_t) L.phi[t] = 1.0 - others;
    }
    A.cnt_out[cc] = static_cast<uint8_t>(IMPL ? (w | (imp_w << 4)) : w);
    return true;
}

template <int K, typename TC, bool GRAD>
__device__ __forceinline__ void report_overflow(const StepArgs& A, int64_t c, CellList<K, TC, GRAD>& L)
{
    for (int t = 0; t < K; ++t) L.keep[t] = false;
    atomicOr(&A.status[kStatusOverflow], 1);
    atomicMin(A.first_overflow, static_cast<unsigned long long>(c));
}

// Domain of block b in a launch over several domains: first[d] <= b < first[d+1] (first holds
// ndom + 1 prefix sums of the domains' block counts).
__device__ __forceinline__ int batch_domain(const uint32_t* __restrict__ first, int ndom, unsigned b)
{
    int lo = 0, hi = ndom - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Workgroups are dispatched round-robin over the eight XCDs, each with its own L2. Renumbering
// them so that every XCD walks one contiguous range of cells keeps the neighbour rows a block
// reads mostly in the L2 that already holds them. A bijection of [0, nb).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    constexpr unsigned kXcds = 8;
    const unsigned q = nb / kXcds, r = nb % kXcds, x = b % kXcds, i = b / kXcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// wave_volume_add for entries of exactly 1.0 (single-entry cells, most of the grid) by integer
// popcounts: 1.0 is 2^116 in the fixed-point sums, bit 20 of limb 3. Other values take the
// FP64 butterfly. Must be reached by whole wavefronts.
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
    wave_volume_add<false>(act && v != 1.0, g, v, acc, bad);
}

// Volume contributions of whole lists: the wavefront walks the distinct grains of all lanes'
// kept entries once (not once per slot); each lane sums its own entries of that grain in slot
// order, the lanes are summed by the FP64 butterfly of wave_volume_add (fixed lane order, so
// deterministic when the cell-to-lane mapping is) and the leader adds the fixed-point image.
template <int K>
__device__ __forceinline__ void wave_volume_add_lists(const bool keep[K], const uint32_t eid[K], const double phi[K],
                                                      unsigned long long* acc, bool& bad)
{
    const int lane = static_cast<int>(threadIdx.x) & (warpSize - 1);
    bool pend[K];
#pragma unroll
    for (int t = 0; t < K; ++t) pend[t] = keep[t];
    while (true) {
        bool any = false;
        uint32_t mine = 0u;
#pragma unroll
        for (int t = K - 1; t >= 0; --t)
            if (pend[t]) { any = true; mine = eid[t]; }
        const unsigned long long m = __ballot(any);
        if (m == 0ull) break;
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        const uint32_t g0 = static_cast<uint32_t>(__shfl(static_cast<int>(mine), leader));
        double v = 0.0;
#pragma unroll
        for (int t = 0; t < K; ++t)
            if (pend[t] && eid[t] == g0) { v += phi[t]; pend[t] = false; }
#pragma unroll
        for (int k = 32; k >= 1; k >>= 1) v += __shfl_xor(v, k);
        if (lane == leader) {
            unsigned long long l[kLimbs];
            value_limbs(v, l, bad);
#pragma unroll
            for (int i = 0; i < kLimbs; ++i)
                if (l[i]) atomicAdd(&acc[static_cast<int64_t>(g0) * kLimbs + i], l[i]);
        }
    }
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
    // whole-list walk for FP32 states, per-slot walk for FP64 ones (whose extra live registers
    // make the whole-list loop spill): 3.14 vs 3.60 ms and 3.63 vs 3.86 ms per 256^3 step
    if constexpr (sizeof(TV) == 4) {
        wave_volume_add_lists<K>(L.keep, L.eid, L.phi, A.vol_acc, bad_volume);
    } else {
#pragma unroll
        for (int t = 0; t < K; ++t) wave_volume_add_split(L.keep[t], L.eid[t], L.phi[t], A.vol_acc, bad_volume);
    }
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
// the plain arithmetic at run time).
int launch_step(const StepArgs& one, const StepArgs* args, const uint32_t* first, int ndom, unsigned blocks, bool fp32,
                bool full, bool uni, void* stream)
{
    const dim3 grid(blocks);
    hipStream_t s = static_cast<hipStream_t>(stream);
#define MPF_STEP(TV, G, D, U) hipLaunchKernelGGL((step_kernel<TV, TV, G, D, U>), grid, dim3(kBlock), 0, s, one, args, first, ndom)
    if (fp32) {
        if (full)     MPF_STEP(float, true, true, false);
        else if (uni) MPF_STEP(float, false, false, true);
        else          MPF_STEP(float, false, false, false);
    } else {
        if (full)     MPF_STEP(double, true, true, false);
        else if (uni) MPF_STEP(double, false, false, true);
        else          MPF_STEP(double, false, false, false);
    }
#undef MPF_STEP
    return static_cast<int>(hipGetLastError());
}

int launch_deferred(const StepArgs* args, int ndom, int per_domain, bool fp32, bool full, void* stream)
{
    const dim3 grid(static_cast<unsigned>(per_domain) * static_cast<unsigned>(ndom));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid2(2u * static_cast<unsigned>(per_domain) * static_cast<unsigned>(ndom));
#define MPF_DEFERRED(TV, F)                                                                              \
    hipLaunchKernelGGL((mid_kernel<TV, TV, F, F, 0>), grid, dim3(kWideBlock), 0, s, args, ndom);    \
    hipLaunchKernelGGL((mid_tail_kernel<TV, TV, F, F>), grid2, dim3(kWideBlock), 0, s, args, ndom); \
    hipLaunchKernelGGL((wide_kernel<TV, TV, F, F>), grid, dim3(kWideBlock), 0, s, args, ndom);
    if (fp32) {
        if (full) { MPF_DEFERRED(float, true) } else { MPF_DEFERRED(float, false) }
    } else {
        if (full) { MPF_DEFERRED(double, true) } else { MPF_DEFERRED(double, false) }
    }
#undef MPF_DEFERRED
    return static_cast<int>(hipGetLastError());
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
