// Per-cell operations of the step shared by the step kernels (mpf_kernels.hip, mpf_tile.hip):
// interface properties, double-obstacle pair increments, driving force, limiter, merge and
// finalize of one cell whose merged list is in registers (compute_cell), the store of the next
// list (store_cell), and launch-order helpers. Everything sits in an anonymous namespace.
#pragma once

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
                const TC en = ip_pair_energy<TC>(A, k, L.eid[a], L.eid[b], L.ev(a), L.ev(b), L.gr[GRAD ? a : 0],
                                                 L.gr[GRAD ? b : 0]);
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
            if (DG && A.reg) {
                // DoubleObstacle's curvature overload (src/DoubleObstacle.cpp:753-767), then
                // InterfaceRegularization::MergePhaseFieldIncrementsSR (src/InterfaceRegularization.cpp:425-456)
                // on interface cells for the pairs of own entries (non-zero values)
                d = d * mob * static_cast<TC>(norm_1) * scale;
                if (A.reg_active) {
                    d *= static_cast<TC>(A.reg_factor);
                    if (A.curv != nullptr && n0 > 1 && b < n0) {
                        const TC kappa = static_cast<TC>(A.curv[static_cast<int64_t>(b * (b - 1) / 2 + a) * A.ncells + cell]);
                        d += kappa * mob * sqrt(L.ev(a) * L.ev(b)) * static_cast<TC>(A.dg_prefactor);
                    }
                }
            } else {
                d *= mob * static_cast<TC>(norm_1) * scale;
            }
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

}  // namespace

}  // namespace mpf
