// Reverse sweep of the multiphase-field cell update for one interface cell (the adjoint's per-cell
// part), host- and device-callable: the kernels of mpf_adjoint.hip run it per cell; a host build
// checks it against finite differences of the same forward (cell_forward).
//
// Forward (cell_forward), as the step kernels' compute_cell in its general form: pair energies
// (InterfaceProperties::SetSR, src/InterfaceProperties.cpp:603-762), double-obstacle pair
// increments psi (src/DoubleObstacle.cpp:502-571) with the triple-junction terms and the constant
// driving force (src/DrivingForce.cpp:497-602), then part2: the sum-to-one limiter
// (PhaseField::NormalizeIncrementsSR, src/PhaseField.cpp:1351-1567), the merge (:1268-1299) and
// NodePF::finalize (include/Containers/NodePF.h:214-255).
#pragma once

#include <float.h>
#include <math.h>
#include <stdint.h>

#include "mpf_device.h"
#include "mpf_lists.h"

#define MPF_HD __host__ __device__

namespace mpf {


template <int C> struct NPairs { static constexpr int value = C * (C - 1) / 2; };

// What the reverse sweep needs of the forward limiter, merge and finalize of one cell.
template <int C> struct Part2Rec {
    uint32_t live;     // bit p: pair p's increment reaches the merge (not zeroed at a field at 0 or 1,
                       // not dropped); the merge's skip of |psi dt| < DBL_EPSILON changes a value by
                       // less than DBL_EPSILON and counts as passing
    uint32_t keep;     // bit t: finalize kept entry t
    uint32_t fixed;    // bit t: finalize replaced the merged value (1 or erased)
    uint32_t alive;    // bit p: pair p in the limiter loop (live and nonzero)
    bool limited;      // the limiter of three or more entries scaled a pair
    int ne2;           // two entries: 0 as computed, 1 scaled to reach 0 or 1, 2 dropped
    int nit;           // limiter iterations
    double vfin[C];    // values after the clamps, before renormalization
    double total;
    int ns;
    double psi0[NPairs<C>::value];  // increments entering the limiter loop
};

// One pass of the sum-to-one limiter of three or more entries (PhaseField::NormalizeIncrementsSR,
// src/PhaseField.cpp:1451-1519) over the pairs in `alive`: per entry the sums pos / neg of its
// positive / negative increments; an entry that would fall below 0 (above 1) limits its
// increments by r2 = -(phi + pos dt) / (neg dt) (r1 = (1 - (phi + neg dt)) / (pos dt)); each pair
// takes the smaller limit of its two entries. Returns whether a limit below 1 was applied.
template <int C>
MPF_HD __forceinline__ bool limiter_pass(double (&psi)[NPairs<C>::value], const double (&phi)[C], double dt, uint32_t alive)
{
    constexpr int UA = C <= 5 ? C : 1;
    bool needed = false;
    double pos[C], neg[C], lim1[C], lim2[C];
#pragma unroll UA
    for (int t = 0; t < C; ++t) { pos[t] = 0.0; neg[t] = 0.0; }
#pragma unroll UA
    for (int a = 0; a < C; ++a)
#pragma unroll UA
    for (int b = a + 1; b < C; ++b) {
        const int p = pair_index<C>(a, b);
        if (!(alive >> p & 1u)) continue;
        const double v = psi[p];
        if (v < 0.0) { neg[a] += v; pos[b] += -v; }
        if (v > 0.0) { pos[a] += v; neg[b] += -v; }
    }
#pragma unroll UA
    for (int t = 0; t < C; ++t) {
        lim1[t] = 1.0;
        lim2[t] = 1.0;
        const double npf = phi[t] + (pos[t] + neg[t]) * dt;
        if (npf < 0.0) lim2[t] = smin(lim2[t], -(phi[t] + pos[t] * dt) / (neg[t] * dt));
        if (npf > 1.0) lim1[t] = smin(lim1[t], (1.0 - (phi[t] + neg[t] * dt)) / (pos[t] * dt));
    }
#pragma unroll UA
    for (int a = 0; a < C; ++a)
#pragma unroll UA
    for (int b = a + 1; b < C; ++b) {
        const int p = pair_index<C>(a, b);
        if (!(alive >> p & 1u)) continue;
        if (psi[p] < 0.0) {
            const double f = smin(lim2[a], lim1[b]);
            psi[p] *= f;
            if (f < 1.0) needed = true;
        }
        if (psi[p] > 0.0) {
            const double f = smin(lim1[a], lim2[b]);
            psi[p] *= f;
            if (f < 1.0) needed = true;
        }
    }
    return needed;
}

// Limiter, merge and finalize of a cell of C entries (the FP64 part of the step kernels'
// compute_cell). psi is modified; out gets the finalized values.
template <int C>
MPF_HD __forceinline__ void part2(double dt, double (&psi)[NPairs<C>::value], const double (&phi)[C], double (&out)[C],
                         Part2Rec<C>& R)
{
    constexpr int NPC = NPairs<C>::value;
    constexpr int UA = C <= 5 ? C : 1;  // full unroll only for short lists (code size)
    uint32_t alive = 0u;
    R.limited = false;
    R.ne2 = 0;
    R.nit = 0;
    R.live = R.keep = R.fixed = R.alive = 0u;
    if (C == 2) {
        const double va = psi[0] * dt;
        const double old = phi[0];
        alive = 1u;
        if ((old == 0.0 && va < 0.0) || (old == 1.0 && va > 0.0)) {
            alive = 0u;
            R.ne2 = 2;
        } else {
            const double nv = old + va;
            double norm = 1.0;
            bool scaled = false;
            if (nv < 0.0) { norm *= -old / va; scaled = true; }
            else if (nv > 1.0) { norm *= (1.0 - old) / va; scaled = true; }
            if (norm > DBL_EPSILON) {
                psi[0] *= norm;
                R.ne2 = scaled ? 1 : 0;
                R.live = 1u;
            } else {
                alive = 0u;
                R.ne2 = 2;
            }
        }
    } else {
        // pairs that would push a field at 0 below or at 1 above are zeroed
        uint32_t zeroed = 0u;
#pragma unroll UA
        for (int a = 0; a < C; ++a) {
            double dpsi = 0.0;
#pragma unroll UA
            for (int b = 0; b < C; ++b) {
                if (b == a) continue;
                dpsi += a < b ? psi[pair_index<C>(a, b)] : psi[pair_index<C>(b, a)] * (-1.0);
            }
            if ((phi[a] == 0.0 && dpsi < 0.0) || (phi[a] == 1.0 && dpsi > 0.0)) {
#pragma unroll UA
                for (int b = 0; b < C; ++b) {
                    if (b == a) continue;
                    const int p = a < b ? pair_index<C>(a, b) : pair_index<C>(b, a);
                    psi[p] = 0.0;
                    zeroed |= 1u << p;
                }
            }
        }
#pragma unroll UA
        for (int p = 0; p < NPC; ++p) {
            R.psi0[p] = psi[p];
            if (psi[p] != 0.0) alive |= 1u << p;
        }
        R.alive = alive;
        R.live = ~zeroed & ((1u << NPC) - 1u);
        if (alive) {
            bool needed = true;
            int iter = 0;
            while (needed) {
                ++iter;
                needed = limiter_pass<C>(psi, phi, dt, alive);
                R.limited = R.limited || needed;
                if (iter > 24) needed = false;
            }
            R.nit = iter;
#pragma unroll UA
            for (int p = 0; p < NPC; ++p)
                if ((alive >> p & 1u) && fabs(psi[p] * dt) < DBL_EPSILON) alive &= ~(1u << p);
        }
    }
    // merge
    double m[C];
#pragma unroll UA
    for (int t = 0; t < C; ++t) m[t] = phi[t];
#pragma unroll UA
    for (int a = 0; a < C; ++a)
#pragma unroll UA
    for (int b = a + 1; b < C; ++b) {
        const int p = pair_index<C>(a, b);
        if (!(alive >> p & 1u)) continue;
        const double value = 1.0 * psi[p] * dt;
        if (fabs(value) >= DBL_EPSILON) {
            m[a] += value;
            m[b] += -value;
        }
    }
    // finalize
    double total = 0.0;
    int ns = 0;
#pragma unroll UA
    for (int t = 0; t < C; ++t) {
        double v = m[t];
        bool erase = false, fixed = false;
        if (v >= 1.0 - DBL_EPSILON) { v = 1.0; fixed = true; }
        if (v <= 0.0 + DBL_EPSILON) { v = 0.0; erase = true; fixed = true; }
        total += v;
        out[t] = v;
        R.vfin[t] = v;
        if (!erase) { R.keep |= 1u << t; ++ns; }
        if (fixed) R.fixed |= 1u << t;
    }
    R.total = total;
    R.ns = ns;
    if (ns > 1) {
        const double norm = 1.0 / total;
#pragma unroll UA
        for (int t = 0; t < C; ++t) if (R.keep >> t & 1u) out[t] *= norm;
    } else if (ns == 1) {
#pragma unroll UA
        for (int t = 0; t < C; ++t) if (R.keep >> t & 1u) out[t] = 1.0;
    }
}

MPF_HD inline void gadd(double* g, int e, double v) { g[e] += v; }

// Cotangent of the next state's value of grain g at cell c: by grain through S_{t+1}'s list, or
// (no list given) by the position w among the entries the step keeps.
MPF_HD inline double lam_of(const AdjointArgs& J, bool fp32, int64_t c, uint32_t g, int w)
{
    const StepArgs& A = J.s;
    const int64_t N = A.ncells;
    if (J.next_cnt != nullptr) {
        const uint8_t b = J.next_cnt[c];
        const int n = fp32 ? (b & 15) : b;
        for (int s = 0; s < n; ++s)
            if (J.next_id[static_cast<int64_t>(s) * N + c] == g) return J.lam_next[static_cast<int64_t>(s) * N + c];
        return 0.0;
    }
    return w < A.slots ? J.lam_next[static_cast<int64_t>(w) * N + c] : 0.0;
}

// Reverse sweep through the limiter iterations of a cell of three or more entries, each
// psi_p <- psi_p f_p (limiter_pass); the increments entering iteration k are recomputed from
// those entering the loop. lam: in, the cotangent of psi after the loop; out, before it (pairs
// zeroed before the loop: 0). lam_phi gets the parts through the limits. An increment of 0 is
// differentiated from the side of the sign it entered the loop with; one that entered as exactly
// 0 (a symmetric configuration, e.g. equal energies of the pairs around it) by the mean of the
// two one-sided derivatives, the limit of a central difference.
template <int C>
MPF_HD __forceinline__ void limiter_reverse(const Part2Rec<C>& R, const double (&phi)[C], double dt,
                                   double (&lam)[NPairs<C>::value], double (&lam_phi)[C])
{
    constexpr int NPC = NPairs<C>::value;
    constexpr int UA = C <= 5 ? C : 1;
#pragma unroll UA
    for (int p = 0; p < NPC; ++p)
        if (!(R.live >> p & 1u)) lam[p] = 0.0;
    for (int k = R.nit - 1; k >= 0; --k) {
        double ps[NPC];
#pragma unroll UA
        for (int p = 0; p < NPC; ++p) ps[p] = R.psi0[p];
        for (int j = 0; j < k; ++j) limiter_pass<C>(ps, phi, dt, R.alive);
        double pos[C], neg[C], r1[C], r2[C];
        bool c1[C], c2[C];
#pragma unroll UA
        for (int t = 0; t < C; ++t) { pos[t] = 0.0; neg[t] = 0.0; }
#pragma unroll UA
        for (int a = 0; a < C; ++a)
#pragma unroll UA
        for (int b = a + 1; b < C; ++b) {
            const int p = pair_index<C>(a, b);
            if (!(R.alive >> p & 1u)) continue;
            const double v = ps[p];
            if (v < 0.0) { neg[a] += v; pos[b] += -v; }
            if (v > 0.0) { pos[a] += v; neg[b] += -v; }
        }
#pragma unroll UA
        for (int t = 0; t < C; ++t) {
            const double npf = phi[t] + (pos[t] + neg[t]) * dt;
            r2[t] = 1.0;
            r1[t] = 1.0;
            c2[t] = c1[t] = false;
            if (npf < 0.0) {
                const double r = -(phi[t] + pos[t] * dt) / (neg[t] * dt);
                if (r < 1.0) { r2[t] = r; c2[t] = true; }
            }
            if (npf > 1.0) {
                const double r = (1.0 - (phi[t] + neg[t] * dt)) / (pos[t] * dt);
                if (r < 1.0) { r1[t] = r; c1[t] = true; }
            }
        }
        double lam_r1[C], lam_r2[C], nl[NPC];
#pragma unroll UA
        for (int t = 0; t < C; ++t) { lam_r1[t] = 0.0; lam_r2[t] = 0.0; }
#pragma unroll UA
        for (int a = 0; a < C; ++a)
#pragma unroll UA
        for (int b = a + 1; b < C; ++b) {
            const int p = pair_index<C>(a, b);
            nl[p] = lam[p];
            if (!(R.live >> p & 1u)) continue;
            const double v = ps[p], v0 = R.psi0[p];
            if (v == 0.0 && v0 == 0.0) {
                nl[p] = lam[p] * 0.5 * (smin(r2[a], r1[b]) + smin(r1[a], r2[b]));
                continue;
            }
            // f = min(first, second), the first winning ties
            const bool neg_pair = v != 0.0 ? v < 0.0 : v0 < 0.0;
            const double first = neg_pair ? r2[a] : r1[a], second = neg_pair ? r1[b] : r2[b];
            const bool take_second = second < first;
            nl[p] = lam[p] * (take_second ? second : first);
            if (v == 0.0) continue;
            const double lf = lam[p] * v;
            if (neg_pair) {
                if (take_second) { if (c1[b]) lam_r1[b] += lf; }
                else if (c2[a]) lam_r2[a] += lf;
            } else {
                if (take_second) { if (c2[b]) lam_r2[b] += lf; }
                else if (c1[a]) lam_r1[a] += lf;
            }
        }
        double lam_pos[C], lam_neg[C];
#pragma unroll UA
        for (int t = 0; t < C; ++t) {
            lam_pos[t] = 0.0;
            lam_neg[t] = 0.0;
            if (c2[t]) {
                lam_phi[t] += lam_r2[t] * (-1.0 / (neg[t] * dt));
                lam_pos[t] += lam_r2[t] * (-1.0 / neg[t]);
                lam_neg[t] += lam_r2[t] * (-r2[t] / neg[t]);
            }
            if (c1[t]) {
                lam_phi[t] += lam_r1[t] * (-1.0 / (pos[t] * dt));
                lam_neg[t] += lam_r1[t] * (-1.0 / pos[t]);
                lam_pos[t] += lam_r1[t] * (-r1[t] / pos[t]);
            }
        }
#pragma unroll UA
        for (int a = 0; a < C; ++a)
#pragma unroll UA
        for (int b = a + 1; b < C; ++b) {
            const int p = pair_index<C>(a, b);
            if (!(R.live >> p & 1u)) continue;
            const double v = ps[p], v0 = R.psi0[p];
            const double side = v != 0.0 ? v : v0;
            if (side < 0.0) nl[p] += lam_neg[a] - lam_pos[b];
            else if (side > 0.0) nl[p] += lam_pos[a] - lam_neg[b];
            else nl[p] += 0.5 * ((lam_neg[a] - lam_pos[b]) + (lam_pos[a] - lam_neg[b]));
        }
#pragma unroll UA
        for (int p = 0; p < NPC; ++p) lam[p] = nl[p];
    }
}

// Forward of one wide cell whose merged list has C entries, with what the reverse sweep reads.
template <int C, bool GRAD> struct CellFwd {
    static constexpr int NPC = NPairs<C>::value;
    int ph[C];
    double vr[C], phi[C], lap[C];
    double norm_1;
    double sig[GRAD ? NPC : 1];  // pair energies (orientation-dependent models)
    double psiB[NPC];            // increments before the limiter
    double dpre[NPC];            // their double-obstacle part before the mobility factor
    double out[C];               // finalized values
    Part2Rec<C> R;
};

template <int C, bool GRAD, int K>
MPF_HD __forceinline__ void cell_forward(const AdjointArgs& J, const CellList<K, double, GRAD>& L, CellFwd<C, GRAD>& F)
{
    constexpr int NPC = NPairs<C>::value;
    constexpr int UA = C <= 5 ? C : 1;  // full unroll only for short lists (code size)
    const StepArgs& A = J.s;
    const int NP = A.nphases;
    const int n0 = L.n0;
    const double Pf = A.prefactor, dt = A.dt;
    int (&ph)[C] = F.ph;
    double (&vr)[C] = F.vr;
    double (&phi)[C] = F.phi;
    double (&lap)[C] = F.lap;
#pragma unroll UA
    for (int t = 0; t < C; ++t) {
        ph[t] = A.gphase[L.eid[t]];
        vr[t] = A.gvr[L.eid[t]];
        phi[t] = L.phi[t];
        lap[t] = L.lap[t];
    }
    double localn = 0.0;
    if (A.consider_nucleus) {
#pragma unroll UA
        for (int t = 0; t < C; ++t) localn += vr[t];
    } else {
        localn = static_cast<double>(C);
    }
    const double norm_1 = 1.0 / localn;
    F.norm_1 = norm_1;
    double (&sig)[GRAD ? NPC : 1] = F.sig;
    if (GRAD) {
#pragma unroll UA
        for (int a = 0; a < C; ++a)
#pragma unroll UA
        for (int b = a + 1; b < C; ++b)
            sig[GRAD ? pair_index<C>(a, b) : 0] =
                ip_pair_energy<double>(A, ph[a] * NP + ph[b], L.eid[a], L.eid[b], phi[a], phi[b],
                                       L.gr[GRAD ? a : 0], L.gr[GRAD ? b : 0]);
    }
    auto sg = [&](int x, int y) -> double {
        if (GRAD) return x < y ? sig[GRAD ? pair_index<C>(x, y) : 0] : sig[GRAD ? pair_index<C>(y, x) : 0];
        return A.sigma[ph[x] * NP + ph[y]];
    };
    auto tj_on = [&](double scaled) { return A.tj_model == 0 ? (A.tj_factor != 0.0 || scaled < 1.0) : A.tj_model == 1; };
    auto tj_coef = [&](double scaled) { return A.tj_model == 0 ? smax(A.tj_factor, (1.0 - scaled)) * Pf : A.tj_energy * Pf; };
#pragma unroll UA
    for (int a = 0; a < C; ++a)
#pragma unroll UA
    for (int b = a + 1; b < C; ++b) {
        const int p = pair_index<C>(a, b);
        const int k = ph[a] * NP + ph[b];
        const double scale = sqrt(vr[a] * vr[b]);
        const double cf = 0.5 * (1.0 + scale) * Pf;
        double d = sg(a, b) * ((lap[a] + cf * phi[a]) - (lap[b] + cf * phi[b]));
#pragma unroll UA
        for (int g = 0; g < C; ++g) {
            if (C < 3 || g == a || g == b) continue;
            d += (sg(b, g) - sg(a, g)) * (lap[g] + cf * phi[g]);
            if (tj_on(scale)) {
                const double e3 = A.sigma[ph[a] * NP + ph[b]] + A.sigma[ph[b] * NP + ph[g]] +
                                  A.sigma[ph[a] * NP + ph[g]];
                d += A.tj_model == 0 ? tj_coef(scale) * e3 * (phi[a] * phi[g] - phi[b] * phi[g])
                                     : tj_coef(scale) * (phi[a] * phi[g] - phi[b] * phi[g]);
            }
        }
        F.dpre[p] = d;
        const double mob = A.mu[k];
        d *= mob * norm_1 * scale;
        if (J.driving_force && n0 > 1 && A.dg_mode[k] == 1) {
            const double weight = sqrt(phi[a] * phi[b]);
            const double allowed = A.dg_allowed[k];
            const double dgl = allowed * tanh(A.dg_value[k] / allowed);
            d += dgl * mob * weight * A.dg_prefactor;
        }
        F.psiB[p] = d;
    }
    double psiF[NPC];
#pragma unroll UA
    for (int p = 0; p < NPC; ++p) psiF[p] = F.psiB[p];
    part2<C>(dt, psiF, phi, F.out, F.R);
}

// Reverse sweep of one wide cell whose merged list has C entries (list capacity K).
template <int C, bool GRAD, int K>
MPF_HD __forceinline__ void cell_adjoint(const AdjointArgs& J, bool fp32, int64_t c, const CellList<K, double, GRAD>& L,
                         double (&lam_phi)[K], double (&lam_lap)[K], double (&lam_gr)[GRAD ? K : 1][3], double* gsl)
{
    constexpr int NPC = NPairs<C>::value;
    constexpr int UA = C <= 5 ? C : 1;  // full unroll only for short lists (code size)
    const StepArgs& A = J.s;
    const int NP = A.nphases, NP2 = NP * NP;
    const int n0 = L.n0;
    const double Pf = A.prefactor, dt = A.dt;
    CellFwd<C, GRAD> F;
    cell_forward<C, GRAD, K>(J, L, F);
    const int (&ph)[C] = F.ph;
    const double (&vr)[C] = F.vr;
    const double (&phi)[C] = F.phi;
    const double (&lap)[C] = F.lap;
    const double norm_1 = F.norm_1;
    const double (&sig)[GRAD ? NPC : 1] = F.sig;
    const double (&dpre)[NPC] = F.dpre;
    const Part2Rec<C>& R = F.R;
    auto sg = [&](int x, int y) -> double {
        if (GRAD) return x < y ? sig[GRAD ? pair_index<C>(x, y) : 0] : sig[GRAD ? pair_index<C>(y, x) : 0];
        return A.sigma[ph[x] * NP + ph[y]];
    };
    auto tj_on = [&](double scaled) { return A.tj_model == 0 ? (A.tj_factor != 0.0 || scaled < 1.0) : A.tj_model == 1; };
    auto tj_coef = [&](double scaled) { return A.tj_model == 0 ? smax(A.tj_factor, (1.0 - scaled)) * Pf : A.tj_energy * Pf; };
    auto dg_on = [&](int k) { return J.driving_force && n0 > 1 && A.dg_mode[k] == 1; };

    // cotangents of the finalized values (kept entries only)
    double lam_out[C];
    {
        int w = 0;
#pragma unroll UA
        for (int t = 0; t < C; ++t) {
            lam_out[t] = 0.0;
            if (R.keep >> t & 1u) {
                lam_out[t] = lam_of(J, fp32, c, L.eid[t], w);
                ++w;
            }
        }
    }

    // ---- reverse: finalize, merge, limiter -> cotangents of psiB and (directly) of phi
    double lam_psiB[NPC], lam_pd[C];
    {
        double lam_m[C];
#pragma unroll UA
        for (int t = 0; t < C; ++t) lam_m[t] = 0.0;
        if (R.ns > 1) {
            const double norm = 1.0 / R.total;
            double s = 0.0;
#pragma unroll UA
            for (int t = 0; t < C; ++t) if (R.keep >> t & 1u) s += lam_out[t] * R.vfin[t];
            const double lam_total = -s * norm * norm;
#pragma unroll UA
            for (int t = 0; t < C; ++t) {
                const double lv = ((R.keep >> t & 1u) ? lam_out[t] * norm : 0.0) + lam_total;
                lam_m[t] = (R.fixed >> t & 1u) ? 0.0 : lv;
            }
        }
#pragma unroll UA
        for (int t = 0; t < C; ++t) lam_pd[t] = lam_m[t];
#pragma unroll UA
        for (int a = 0; a < C; ++a)
#pragma unroll UA
        for (int b = a + 1; b < C; ++b) {
            const int p = pair_index<C>(a, b);
            lam_psiB[p] = (R.live >> p & 1u) ? dt * (lam_m[a] - lam_m[b]) : 0.0;
        }
        if (C == 2 && R.ne2 == 1) {
            // psi scaled to -phi_0 / dt or (1 - phi_0) / dt: no dependence on psi
            lam_pd[0] += lam_psiB[0] * (-1.0 / dt);
            lam_psiB[0] = 0.0;
        } else if (R.limited) {
            limiter_reverse<C>(R, phi, dt, lam_psiB, lam_pd);
        }
    }

    // ---- reverse: increments -> phi, Laplacians, gradients, tables
#pragma unroll UA
    for (int t = 0; t < C; ++t) {
        lam_phi[t] = lam_pd[t];
        lam_lap[t] = 0.0;
        if (GRAD) lam_gr[GRAD ? t : 0][0] = lam_gr[GRAD ? t : 0][1] = lam_gr[GRAD ? t : 0][2] = 0.0;
    }
    double lam_sig[GRAD ? NPC : 1];
#pragma unroll UA
    for (int p = 0; p < (GRAD ? NPC : 1); ++p) lam_sig[p] = 0.0;
    // cotangent of the energy sg(x, y) as the forward reads it
    auto sig_bar = [&](int x, int y, double v) {
        if (GRAD) lam_sig[GRAD ? (x < y ? pair_index<C>(x, y) : pair_index<C>(y, x)) : 0] += v;
        else gadd(gsl, ph[x] * NP + ph[y], v);
    };
#pragma unroll UA
    for (int a = 0; a < C; ++a)
#pragma unroll UA
    for (int b = a + 1; b < C; ++b) {
        const int p = pair_index<C>(a, b);
        const double Lp = lam_psiB[p];
        if (Lp == 0.0) continue;
        const int k = ph[a] * NP + ph[b];
        const double scale = sqrt(vr[a] * vr[b]);
        const double cf = 0.5 * (1.0 + scale) * Pf;
        const double mob = A.mu[k];
        gadd(gsl, NP2 + k, Lp * dpre[p] * norm_1 * scale);
        if (dg_on(k)) {
            const double weight = sqrt(phi[a] * phi[b]);
            const double allowed = A.dg_allowed[k];
            const double u = A.dg_value[k] / allowed;
            const double th = tanh(u);
            const double dgl = allowed * th;
            const double pre = A.dg_prefactor;
            gadd(gsl, NP2 + k, Lp * dgl * weight * pre);
            gadd(gsl, 2 * NP2 + k, Lp * mob * weight * pre * (1.0 - th * th));
            // the tanh amplitude: Limit * Pi/Eta * MaxEnergy * RegularizationFactor
            const double dmax = A.emodel[k] == 1 ? 1.0 + A.eps[k] * 1.41 : 1.0;
            gadd(gsl, k, Lp * mob * weight * pre * (th - u * (1.0 - th * th)) * J.dg_limit * pre * J.regularization * dmax);
            if (weight > 0.0) {
                const double lw = Lp * dgl * mob * pre / (2.0 * weight);
                lam_phi[a] += lw * phi[b];
                lam_phi[b] += lw * phi[a];
            }
        }
        const double Ld = Lp * (mob * norm_1 * scale);
        const double sab = sg(a, b);
        sig_bar(a, b, Ld * ((lap[a] + cf * phi[a]) - (lap[b] + cf * phi[b])));
        lam_lap[a] += Ld * sab;
        lam_phi[a] += Ld * sab * cf;
        lam_lap[b] -= Ld * sab;
        lam_phi[b] -= Ld * sab * cf;
#pragma unroll UA
        for (int g = 0; g < C; ++g) {
            if (C < 3 || g == a || g == b) continue;
            const double xg = lap[g] + cf * phi[g];
            const double cg = sg(b, g) - sg(a, g);
            lam_lap[g] += Ld * cg;
            lam_phi[g] += Ld * cg * cf;
            sig_bar(b, g, Ld * xg);
            sig_bar(a, g, -Ld * xg);
            if (tj_on(scale)) {
                const double cc = tj_coef(scale);
                double f = cc;
                if (A.tj_model == 0) {
                    const double e3 = A.sigma[ph[a] * NP + ph[b]] + A.sigma[ph[b] * NP + ph[g]] +
                                      A.sigma[ph[a] * NP + ph[g]];
                    const double q = Ld * cc * (phi[a] * phi[g] - phi[b] * phi[g]);
                    gadd(gsl, ph[a] * NP + ph[b], q);
                    gadd(gsl, ph[b] * NP + ph[g], q);
                    gadd(gsl, ph[a] * NP + ph[g], q);
                    f = cc * e3;
                }
                lam_phi[a] += Ld * f * phi[g];
                lam_phi[b] -= Ld * f * phi[g];
                lam_phi[g] += Ld * f * (phi[a] - phi[b]);
            }
        }
    }
    // pair energies of orientation-dependent models -> table, values, gradients
    if (GRAD) {
#pragma unroll UA
        for (int a = 0; a < C; ++a)
#pragma unroll UA
        for (int b = a + 1; b < C; ++b) {
            const double v = lam_sig[GRAD ? pair_index<C>(a, b) : 0];
            if (v == 0.0) continue;
            const int k = ph[a] * NP + ph[b];
            if (A.emodel[k] == 0) { gadd(gsl, k, v); continue; }
            const double E = A.sigma[k], e1 = A.eps[k];
            // per side: weight, cubic energy, normal
            double w[2] = {0.0, 0.0}, ce[2] = {0.0, 0.0}, nv[2][3], vv[2][3], len[2] = {0.0, 0.0};
            const int xs[2] = {a, b};
#pragma unroll UA
            for (int s = 0; s < 2; ++s) {
                const int x = xs[s];
                nv[s][0] = nv[s][1] = nv[s][2] = 0.0;
                vv[s][0] = vv[s][1] = vv[s][2] = 0.0;
                if (!A.gsolid[L.eid[x]]) continue;
                const double wx = fabs(phi[x] * (1.0 - phi[x]));
                if (!(wx > DBL_EPSILON)) continue;
                w[s] = wx;
                const double* Rm = A.grot + 9 * static_cast<int64_t>(L.eid[x]);
                double v0 = L.gr[GRAD ? x : 0][0] * -1.0, v1 = L.gr[GRAD ? x : 0][1] * -1.0, v2 = L.gr[GRAD ? x : 0][2] * -1.0;
                const double l = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                if (l > DBL_EPSILON) {
                    const double inv = 1.0 / l;
                    v0 *= inv; v1 *= inv; v2 *= inv;
                    len[s] = l;
                } else {
                    v0 = v1 = v2 = 0.0;
                }
                vv[s][0] = v0; vv[s][1] = v1; vv[s][2] = v2;
                for (int i = 0; i < 3; ++i) nv[s][i] = Rm[i] * v0 + Rm[3 + i] * v1 + Rm[6 + i] * v2;
                const double q = nv[s][0] * nv[s][0] * nv[s][0] * nv[s][0] + nv[s][1] * nv[s][1] * nv[s][1] * nv[s][1] +
                                 nv[s][2] * nv[s][2] * nv[s][2] * nv[s][2];
                ce[s] = E * (1.0 + e1 * (1.5 - 2.5 * q));
            }
            const double W = w[0] + w[1];
            if (!(W > DBL_EPSILON)) {
                gadd(gsl, k, v * (1.0 + e1 * 1.41));  // MaxEnergy
                continue;
            }
            const double sv = (ce[0] * w[0] + ce[1] * w[1]) / W;
#pragma unroll UA
            for (int s = 0; s < 2; ++s) {
                if (w[s] == 0.0) continue;
                const int x = xs[s];
                const double lc = v * w[s] / W;
                const double lw = v * (ce[s] - sv) / W;
                const double q = nv[s][0] * nv[s][0] * nv[s][0] * nv[s][0] + nv[s][1] * nv[s][1] * nv[s][1] * nv[s][1] +
                                 nv[s][2] * nv[s][2] * nv[s][2] * nv[s][2];
                gadd(gsl, k, lc * (1.0 + e1 * (1.5 - 2.5 * q)));
                const double px = phi[x] * (1.0 - phi[x]);
                lam_phi[x] += lw * (px < 0.0 ? -1.0 : 1.0) * (1.0 - 2.0 * phi[x]);
                if (len[s] > 0.0) {
                    const double* Rm = A.grot + 9 * static_cast<int64_t>(L.eid[x]);
                    double ln[3], lv[3];
                    for (int i = 0; i < 3; ++i) ln[i] = lc * (-10.0 * E * e1) * nv[s][i] * nv[s][i] * nv[s][i];
                    for (int j = 0; j < 3; ++j) lv[j] = ln[0] * Rm[3 * j] + ln[1] * Rm[3 * j + 1] + ln[2] * Rm[3 * j + 2];
                    const double vl = lv[0] * vv[s][0] + lv[1] * vv[s][1] + lv[2] * vv[s][2];
                    for (int l = 0; l < 3; ++l) lam_gr[GRAD ? x : 0][l] += -(lv[l] - vl * vv[s][l]) / len[s];
                }
            }
        }
    }
}


}  // namespace mpf
