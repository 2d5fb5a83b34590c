// The update of one interface cell as a function of its neighbour-merged list, over a scalar type
// T: double reproduces the step kernels' FP64 arithmetic operation for operation; Dual<D> carries
// D directional derivatives alongside (forward mode), which the adjoint kernels contract with the
// cotangent of the next state.
//
// Inputs per merged entry e < ne: grain, value phi_e (0 for entries appended from neighbours),
// Laplacian lap_e and gradient gr_e (CalculateDerivativesSR, src/PhaseField.cpp:566-615).
// Pipeline, as in the step kernels: pair energies (InterfaceProperties::SetSR,
// src/InterfaceProperties.cpp:603-762), double-obstacle pair increments
// (src/DoubleObstacle.cpp:502-571) with the constant driving force (src/DrivingForce.cpp:497-602),
// the sum-to-one limiter (PhaseField::NormalizeIncrementsSR, src/PhaseField.cpp:1351-1567), the
// merge (:1268-1299) and NodePF::finalize (include/Containers/NodePF.h:214-255).
//
// Derivatives are those of the branch taken (active-set masks): a clamped or erased entry has
// derivative 0, limiter factors are differentiated as evaluated, VolumeRatio is held fixed.
#pragma once

#include <float.h>
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define MPF_HD __host__ __device__
#else
#define MPF_HD
#endif

namespace mpf {

constexpr int kCellMax = 8;      // merged entries handled by the cell function
constexpr int kPhaseMax = 4;     // thermodynamic phases handled by the adjoint

// ---------------------------------------------------------------- scalar types
template <int D> struct Dual {
    double v;
    double d[D];
    MPF_HD Dual() : v(0.0) { for (int i = 0; i < D; ++i) d[i] = 0.0; }
    MPF_HD Dual(double x) : v(x) { for (int i = 0; i < D; ++i) d[i] = 0.0; }
};

MPF_HD inline double value_of(double x) { return x; }
template <int D> MPF_HD inline double value_of(const Dual<D>& x) { return x.v; }

template <int D> MPF_HD inline Dual<D> operator+(const Dual<D>& a, const Dual<D>& b)
{
    Dual<D> r; r.v = a.v + b.v;
    for (int i = 0; i < D; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int D> MPF_HD inline Dual<D> operator-(const Dual<D>& a, const Dual<D>& b)
{
    Dual<D> r; r.v = a.v - b.v;
    for (int i = 0; i < D; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int D> MPF_HD inline Dual<D> operator-(const Dual<D>& a)
{
    Dual<D> r; r.v = -a.v;
    for (int i = 0; i < D; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int D> MPF_HD inline Dual<D> operator*(const Dual<D>& a, const Dual<D>& b)
{
    Dual<D> r; r.v = a.v * b.v;
    for (int i = 0; i < D; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int D> MPF_HD inline Dual<D> operator/(const Dual<D>& a, const Dual<D>& b)
{
    Dual<D> r; r.v = a.v / b.v;
    for (int i = 0; i < D; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
template <int D> MPF_HD inline Dual<D> operator+(const Dual<D>& a, double b) { return a + Dual<D>(b); }
template <int D> MPF_HD inline Dual<D> operator-(const Dual<D>& a, double b) { return a - Dual<D>(b); }
template <int D> MPF_HD inline Dual<D> operator*(const Dual<D>& a, double b) { return a * Dual<D>(b); }
template <int D> MPF_HD inline Dual<D> operator/(const Dual<D>& a, double b) { return a / Dual<D>(b); }
template <int D> MPF_HD inline Dual<D> operator+(double a, const Dual<D>& b) { return Dual<D>(a) + b; }
template <int D> MPF_HD inline Dual<D> operator-(double a, const Dual<D>& b) { return Dual<D>(a) - b; }
template <int D> MPF_HD inline Dual<D> operator*(double a, const Dual<D>& b) { return Dual<D>(a) * b; }
template <int D> MPF_HD inline Dual<D> operator/(double a, const Dual<D>& b) { return Dual<D>(a) / b; }
template <int D> MPF_HD inline Dual<D>& operator+=(Dual<D>& a, const Dual<D>& b) { a = a + b; return a; }
template <int D> MPF_HD inline Dual<D>& operator*=(Dual<D>& a, const Dual<D>& b) { a = a * b; return a; }
template <int D> MPF_HD inline Dual<D>& operator/=(Dual<D>& a, const Dual<D>& b) { a = a / b; return a; }

MPF_HD inline double d_sqrt(double x) { return sqrt(x); }
MPF_HD inline double d_tanh(double x) { return tanh(x); }
MPF_HD inline double d_fabs(double x) { return fabs(x); }
template <int D> MPF_HD inline Dual<D> d_sqrt(const Dual<D>& a)
{
    Dual<D> r; r.v = sqrt(a.v);
    // sqrt(phi_a phi_b) with an appended (structurally zero) entry: no derivative
    for (int i = 0; i < D; ++i) r.d[i] = r.v > 0.0 ? a.d[i] / (2.0 * r.v) : 0.0;
    return r;
}
template <int D> MPF_HD inline Dual<D> d_tanh(const Dual<D>& a)
{
    Dual<D> r; r.v = tanh(a.v);
    for (int i = 0; i < D; ++i) r.d[i] = a.d[i] * (1.0 - r.v * r.v);
    return r;
}
template <int D> MPF_HD inline Dual<D> d_fabs(const Dual<D>& a) { return a.v < 0.0 ? -a : a; }

// std::min / std::max semantics on values (first argument wins ties)
template <typename T> MPF_HD inline T t_min(const T& a, const T& b) { return (value_of(b) < value_of(a)) ? b : a; }
template <typename T> MPF_HD inline T t_max(const T& a, const T& b) { return (value_of(a) < value_of(b)) ? b : a; }

// ---------------------------------------------------------------- inputs
// Pair tables (nphases^2) in T: interface energy, mobility, driving force; derived tables (MaxEnergy
// and the tanh amplitude) are recomputed from them as the host does.
template <typename T> struct CellTables {
    int np;
    T sigma[kPhaseMax * kPhaseMax];
    T mu[kPhaseMax * kPhaseMax];
    T dg[kPhaseMax * kPhaseMax];
    const double* eps;          // cubic anisotropy strength per pair
    const int32_t* emodel;      // 0 ISO, 1 CUBIC
    const int32_t* dg_mode;     // 1 = constant driving force
    double dg_limit;            // DrivingForce Limit
    double regularization;
    double prefactor;           // pi^2 / eta^2
    double dg_prefactor;        // pi / eta
    double dt;
    int consider_nucleus;
    int tj_model;               // 0 MODEL, 1 VALUE, 2 NONE
    double tj_factor;
    double tj_energy;
    int dg_on;                  // driving force present
    int grad;                   // orientation-dependent energies present
};

template <typename T> struct CellState {
    int ne, n0;
    int ph[kCellMax];
    double vr[kCellMax];
    int solid[kCellMax];
    const double* rot[kCellMax];    // row-major rotation of the grain
    T phi[kCellMax];
    T lap[kCellMax];
    T gr[kCellMax][3];
    bool keep[kCellMax];            // out: entry survives finalize
};

// MaxEnergy of InterfaceEnergyModel::ReadInputIso/ReadInputCubic (InterfaceEnergyModel.h:78-94)
template <typename T> MPF_HD inline T max_energy_of(const CellTables<T>& P, int k)
{
    return P.emodel[k] == 1 ? P.sigma[k] * (1.0 + P.eps[k] * 1.41) : P.sigma[k];
}

// InterfaceOrientation (src/InterfaceProperties.cpp:39-45) and the cubic energy
// (include/InterfaceProperties/InterfaceEnergyModel.h:325-345)
template <typename T>
MPF_HD inline T cubic_energy_t(const double* R, const T& gx, const T& gy, const T& gz, const T& energy, double eps1)
{
    T v0 = gx * -1.0, v1 = gy * -1.0, v2 = gz * -1.0;
    const T len = d_sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (value_of(len) > DBL_EPSILON) {
        const T inv = 1.0 / len;
        v0 = v0 * inv; v1 = v1 * inv; v2 = v2 * inv;
    } else {
        v0 = T(0.0); v1 = T(0.0); v2 = T(0.0);
    }
    const T n0 = R[0] * v0 + R[3] * v1 + R[6] * v2;
    const T n1 = R[1] * v0 + R[4] * v1 + R[7] * v2;
    const T n2 = R[2] * v0 + R[5] * v1 + R[8] * v2;
    const T q0 = n0 * n0, q1 = n1 * n1, q2 = n2 * n2;
    return energy * (1.0 + eps1 * (1.5 - 2.5 * (q0 * q0 + q1 * q1 + q2 * q2)));
}

MPF_HD inline int pair_at(int a, int b) { return a * kCellMax - a * (a + 1) / 2 + (b - a - 1); }

// One cell: overwrites S.phi with the finalized values and sets S.keep. Returns the number of
// kept entries.
template <typename T> MPF_HD int cell_update(const CellTables<T>& P, CellState<T>& S)
{
    constexpr int NPM = kCellMax * (kCellMax - 1) / 2;
    const int ne = S.ne, NP = P.np;
    double localn = 0.0;
    if (P.consider_nucleus) {
        for (int t = 0; t < ne; ++t) localn += S.vr[t];
    } else {
        localn = static_cast<double>(ne);
    }
    const double norm_1 = 1.0 / localn;
    const double Pf = P.prefactor;

    // pair energies
    T sig[NPM];
    for (int a = 0; a < ne; ++a)
        for (int b = a + 1; b < ne; ++b) {
            const int k = S.ph[a] * NP + S.ph[b];
            T en(0.0);
            if (!P.grad || P.emodel[k] == 0) {
                en = P.sigma[k];
            } else {
                const T E = P.sigma[k];
                const double e1 = P.eps[k];
                T wa(0.0), wb(0.0);
                if (S.solid[a]) {
                    wa = d_fabs(S.phi[a] * (1.0 - S.phi[a]));
                    if (value_of(wa) > DBL_EPSILON)
                        en += cubic_energy_t<T>(S.rot[a], S.gr[a][0], S.gr[a][1], S.gr[a][2], E, e1) * wa;
                    else
                        wa = T(0.0);
                }
                if (S.solid[b]) {
                    wb = d_fabs(S.phi[b] * (1.0 - S.phi[b]));
                    if (value_of(wb) > DBL_EPSILON)
                        en += cubic_energy_t<T>(S.rot[b], S.gr[b][0], S.gr[b][1], S.gr[b][2], E, e1) * wb;
                    else
                        wb = T(0.0);
                }
                if (value_of(wa + wb) > DBL_EPSILON) en /= wa + wb;
                else en = max_energy_of(P, k);
            }
            sig[pair_at(a, b)] = en;
        }
    auto sigma_of = [&](int a, int b) -> T {
        if (P.grad) return a < b ? sig[pair_at(a, b)] : sig[pair_at(b, a)];
        return P.sigma[S.ph[a] * NP + S.ph[b]];
    };

    // double-obstacle pair increments (the general form; the equal-energy form of the step kernels
    // differs from it only by exact zero terms)
    T psi[NPM];
    for (int a = 0; a < ne; ++a)
        for (int b = a + 1; b < ne; ++b) {
            const double scaled = sqrt(S.vr[a] * S.vr[b]);
            const double scale = scaled;
            const double cf = 0.5 * (1.0 + scale) * Pf;
            T d = sigma_of(a, b) * ((S.lap[a] + cf * S.phi[a]) - (S.lap[b] + cf * S.phi[b]));
            if (ne > 2) {
                for (int g = 0; g < ne; ++g) {
                    if (g == a || g == b) continue;
                    d += (sigma_of(b, g) - sigma_of(a, g)) * (S.lap[g] + cf * S.phi[g]);
                    if (P.tj_model == 0) {
                        if (P.tj_factor != 0.0 || scaled < 1.0) {
                            const T e3 = P.sigma[S.ph[a] * NP + S.ph[b]] + P.sigma[S.ph[b] * NP + S.ph[g]] +
                                         P.sigma[S.ph[a] * NP + S.ph[g]];
                            const double f = (P.tj_factor < (1.0 - scaled)) ? (1.0 - scaled) : P.tj_factor;
                            d += f * Pf * e3 * (S.phi[a] * S.phi[g] - S.phi[b] * S.phi[g]);
                        }
                    } else if (P.tj_model == 1) {
                        d += P.tj_energy * Pf * (S.phi[a] * S.phi[g] - S.phi[b] * S.phi[g]);
                    }
                }
            }
            const int k = S.ph[a] * NP + S.ph[b];
            const T mob = P.mu[k];
            d *= mob * norm_1 * scale;
            if (P.dg_on && S.n0 > 1 && P.dg_mode[k] == 1) {
                const T weight = d_sqrt(S.phi[a] * S.phi[b]);
                T allowed = P.dg_limit * P.dg_prefactor * max_energy_of(P, k);
                allowed *= T(P.regularization);
                const T dgl = allowed * d_tanh(P.dg[k] / allowed);
                d += dgl * mob * weight * P.dg_prefactor;
            }
            psi[pair_at(a, b)] = d;
        }

    // normalization / limiter
    const double dt = P.dt;
    bool alive[NPM];
    for (int p = 0; p < NPM; ++p) alive[p] = false;
    if (ne == 2) {
        const T va = psi[0] * dt;
        const T old = S.phi[0];
        alive[0] = true;
        if ((value_of(old) == 0.0 && value_of(va) < 0.0) || (value_of(old) == 1.0 && value_of(va) > 0.0)) {
            alive[0] = false;
        } else {
            const T nv = old + va;
            T norm(1.0);
            if (value_of(nv) < 0.0) norm *= -old / va;
            else if (value_of(nv) > 1.0) norm *= (1.0 - old) / va;
            if (value_of(norm) > DBL_EPSILON) psi[0] *= norm;
            else alive[0] = false;
        }
    } else if (ne > 2) {
        for (int a = 0; a < ne; ++a) {
            T dpsi(0.0);
            for (int b = 0; b < ne; ++b) {
                if (b == a) continue;
                dpsi += a < b ? psi[pair_at(a, b)] : psi[pair_at(b, a)] * -1.0;
            }
            if ((value_of(S.phi[a]) == 0.0 && value_of(dpsi) < 0.0) || (value_of(S.phi[a]) == 1.0 && value_of(dpsi) > 0.0)) {
                for (int b = 0; b < ne; ++b) {
                    if (b == a) continue;
                    psi[a < b ? pair_at(a, b) : pair_at(b, a)] = T(0.0);
                }
            }
        }
        bool any = false;
        for (int a = 0; a < ne; ++a)
            for (int b = a + 1; b < ne; ++b) {
                const int p = pair_at(a, b);
                alive[p] = value_of(psi[p]) != 0.0;
                any = any || alive[p];
            }
        if (any) {
            bool needed = true;
            int iter = 0;
            while (needed) {
                ++iter;
                needed = false;
                T pos[kCellMax], neg[kCellMax], lim1[kCellMax], lim2[kCellMax];
                for (int t = 0; t < ne; ++t) { pos[t] = T(0.0); neg[t] = T(0.0); }
                for (int a = 0; a < ne; ++a)
                    for (int b = a + 1; b < ne; ++b) {
                        const int p = pair_at(a, b);
                        if (!alive[p]) continue;
                        const T v = psi[p];
                        if (value_of(v) < 0.0) { neg[a] += v; pos[b] += -v; }
                        if (value_of(v) > 0.0) { pos[a] += v; neg[b] += -v; }
                    }
                for (int t = 0; t < ne; ++t) {
                    lim1[t] = T(1.0); lim2[t] = T(1.0);
                    const T npf = S.phi[t] + (pos[t] + neg[t]) * dt;
                    if (value_of(npf) < 0.0) lim2[t] = t_min(lim2[t], -(S.phi[t] + pos[t] * dt) / (neg[t] * dt));
                    if (value_of(npf) > 1.0) lim1[t] = t_min(lim1[t], (1.0 - (S.phi[t] + neg[t] * dt)) / (pos[t] * dt));
                }
                for (int a = 0; a < ne; ++a)
                    for (int b = a + 1; b < ne; ++b) {
                        const int p = pair_at(a, b);
                        if (!alive[p]) continue;
                        if (value_of(psi[p]) < 0.0) {
                            const T f = t_min(lim2[a], lim1[b]);
                            psi[p] *= f;
                            if (value_of(f) < 1.0) needed = true;
                        }
                        if (value_of(psi[p]) > 0.0) {
                            const T f = t_min(lim1[a], lim2[b]);
                            psi[p] *= f;
                            if (value_of(f) < 1.0) needed = true;
                        }
                    }
                if (iter > 24) needed = false;
            }
            for (int a = 0; a < ne; ++a)
                for (int b = a + 1; b < ne; ++b) {
                    const int p = pair_at(a, b);
                    if (alive[p] && d_fabs(value_of(psi[p]) * dt) < DBL_EPSILON) alive[p] = false;
                }
        }
    }

    // merge
    T phi[kCellMax];
    for (int t = 0; t < ne; ++t) phi[t] = S.phi[t];
    for (int a = 0; a < ne; ++a)
        for (int b = a + 1; b < ne; ++b) {
            const int p = pair_at(a, b);
            if (!alive[p]) continue;
            const T value = 1.0 * psi[p] * dt;
            if (d_fabs(value_of(value)) >= DBL_EPSILON) {
                phi[a] += value;
                phi[b] += -value;
            }
        }

    // finalize: clamp, erase, renormalize
    T total(0.0);
    int ns = 0;
    for (int t = 0; t < ne; ++t) {
        T v = phi[t];
        bool erase = false;
        if (value_of(v) >= 1.0 - DBL_EPSILON) v = T(1.0);
        if (value_of(v) <= 0.0 + DBL_EPSILON) { v = T(0.0); erase = true; }
        total += v;
        phi[t] = v;
        S.keep[t] = !erase;
        ns += erase ? 0 : 1;
    }
    if (ns > 1) {
        const T norm = 1.0 / total;
        for (int t = 0; t < ne; ++t) if (S.keep[t]) phi[t] *= norm;
    } else if (ns == 1) {
        for (int t = 0; t < ne; ++t) if (S.keep[t]) phi[t] = T(1.0);
    }
    for (int t = 0; t < ne; ++t) S.phi[t] = phi[t];
    return ns;
}

}  // namespace mpf
