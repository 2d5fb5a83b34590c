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
// The common case (merged list of at most 8 entries) runs fully unrolled in registers. Cells whose
// merged list is longer (junctions of many grains) are deferred to a work list and finished by a
// second kernel that holds up to 16 entries per cell in private memory.
//
// The per-grain volume sums (PhaseField::CalculateGrainsVolume, src/PhaseField.cpp:812-951) are
// accumulated exactly as 128-bit fixed point in 32-bit limbs, so they are deterministic.
//
// Everything is built with -ffp-contract=off: the FP64 path performs the reference's operations
// in the reference's order, so active-phase sets match bit for bit.

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "mpf_device.h"

namespace mpf {

namespace {

constexpr int kFastSlots = 8;    // merged-list capacity of the register path
constexpr int kWideSlots = 16;   // merged-list capacity of the deferred path

template <int K> __host__ __device__ constexpr int pair_index(int a, int b)
{
    return a * K - a * (a + 1) / 2 + (b - a - 1);
}

__device__ __forceinline__ int wrap1(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

template <typename T> __device__ __forceinline__ T load_val(const void* p, int64_t i)
{
    return static_cast<const T*>(p)[i];
}

// std::min / std::max semantics (first argument wins ties and unordered compares).
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// Exact fixed-point image V = v * 2^116 of a value in [2^-64, 2) as four 32-bit limbs.
__device__ __forceinline__ void value_limbs(double v, unsigned long long l[kLimbs], bool& bad)
{
    unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
    int e = static_cast<int>((bits >> 52) & 0x7ff);
    l[0] = l[1] = l[2] = l[3] = 0ull;
    if ((bits >> 63) || e == 0x7ff) { bad = true; return; }
    if (e == 0) return;  // zero or subnormal: below the fixed-point resolution
    unsigned long long mant = (bits & 0xFFFFFFFFFFFFFull) | (1ull << 52);
    int shift = e - 959;  // (e - 1023) - 52 + 116
    unsigned long long lo, hi;
    if (shift > 64) { bad = true; return; }
    if (shift < 0) {
        lo = shift > -64 ? (mant >> (-shift)) : 0ull;
        hi = 0ull;
    } else if (shift == 0) {
        lo = mant; hi = 0ull;
    } else if (shift < 64) {
        lo = mant << shift; hi = mant >> (64 - shift);
    } else {
        lo = 0ull; hi = mant;
    }
    l[0] = lo & 0xFFFFFFFFull;
    l[1] = lo >> 32;
    l[2] = hi & 0xFFFFFFFFull;
    l[3] = hi >> 32;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// Adds the (grain, value) entries of all lanes of the wavefront to the grain accumulators: one
// atomic per distinct grain and limb instead of one per entry. Must be reached by whole waves.
__device__ __forceinline__ void wave_volume_add(bool act, uint32_t g, double v,
                                                unsigned long long* acc, bool& bad)
{
    unsigned long long l[kLimbs];
    if (act) value_limbs(v, l, bad);
    else l[0] = l[1] = l[2] = l[3] = 0ull;
    const int lane = static_cast<int>(threadIdx.x) & (warpSize - 1);
    while (true) {
        unsigned long long m = __ballot(act);
        if (m == 0ull) break;
        int leader = __ffsll(static_cast<long long>(m)) - 1;
        uint32_t g0 = static_cast<uint32_t>(__shfl(static_cast<int>(g), leader));
        bool sel = act && (g == g0);
#pragma unroll
        for (int i = 0; i < kLimbs; ++i) {
            unsigned long long s = wave_sum(sel ? l[i] : 0ull);
            if (lane == leader && s) atomicAdd(&acc[static_cast<int64_t>(g0) * kLimbs + i], s);
        }
        act = act && !sel;
    }
}

// InterfaceOrientation (src/InterfaceProperties.cpp:39-45) and the cubic energy of
// InterfaceEnergyModel::Calculate (include/InterfaceProperties/InterfaceEnergyModel.h:325-345).
template <typename TC>
__device__ __forceinline__ TC cubic_energy(const double* R, TC gx, TC gy, TC gz, TC energy, TC eps1)
{
    TC v0 = gx * TC(-1.0), v1 = gy * TC(-1.0), v2 = gz * TC(-1.0);
    TC len = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (len > TC(DBL_EPSILON)) {
        TC inv = TC(1.0) / len;
        v0 *= inv; v1 *= inv; v2 *= inv;
    } else {
        v0 = v1 = v2 = TC(0.0);
    }
    TC n0 = TC(R[0]) * v0 + TC(R[3]) * v1 + TC(R[6]) * v2;
    TC n1 = TC(R[1]) * v0 + TC(R[4]) * v1 + TC(R[7]) * v2;
    TC n2 = TC(R[2]) * v0 + TC(R[5]) * v1 + TC(R[8]) * v2;
    TC q0 = n0 * n0, q1 = n1 * n1, q2 = n2 * n2;
    return energy * (TC(1.0) + eps1 * (TC(1.5) - TC(2.5) * (q0 * q0 + q1 * q1 + q2 * q2)));
}

// finalize clamp/erase/renormalize of the first n entries (NodePF::finalize)
template <int K, int U>
__device__ __forceinline__ void finalize_list(int n, double (&phi)[K], bool (&keep)[K], int& ns, bool& empty)
{
    double total = 0.0;
    ns = 0;
#pragma unroll U
    for (int t = 0; t < K; ++t) {
        keep[t] = false;
        if (t >= n) continue;
        double v = phi[t];
        bool erase = false;
        if (v >= 1.0 - DBL_EPSILON) v = 1.0;
        if (v <= 0.0 + DBL_EPSILON) { v = 0.0; erase = true; }
        total += v;
        phi[t] = v;
        keep[t] = !erase;
        ns += erase ? 0 : 1;
    }
    empty = ns == 0;
    if (ns > 1) {
        const double norm = 1.0 / total;
#pragma unroll U
        for (int t = 0; t < K; ++t) if (keep[t]) phi[t] *= norm;
    } else if (ns == 1) {
#pragma unroll U
        for (int t = 0; t < K; ++t) if (keep[t]) phi[t] = 1.0;
    }
}

enum CellResult { kCellDone = 0, kCellDeferred = 1, kCellOverflow = 2 };

// Advances cell cc by one step with a merged-list capacity K. On kCellDone, eid/phi/keep hold the
// finalized list (in order) and the next state is written; on kCellDeferred (merged list longer
// than K) nothing is written.
template <int K, typename TV, typename TC, bool GRAD, bool DG>
__device__ __forceinline__ int advance_cell(const StepArgs& A, int64_t cc, uint32_t (&eid)[K],
                                            double (&phi)[K], bool (&keep)[K], bool& empty)
{
    constexpr int U = K <= kFastSlots ? K : 1;                 // full unroll only in registers
    constexpr int NPK = K * (K - 1) / 2;
    const int64_t N = A.ncells;
    const int NY = A.ny, NZ = A.nz;
    const int x = static_cast<int>(cc / (static_cast<int64_t>(NY) * NZ));
    const int y = static_cast<int>((cc / NZ) % NY);
    const int z = static_cast<int>(cc % NZ);
    const int n0 = A.cnt_in[cc];
    empty = false;
    if (n0 > K) return kCellDeferred;

    // wide_interface(): flag 2 (more than one entry) or a flag-2 cell in the 3^d box
    bool wide = n0 > 1;
    if (!wide) {
        for (int ii = -A.dnx; ii <= A.dnx && !wide; ++ii) {
            const int xx = wrap1(x + ii, A.nx);
            for (int jj = -A.dny; jj <= A.dny && !wide; ++jj) {
                const int yy = wrap1(y + jj, NY);
                for (int kk = -A.dnz; kk <= A.dnz; ++kk) {
                    const int zz = wrap1(z + kk, NZ);
                    const int64_t nb = (static_cast<int64_t>(xx) * NY + yy) * NZ + zz;
                    if (A.cnt_in[nb] > 1) { wide = true; break; }
                }
            }
        }
    }

    TC ev[K];
    TC lap[K];
    TC gr[GRAD ? K : 1][3];
#pragma unroll U
    for (int t = 0; t < K; ++t) {
        eid[t] = 0u; phi[t] = 0.0; ev[t] = TC(0); lap[t] = TC(0); keep[t] = false;
        if (GRAD) { gr[GRAD ? t : 0][0] = gr[GRAD ? t : 0][1] = gr[GRAD ? t : 0][2] = TC(0); }
        if (t < n0) {
            const TV v = load_val<TV>(A.val_in, static_cast<int64_t>(t) * N + cc);
            eid[t] = A.id_in[static_cast<int64_t>(t) * N + cc];
            phi[t] = static_cast<double>(v);
            ev[t] = static_cast<TC>(v);
            keep[t] = true;
        }
    }
    if (!wide) return kCellDone;  // bulk cell: unchanged

    // ---- derivatives + neighbour-list merge, stencil order x outer, z inner
    int ne = n0;
    for (int ii = -A.dnx; ii <= A.dnx; ++ii) {
        const int xx = wrap1(x + ii, A.nx);
        for (int jj = -A.dny; jj <= A.dny; ++jj) {
            const int yy = wrap1(y + jj, NY);
            for (int kk = -A.dnz; kk <= A.dnz; ++kk) {
                const int zz = wrap1(z + kk, NZ);
                const int64_t nb = (static_cast<int64_t>(xx) * NY + yy) * NZ + zz;
                const int nnz = (ii != 0) + (jj != 0) + (kk != 0);
                const TC w = static_cast<TC>(A.wl[nnz]);
                const TC gw = static_cast<TC>(A.wg[nnz]);
                const TC gwx = ii == 0 ? TC(0) : (ii < 0 ? -gw : gw);
                const TC gwy = jj == 0 ? TC(0) : (jj < 0 ? -gw : gw);
                const TC gwz = kk == 0 ? TC(0) : (kk < 0 ? -gw : gw);
                const int m = A.cnt_in[nb];
                for (int q = 0; q < m; ++q) {
                    const TC gv = static_cast<TC>(load_val<TV>(A.val_in, static_cast<int64_t>(q) * N + nb));
                    if (gv == TC(0)) continue;
                    const uint32_t gid = A.id_in[static_cast<int64_t>(q) * N + nb];
                    const TC l = w * gv;
                    bool found = false;
#pragma unroll U
                    for (int t = 0; t < K; ++t)
                        if (t < ne && eid[t] == gid) {
                            lap[t] += l;
                            if (GRAD && nnz) {
                                gr[GRAD ? t : 0][0] = gr[GRAD ? t : 0][0] + gwx * gv;
                                gr[GRAD ? t : 0][1] = gr[GRAD ? t : 0][1] + gwy * gv;
                                gr[GRAD ? t : 0][2] = gr[GRAD ? t : 0][2] + gwz * gv;
                            }
                            found = true;
                        }
                    if (!found) {
                        if (ne >= K) return kCellDeferred;
#pragma unroll U
                        for (int t = 0; t < K; ++t)
                            if (t == ne) {
                                eid[t] = gid;
                                phi[t] = 0.0;
                                ev[t] = TC(0);
                                lap[t] = l;
                                if (GRAD) {
                                    gr[GRAD ? t : 0][0] = nnz ? gwx * gv : TC(0);
                                    gr[GRAD ? t : 0][1] = nnz ? gwy * gv : TC(0);
                                    gr[GRAD ? t : 0][2] = nnz ? gwz * gv : TC(0);
                                }
                            }
                        ++ne;
                    }
                }
            }
        }
    }

    // ---- per-entry grain data
    int ph[K];
    double vr[K];
#pragma unroll U
    for (int t = 0; t < K; ++t) {
        ph[t] = 0; vr[t] = 0.0;
        if (t < ne) { ph[t] = A.gphase[eid[t]]; vr[t] = A.gvr[eid[t]]; }
    }
    double localn = 0.0;
    if (A.consider_nucleus) {
#pragma unroll U
        for (int t = 0; t < K; ++t) if (t < ne) localn += vr[t];
    } else {
        localn = static_cast<double>(ne);
    }
    const double norm_1 = 1.0 / localn;
    const int NP = A.nphases;

    // ---- interface energies of the pairs (ISO: phase-pair table; CUBIC: orientation dependent)
    TC sig[GRAD ? NPK : 1];
    if (GRAD) {
#pragma unroll U
        for (int a = 0; a < K; ++a)
#pragma unroll U
        for (int b = 0; b < K; ++b) {
            if (b <= a || b >= ne) continue;
            const int k = ph[a] * NP + ph[b];
            TC en = TC(0);
            if (A.emodel[k] == 0) {
                en = static_cast<TC>(A.sigma[k]);
            } else {
                const TC E = static_cast<TC>(A.sigma[k]), e1 = static_cast<TC>(A.eps[k]);
                TC wa = TC(0), wb = TC(0);
                if (A.gsolid[eid[a]]) {
                    wa = fabs(ev[a] * (TC(1.0) - ev[a]));
                    if (wa > TC(DBL_EPSILON))
                        en += cubic_energy<TC>(A.grot + 9 * static_cast<int64_t>(eid[a]), gr[GRAD ? a : 0][0],
                                               gr[GRAD ? a : 0][1], gr[GRAD ? a : 0][2], E, e1) * wa;
                    else
                        wa = TC(0);
                }
                if (A.gsolid[eid[b]]) {
                    wb = fabs(ev[b] * (TC(1.0) - ev[b]));
                    if (wb > TC(DBL_EPSILON))
                        en += cubic_energy<TC>(A.grot + 9 * static_cast<int64_t>(eid[b]), gr[GRAD ? b : 0][0],
                                               gr[GRAD ? b : 0][1], gr[GRAD ? b : 0][2], E, e1) * wb;
                    else
                        wb = TC(0);
                }
                if (wa + wb > TC(DBL_EPSILON)) en /= wa + wb;
                else en = static_cast<TC>(A.max_energy[k]);
            }
            sig[GRAD ? pair_index<K>(a, b) : 0] = en;
        }
    }
    auto sigma_of = [&](int a, int b) -> TC {
        if (GRAD) return a < b ? sig[GRAD ? pair_index<K>(a, b) : 0] : sig[GRAD ? pair_index<K>(b, a) : 0];
        return static_cast<TC>(A.sigma[ph[a] * NP + ph[b]]);
    };

    // ---- double-obstacle pair increments
    const TC P = static_cast<TC>(A.prefactor);
    double psi[NPK];
#pragma unroll U
    for (int p = 0; p < NPK; ++p) psi[p] = 0.0;
#pragma unroll U
    for (int a = 0; a < K; ++a)
#pragma unroll U
    for (int b = 0; b < K; ++b) {
        if (b <= a || b >= ne) continue;
        const double scaled = sqrt(vr[a] * vr[b]);
        const TC scale = static_cast<TC>(scaled);
        const TC cf = TC(0.5) * (TC(1.0) + scale) * P;
        TC d = sigma_of(a, b) * ((lap[a] + cf * ev[a]) - (lap[b] + cf * ev[b]));
        if (ne > 2) {
#pragma unroll U
            for (int g = 0; g < K; ++g) {
                if (g == a || g == b || g >= ne) continue;
                d += (sigma_of(b, g) - sigma_of(a, g)) * (lap[g] + cf * ev[g]);
                if (A.tj_model == 0) {
                    if (A.tj_factor != 0.0 || scaled < 1.0) {
                        const TC e3 = static_cast<TC>(A.sigma[ph[a] * NP + ph[b]]) +
                                      static_cast<TC>(A.sigma[ph[b] * NP + ph[g]]) +
                                      static_cast<TC>(A.sigma[ph[a] * NP + ph[g]]);
                        d += static_cast<TC>(smax(A.tj_factor, (1.0 - scaled))) * P * e3 *
                             (ev[a] * ev[g] - ev[b] * ev[g]);
                    }
                } else if (A.tj_model == 1) {
                    d += static_cast<TC>(A.tj_energy) * P * (ev[a] * ev[g] - ev[b] * ev[g]);
                }
            }
        }
        const TC mob = static_cast<TC>(A.mu[ph[a] * NP + ph[b]]);
        d *= mob * static_cast<TC>(norm_1) * scale;
        // constant driving force on interface cells (flag 2 = more than one own entry)
        if (DG && n0 > 1 && A.dg_mode[ph[a] * NP + ph[b]] == 1) {
            const TC raw = static_cast<TC>(A.dg_value[ph[a] * NP + ph[b]]);
            const TC weight = sqrt(ev[a] * ev[b]);
            const TC allowed = static_cast<TC>(A.dg_allowed[ph[a] * NP + ph[b]]);
            const TC dgl = allowed * tanh(raw / allowed);
            d += dgl * mob * weight * static_cast<TC>(A.dg_prefactor);
        }
        psi[pair_index<K>(a, b)] = static_cast<double>(d);
    }

    // ---- normalization / limiter (FP64); alive[p]: pair still in the increment list
    const double dt = A.dt;
    bool alive[NPK];
#pragma unroll U
    for (int p = 0; p < NPK; ++p) alive[p] = false;
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
        for (int a = 0; a < K; ++a) {
            if (a >= ne) continue;
            double dpsi = 0.0;
#pragma unroll U
            for (int b = 0; b < K; ++b) {
                if (b == a || b >= ne) continue;
                dpsi += a < b ? psi[pair_index<K>(a, b)] : psi[pair_index<K>(b, a)] * (-1.0);
            }
            if ((phi[a] == 0.0 && dpsi < 0.0) || (phi[a] == 1.0 && dpsi > 0.0)) {
#pragma unroll U
                for (int b = 0; b < K; ++b) {
                    if (b == a || b >= ne) continue;
                    psi[a < b ? pair_index<K>(a, b) : pair_index<K>(b, a)] = 0.0;
                }
            }
        }
        bool any = false;
#pragma unroll U
        for (int a = 0; a < K; ++a)
#pragma unroll U
        for (int b = 0; b < K; ++b) {
            if (b <= a || b >= ne) continue;
            const int p = pair_index<K>(a, b);
            alive[p] = psi[p] != 0.0;
            any = any || alive[p];
        }

        if (any) {
            bool needed = true;
            int iter = 0;
            while (needed) {
                ++iter;
                needed = false;
                double pos[K], neg[K], lim1[K], lim2[K];
#pragma unroll U
                for (int t = 0; t < K; ++t) { pos[t] = 0.0; neg[t] = 0.0; }
#pragma unroll U
                for (int a = 0; a < K; ++a)
#pragma unroll U
                for (int b = 0; b < K; ++b) {
                    if (b <= a || b >= ne) continue;
                    const int p = pair_index<K>(a, b);
                    if (!alive[p]) continue;
                    const double v = psi[p];
                    if (v < 0.0) { neg[a] += v; pos[b] += -v; }
                    if (v > 0.0) { pos[a] += v; neg[b] += -v; }
                }
#pragma unroll U
                for (int t = 0; t < K; ++t) {
                    lim1[t] = 1.0; lim2[t] = 1.0;
                    if (t >= ne) continue;
                    const double npf = phi[t] + (pos[t] + neg[t]) * dt;
                    if (npf < 0.0) lim2[t] = smin(lim2[t], -(phi[t] + pos[t] * dt) / (neg[t] * dt));
                    if (npf > 1.0) lim1[t] = smin(lim1[t], (1.0 - (phi[t] + neg[t] * dt)) / (pos[t] * dt));
                }
#pragma unroll U
                for (int a = 0; a < K; ++a)
#pragma unroll U
                for (int b = 0; b < K; ++b) {
                    if (b <= a || b >= ne) continue;
                    const int p = pair_index<K>(a, b);
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
            for (int p = 0; p < NPK; ++p)
                if (alive[p] && fabs(psi[p] * dt) < DBL_EPSILON) alive[p] = false;

            // plausibility check of the limited increments (warning count)
            double tv[K];
#pragma unroll U
            for (int t = 0; t < K; ++t) tv[t] = phi[t];
#pragma unroll U
            for (int a = 0; a < K; ++a)
#pragma unroll U
            for (int b = 0; b < K; ++b) {
                if (b <= a || b >= ne) continue;
                const int p = pair_index<K>(a, b);
                if (!alive[p] || psi[p] == 0.0) continue;
                tv[a] += psi[p] * dt;
                tv[b] += -psi[p] * dt;
            }
            unsigned long long warn = 0;
#pragma unroll U
            for (int t = 0; t < K; ++t)
                if (t < ne && (tv[t] < -FLT_EPSILON || tv[t] > 1.0 + FLT_EPSILON)) ++warn;
            if (warn) atomicAdd(A.warnings, warn);
        }
    }

    // ---- merge
#pragma unroll U
    for (int a = 0; a < K; ++a)
#pragma unroll U
    for (int b = 0; b < K; ++b) {
        if (b <= a || b >= ne) continue;
        const int p = pair_index<K>(a, b);
        if (!alive[p]) continue;
        const double value = 1.0 * psi[p] * dt;
        if (fabs(value) >= DBL_EPSILON) {
            phi[a] += value;
            phi[b] += -value;
        }
    }

    // ---- finalize: clamp, erase, renormalize
    int ns = 0;
    finalize_list<K, U>(ne, phi, keep, ns, empty);
    if (ns > A.slots) return kCellOverflow;
    return kCellDone;
}

// Writes the finalized list of cell cc (compacted, order preserved); rounds phi to storage type.
template <int K, typename TV>
__device__ __forceinline__ void store_cell(const StepArgs& A, int64_t cc, const uint32_t (&eid)[K],
                                           double (&phi)[K], const bool (&keep)[K])
{
    constexpr int U = K <= kFastSlots ? K : 1;
    const int64_t N = A.ncells;
    int w = 0;
#pragma unroll U
    for (int t = 0; t < K; ++t)
        if (keep[t]) {
            const TV sv = static_cast<TV>(phi[t]);
            phi[t] = static_cast<double>(sv);
            A.id_out[static_cast<int64_t>(w) * N + cc] = static_cast<uint16_t>(eid[t]);
            static_cast<TV*>(A.val_out)[static_cast<int64_t>(w) * N + cc] = sv;
            ++w;
        }
    A.cnt_out[cc] = static_cast<uint8_t>(w);
}

template <typename TV, typename TC, bool GRAD, bool DG>
__global__ __launch_bounds__(kBlock) void step_kernel(const StepArgs A)
{
    constexpr int K = kFastSlots;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool valid = c < A.ncells;
    uint32_t eid[K];
    double phi[K];
    bool keep[K];
    bool empty = false, bad_volume = false;
#pragma unroll
    for (int t = 0; t < K; ++t) { eid[t] = 0u; phi[t] = 0.0; keep[t] = false; }
    if (valid) {
        const int r = advance_cell<K, TV, TC, GRAD, DG>(A, c, eid, phi, keep, empty);
        if (r == kCellDone) {
            store_cell<K, TV>(A, c, eid, phi, keep);
        } else {
#pragma unroll
            for (int t = 0; t < K; ++t) keep[t] = false;
            if (r == kCellDeferred) {
                const unsigned slot = atomicAdd(A.work_count, 1u);
                A.work_list[slot] = static_cast<uint32_t>(c);
            } else {
                atomicOr(&A.status[kStatusOverflow], 1);
                atomicMin(A.first_overflow, static_cast<unsigned long long>(c));
            }
        }
        if (empty) atomicOr(&A.status[kStatusEmpty], 1);
    }
#pragma unroll
    for (int t = 0; t < K; ++t) wave_volume_add(keep[t], eid[t], phi[t], A.vol_acc, bad_volume);
    if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
}

// Finishes the cells the register kernel deferred (merged list longer than 8 entries).
template <typename TV, typename TC, bool GRAD, bool DG>
__global__ __launch_bounds__(kWideBlock) void wide_kernel(const StepArgs A)
{
    constexpr int K = kWideSlots;
    const unsigned n = *A.work_count;
    const unsigned stride = gridDim.x * kWideBlock;
    for (unsigned base = blockIdx.x * kWideBlock; base < n; base += stride) {
        const unsigned i = base + threadIdx.x;
        const bool valid = i < n;
        uint32_t eid[K];
        double phi[K];
        bool keep[K];
        bool empty = false, bad_volume = false;
        for (int t = 0; t < K; ++t) { eid[t] = 0u; phi[t] = 0.0; keep[t] = false; }
        if (valid) {
            const int64_t c = A.work_list[i];
            const int r = advance_cell<K, TV, TC, GRAD, DG>(A, c, eid, phi, keep, empty);
            if (r == kCellDone) {
                store_cell<K, TV>(A, c, eid, phi, keep);
            } else {
                for (int t = 0; t < K; ++t) keep[t] = false;
                atomicOr(&A.status[kStatusOverflow], 1);
                atomicMin(A.first_overflow, static_cast<unsigned long long>(c));
            }
            if (empty) atomicOr(&A.status[kStatusEmpty], 1);
        }
        for (int t = 0; t < K; ++t) wave_volume_add(keep[t], eid[t], phi[t], A.vol_acc, bad_volume);
        if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
    }
}

// PhaseField::FinalizeInitialization: NodePF::finalize on flag-2 cells (more than one entry after
// a raw read) plus the volume sums; derivatives and flags are rebuilt by the next step kernel.
template <typename TV>
__global__ __launch_bounds__(kBlock) void finalize_init_kernel(const StepArgs A)
{
    constexpr int K = kFastSlots;
    const int64_t N = A.ncells;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool valid = c < N;
    const int64_t cc = valid ? c : 0;
    const int n0 = valid ? A.cnt_in[cc] : 0;
    uint32_t eid[K];
    double phi[K];
    bool keep[K];
    bool bad_volume = false, empty = false;
#pragma unroll
    for (int t = 0; t < K; ++t) {
        eid[t] = 0u; phi[t] = 0.0; keep[t] = false;
        if (t < n0) {
            eid[t] = A.id_in[static_cast<int64_t>(t) * N + cc];
            phi[t] = static_cast<double>(load_val<TV>(A.val_in, static_cast<int64_t>(t) * N + cc));
            keep[t] = true;
        }
    }
    if (valid && n0 > 1) {
        int ns = 0;
        finalize_list<K, K>(n0, phi, keep, ns, empty);
        if (empty) atomicOr(&A.status[kStatusEmpty], 1);
    }
    if (valid) store_cell<K, TV>(A, cc, eid, phi, keep);
#pragma unroll
    for (int t = 0; t < K; ++t) wave_volume_add(keep[t], eid[t], phi[t], A.vol_acc, bad_volume);
    if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
}

// Correctly rounded V * 2^-116 of the exact 4-limb sum.
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
__global__ void grains_kernel(const GrainArgs G)
{
    const int g = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g == 0) *G.work_count = 0u;
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

}  // namespace

// Two variants per precision: the plain isotropic step, and the full one (orientation-dependent
// energies and constant driving force; with ISO tables or no driving-force pairs it reduces to
// the plain arithmetic at run time).
int launch_step(const StepArgs& a, bool fp32, bool full, void* stream)
{
    const dim3 grid(static_cast<unsigned>((a.ncells + kBlock - 1) / kBlock));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (fp32) {
        if (full) hipLaunchKernelGGL((step_kernel<float, float, true, true>), grid, dim3(kBlock), 0, s, a);
        else      hipLaunchKernelGGL((step_kernel<float, float, false, false>), grid, dim3(kBlock), 0, s, a);
    } else {
        if (full) hipLaunchKernelGGL((step_kernel<double, double, true, true>), grid, dim3(kBlock), 0, s, a);
        else      hipLaunchKernelGGL((step_kernel<double, double, false, false>), grid, dim3(kBlock), 0, s, a);
    }
    return static_cast<int>(hipGetLastError());
}

int launch_wide(const StepArgs& a, bool fp32, bool full, void* stream)
{
    const dim3 grid(kWideGrid);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (fp32) {
        if (full) hipLaunchKernelGGL((wide_kernel<float, float, true, true>), grid, dim3(kWideBlock), 0, s, a);
        else      hipLaunchKernelGGL((wide_kernel<float, float, false, false>), grid, dim3(kWideBlock), 0, s, a);
    } else {
        if (full) hipLaunchKernelGGL((wide_kernel<double, double, true, true>), grid, dim3(kWideBlock), 0, s, a);
        else      hipLaunchKernelGGL((wide_kernel<double, double, false, false>), grid, dim3(kWideBlock), 0, s, a);
    }
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

int launch_grains(const GrainArgs& g, void* stream)
{
    const int n = g.ngrains > 0 ? g.ngrains : 1;
    const dim3 grid(static_cast<unsigned>((n + 255) / 256));
    hipLaunchKernelGGL(grains_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream), g);
    return static_cast<int>(hipGetLastError());
}

}  // namespace mpf
