// MI355X (gfx950) kernels of the driving-force averaging over the interface
// (DrivingForce::Average with Averaging on, src/DrivingForce.cpp:242-255).
//
// The reference keeps a Force entry per VALUE-mode phase-field pair of every interface node
// (UserDrivingForce::SetDrivingForce, src/UserDrivingForce.cpp:152-175) with raw = Value(phaseA,
// phaseB) and weight = sqrt(phiA phiB) (SetWeights, :276-291). CollectAverage (:293-379) replaces
// raw by its weighted mean over the sphere of radius Range (tmp); DistributeAverage (:381-478)
// averages tmp again (average), which MergePhaseFieldIncrementsSR then limits and merges. Here no
// Force storage exists: an entry's raw value and weight follow from the node's own list (grains,
// values) and the phase-pair table, so both passes read the step-start state directly. A
// wavefront takes one interface cell at a time and, per own pair, the sphere's offsets x outer,
// z inner (the reference's loop order; every sum is formed in that order); a neighbour contributes
// when it is an interface cell holding both grains. tmp and the averaged value are stored per pair of own
// entries (a < b at plane b(b-1)/2 + a) for the step kernel, which reads them in place of the
// phase-pair value.
//
// Built with -ffp-contract=off like the step kernels: the FP64 sums are the reference's.

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "mpf_device.h"
#include "mpf_lists.h"

namespace mpf {

namespace {

__device__ __forceinline__ int pair_plane(int a, int b)  // a < b
{
    return b * (b - 1) / 2 + a;
}

// Value of own entry t of cell c (count code `code`). FP32 lists recover their implicit entry as
// 1 - (sum of the others in slot order), as the step does.
template <typename TV>
__device__ double entry_value(const StepArgs& A, int64_t c, uint8_t code, int t)
{
    const int64_t N = A.ncells;
    const int imp = ListCode<TV>::implicit(code);
    if (t != imp) return static_cast<double>(load_val<TV>(A.val_in, static_cast<int64_t>(t) * N + c));
    const int n = ListCode<TV>::count(code);
    double others = 0.0;
    for (int s = 0; s < n; ++s)
        if (s != imp) others += static_cast<double>(load_val<TV>(A.val_in, static_cast<int64_t>(s) * N + c));
    return 1.0 - others;
}

// NodeDF::get_weight of pair (ga, gb) at cell nb (include/Containers/NodeDF.h:401-410): a Force
// entry exists on an interface cell (more than one own entry) whose list holds both grains (the
// pair's phases give the same mode everywhere); entries of zero-valued appended grains weigh 0 and
// never count. On success: the weight, the pair plane and whether ga precedes gb in that cell's
// list (the entry's stored order; get_raw / get_tmp flip the sign otherwise).
template <typename TV>
__device__ bool find_pair(const StepArgs& A, int64_t nb, uint32_t ga, uint32_t gb, double& w, int& plane, bool& fwd)
{
    const int64_t N = A.ncells;
    const uint8_t code = A.cnt_in[nb];
    const int n = ListCode<TV>::count(code);
    if (n < 2) return false;
    int ia = -1, ib = -1;
    for (int s = 0; s < n; ++s) {
        const uint32_t g = A.id_in[static_cast<int64_t>(s) * N + nb];
        ia = g == ga ? s : ia;
        ib = g == gb ? s : ib;
    }
    if (ia < 0 || ib < 0) return false;
    w = sqrt(entry_value<TV>(A, nb, code, ia) * entry_value<TV>(A, nb, code, ib));
    fwd = ia < ib;
    plane = fwd ? pair_plane(ia, ib) : pair_plane(ib, ia);
    return true;
}

// Lane l's value of x (l uniform across the wavefront).
__device__ __forceinline__ double lane_value(double x, int l)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(b)), l));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(b >> 32)), l));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

constexpr int kAvgCells = 64;                 // cells per block
constexpr int kAvgWaves = kAvgBlock / 64;     // wavefronts per block

// SECOND = false: CollectAverage into tmp; true: DistributeAverage into avg. A block takes 64
// consecutive cells and compacts their interface cells in LDS; its wavefronts take them in turn.
// For one cell and pair the 64 lanes evaluate 64 sphere offsets at once (the neighbour's weight
// and raw or tmp value), then every lane adds the terms of the lanes that count in lane order,
// chunk after chunk: the sums are formed in the reference's offset order, and a wavefront waits
// on a few memory round trips per chunk instead of one chain per offset.
template <typename TV, bool SECOND>
__global__ __launch_bounds__(kAvgBlock) void dg_average_kernel(const AvgArgs P)
{
    __shared__ uint16_t s_list[kAvgCells];
    __shared__ int s_count;
    const StepArgs& A = P.s;
    const int64_t N = A.ncells;
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kAvgCells;
    if (wave == 0) {
        // Force entries live on interface cells (more than one own entry) only
        const int64_t c0 = base + lane;
        const bool inter = lane < kAvgCells && c0 < N && ListCode<TV>::count(A.cnt_in[c0]) >= 2;
        const unsigned long long m = __ballot(inter);
        if (inter) s_list[__popcll(m & ((1ull << lane) - 1ull))] = static_cast<uint16_t>(lane);
        if (lane == 0) s_count = __popcll(m);
    }
    __syncthreads();
    const int ninter = s_count;
    const int64_t nyz = static_cast<int64_t>(A.ny) * A.nz;
    const int NP = A.nphases;
    for (int i = wave; i < ninter; i += kAvgWaves) {
        const int64_t c = base + s_list[i];
        const uint8_t code = A.cnt_in[c];
        const int n = ListCode<TV>::count(code);
        const int x = static_cast<int>(c / nyz), y = static_cast<int>((c / A.nz) % A.ny), z = static_cast<int>(c % A.nz);
        for (int b = 1; b < n; ++b)
            for (int a = 0; a < b; ++a) {
                const uint32_t ga = A.id_in[static_cast<int64_t>(a) * N + c], gb = A.id_in[static_cast<int64_t>(b) * N + c];
                const int pa = A.gphase[ga], pb = A.gphase[gb];
                if (A.dg_mode[pa * NP + pb] != 1) continue;  // no Force entry for this pair
                const int plane = pair_plane(a, b);
                const double w = sqrt(entry_value<TV>(A, c, code, a) * entry_value<TV>(A, c, code, b));
                const double scale = A.gvr[ga] * A.gvr[gb];
                const double cut = P.threshold * scale;
                const double vab = A.dg_value[pa * NP + pb], vba = A.dg_value[pb * NP + pa];
                // own value the pass falls back to: raw (collect) or tmp (distribute)
                const double own = SECOND ? P.tmp[static_cast<int64_t>(plane) * N + c] : vab;
                double result = 0.0;  // a fresh entry's tmp / average
                if (SECOND ? (w > 0.0) : (w > cut)) {
                    double value = 0.0, sum = 0.0;
                    int counter = 0;
                    for (int t0 = 0; t0 < P.noff; t0 += 64) {
                        const int t = t0 + lane;
                        bool ok = false;
                        double wt = 0.0, term = 0.0;
                        if (t < P.noff) {
                            const int ii = P.off[3 * t], jj = P.off[3 * t + 1], kk = P.off[3 * t + 2];
                            const int64_t nb = (static_cast<int64_t>(wrap1(x + ii, A.nx)) * A.ny + wrap1(y + jj, A.ny)) * A.nz +
                                               wrap1(z + kk, A.nz);
                            double wphi = 0.0;
                            int np = 0;
                            bool fwd = true;
                            if (find_pair<TV>(A, nb, ga, gb, wphi, np, fwd) && (SECOND ? (wphi > cut) : (wphi > 0.0))) {
                                ok = true;
                                wt = P.weights == 1 ? P.off_w[t] : (P.weights == 0 ? wphi : 1.0);
                                double v;
                                if (SECOND) {
                                    const double tn = P.tmp[static_cast<int64_t>(np) * N + nb];
                                    v = fwd ? tn : -tn;
                                } else {
                                    v = fwd ? vab : -vba;
                                }
                                term = wt * v;
                            }
                        }
                        // value += weight * v and sum += weight in offset order
                        unsigned long long okm = __ballot(ok);
                        while (okm) {
                            const int l = __ffsll(static_cast<long long>(okm)) - 1;
                            okm &= okm - 1ull;
                            if (P.weights == 2) counter++;
                            else sum += lane_value(wt, l);
                            value += lane_value(term, l);
                        }
                    }
                    if (P.weights == 2) result = counter != 0 ? value / static_cast<double>(counter) : own;
                    else result = sum > 0.0 ? value / sum : own;
                }
                // DistributeAverage's weight-0 branch (Seed grains) cannot occur: own entries are
                // above the finalize cutoff and no grain is a Seed
                if (lane == 0) (SECOND ? P.avg : P.tmp)[static_cast<int64_t>(plane) * N + c] = result;
            }
    }
}

// ---- interface regularization (FP64 state) -------------------------------------------------

constexpr int kCurvBlock = 256;

// The curvature DoubleObstacle's regularization overload adds per pair of non-zero entries of a
// wide-interface cell (src/DoubleObstacle.cpp:702-763): loc_increment (the pair's increment before
// the mobility, over the neighbour-merged list) * Eta/Pi * scale * norm_1 / sqrt(phi_a phi_b) *
// (CurvatureFactor - RegularizationFactor). Only own entries are non-zero, so only interface cells
// (more than one own entry) hold such pairs. The list of up to K merged entries is gathered as the
// step's deferred path does and kept in registers (loops fully unrolled over K, guarded by the
// counts); energies per pair from the phase-pair table (GRAD false) or the orientation. Returns
// false when the cell has more than K merged entries.
template <int K, bool GRAD>
__device__ __forceinline__ bool raw_cell(const CurvArgs& P, int64_t c)
{
    const StepArgs& A = P.s;
    const int64_t N = A.ncells;
    CellList<K, double, GRAD> L;
#pragma unroll
    for (int t = 0; t < K; ++t) L.eid[t] = 0u;
    const int r = gather_cell<K, double, double, GRAD>(A, c, L);
    if (r == kCellDeferred) return false;
    if (r != kCellWide) return true;
    const int ne = L.ne, n0 = L.n0, NP = A.nphases;
    int ph[K];
    double vr[K];
    double localn = 0.0;
#pragma unroll
    for (int t = 0; t < K; ++t) {
        ph[t] = 0; vr[t] = 0.0;
        if (t < ne) { ph[t] = A.gphase[L.eid[t]]; vr[t] = A.gvr[L.eid[t]]; }
    }
    if (A.consider_nucleus) {
#pragma unroll
        for (int t = 0; t < K; ++t) if (t < ne) localn += vr[t];
    } else {
        localn = static_cast<double>(ne);
    }
    const double norm_1 = 1.0 / localn;
    const double Pf = A.prefactor;
    // InterfaceProperties energy of a pair (the lower list position first)
    auto sigma_of = [&](int x, int y) -> double {
        const int lo = x < y ? x : y, hi = x < y ? y : x;
        const int k = ph[lo] * NP + ph[hi];
        if (!GRAD) return A.sigma[k];
        return ip_pair_energy<double>(A, k, L.eid[lo], L.eid[hi], L.phi[lo], L.phi[hi], L.gr[GRAD ? lo : 0],
                                      L.gr[GRAD ? hi : 0]);
    };
#pragma unroll
    for (int b = 1; b < K; ++b)
#pragma unroll
        for (int a = 0; a < b; ++a) {
            if (b >= n0) continue;
            const double scale = sqrt(vr[a] * vr[b]);
            const double cf = 0.5 * (1.0 + scale) * Pf;
            double d = sigma_of(a, b) * ((L.lap[a] + cf * L.phi[a]) - (L.lap[b] + cf * L.phi[b]));
            if (ne > 2) {
#pragma unroll
                for (int g = 0; g < K; ++g) {
                    if (g == a || g == b || g >= ne) continue;
                    d += (sigma_of(b, g) - sigma_of(a, g)) * (L.lap[g] + cf * L.phi[g]);
                    if (A.tj_model == 0) {
                        if (A.tj_factor != 0.0 || scale < 1.0) {
                            const double e3 = A.sigma[ph[a] * NP + ph[b]] + A.sigma[ph[b] * NP + ph[g]] + A.sigma[ph[a] * NP + ph[g]];
                            d += smax(A.tj_factor, (1.0 - scale)) * Pf * e3 * (L.phi[a] * L.phi[g] - L.phi[b] * L.phi[g]);
                        }
                    } else if (A.tj_model == 1) {
                        d += A.tj_energy * Pf * (L.phi[a] * L.phi[g] - L.phi[b] * L.phi[g]);
                    }
                }
            }
            double kappa = d * P.prefactor2 * scale * norm_1 / sqrt(L.phi[a] * L.phi[b]);
            kappa *= P.factor;
            P.raw[static_cast<int64_t>(pair_plane(a, b)) * N + c] = kappa;
        }
    return true;
}

constexpr int kRawCells = 4 * kCurvBlock;  // cells per block of the raw pass

// A block takes kRawCells consecutive cells, collects its interface cells in LDS (in any order:
// each cell's result is its own) and gives one to each thread in turn, so no lane idles on bulk
// cells. Lists of up to 8 merged entries in registers, longer ones (more than 16: the step
// itself reports the overflow) with 16.
template <bool GRAD>
__global__ __launch_bounds__(kCurvBlock) void curvature_raw_kernel(const CurvArgs P)
{
    __shared__ uint16_t s_list[kRawCells];
    __shared__ int s_n;
    const int tid = static_cast<int>(threadIdx.x);
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kRawCells;
    if (tid == 0) s_n = 0;
    __syncthreads();
    for (int k = tid; k < kRawCells; k += kCurvBlock) {
        const int64_t c = base + k;
        if (c < P.s.ncells && ListCode<double>::count(P.s.cnt_in[c]) >= 2) s_list[atomicAdd(&s_n, 1)] = static_cast<uint16_t>(k);
    }
    __syncthreads();
    const int n = s_n;
    for (int i = tid; i < n; i += kCurvBlock) {
        const int64_t c = base + s_list[i];
        if (!raw_cell<8, GRAD>(P, c)) (void)raw_cell<kWideSlots, GRAD>(P, c);
    }
}

// SECOND = false: InterfaceRegularization::CollectAverageSR (src/InterfaceRegularization.cpp:196-223)
// into tmp; true: DistributeAverageSR (:225-274) into avg. The pair's weight is sqrt(phi_a phi_b)
// (SetWeightsSR, :178-194) on interface cells. Collect: for a weight above the threshold, the mean
// of the neighbours' raw weighted by weight * (Range - |offset|), over terms above DBL_EPSILON, if
// their sum exceeds DBL_EPSILON (else the fresh entry's 0). Distribute: for a weight above
// DBL_EPSILON, the plain mean of the tmp of the neighbours whose weight exceeds the threshold, or
// the own tmp. No grain is a Seed here (grains enter as Stable; only nucleation makes seeds), so
// the entries below those bounds keep 0. Lane-parallel over the sphere, summed in offset order as
// dg_average_kernel does.
template <bool SECOND>
__global__ __launch_bounds__(kAvgBlock) void curvature_average_kernel(const CurvArgs P)
{
    __shared__ uint16_t s_list[kAvgCells];
    __shared__ int s_count;
    const StepArgs& A = P.s;
    const int64_t N = A.ncells;
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kAvgCells;
    if (wave == 0) {
        const int64_t c0 = base + lane;
        const bool inter = lane < kAvgCells && c0 < N && ListCode<double>::count(A.cnt_in[c0]) >= 2;
        const unsigned long long m = __ballot(inter);
        if (inter) s_list[__popcll(m & ((1ull << lane) - 1ull))] = static_cast<uint16_t>(lane);
        if (lane == 0) s_count = __popcll(m);
    }
    __syncthreads();
    const int ninter = s_count;
    const int64_t nyz = static_cast<int64_t>(A.ny) * A.nz;
    const double* src = SECOND ? P.tmp : P.raw;
    for (int i = wave; i < ninter; i += kAvgWaves) {
        const int64_t c = base + s_list[i];
        const uint8_t code = A.cnt_in[c];
        const int n = ListCode<double>::count(code);
        const int x = static_cast<int>(c / nyz), y = static_cast<int>((c / A.nz) % A.ny), z = static_cast<int>(c % A.nz);
        for (int b = 1; b < n; ++b)
            for (int a = 0; a < b; ++a) {
                const uint32_t ga = A.id_in[static_cast<int64_t>(a) * N + c], gb = A.id_in[static_cast<int64_t>(b) * N + c];
                const int plane = pair_plane(a, b);
                const double w = sqrt(entry_value<double>(A, c, code, a) * entry_value<double>(A, c, code, b));
                double result = 0.0;
                if (SECOND ? (w > DBL_EPSILON) : (w > P.threshold)) {
                    double value = 0.0, sum = 0.0;
                    int counter = 0;
                    for (int t0 = 0; t0 < P.noff; t0 += 64) {
                        const int t = t0 + lane;
                        bool ok = false;
                        double wt = 0.0, term = 0.0;
                        if (t < P.noff) {
                            const int ii = P.off[3 * t], jj = P.off[3 * t + 1], kk = P.off[3 * t + 2];
                            const int64_t nb = (static_cast<int64_t>(wrap1(x + ii, A.nx)) * A.ny + wrap1(y + jj, A.ny)) * A.nz +
                                               wrap1(z + kk, A.nz);
                            double wphi = 0.0;
                            int np = 0;
                            bool fwd = true;
                            if (find_pair<double>(A, nb, ga, gb, wphi, np, fwd)) {
                                const double v = src[static_cast<int64_t>(np) * N + nb];
                                if (SECOND) {
                                    ok = wphi > P.threshold;
                                    term = fwd ? v : -v;
                                } else {
                                    wt = wphi * P.off_w[t];
                                    ok = wt > DBL_EPSILON;
                                    term = wt * (fwd ? v : -v);
                                }
                            }
                        }
                        unsigned long long okm = __ballot(ok);
                        while (okm) {
                            const int l = __ffsll(static_cast<long long>(okm)) - 1;
                            okm &= okm - 1ull;
                            if (SECOND) counter++;
                            else sum += lane_value(wt, l);
                            value += lane_value(term, l);
                        }
                    }
                    if (SECOND) result = counter != 0 ? value / static_cast<double>(counter) : P.tmp[static_cast<int64_t>(plane) * N + c];
                    else result = sum > DBL_EPSILON ? value / sum : 0.0;
                }
                if (lane == 0) (SECOND ? P.avg : P.tmp)[static_cast<int64_t>(plane) * N + c] = result;
            }
    }
}

}  // namespace

int launch_curvature(const CurvArgs& a, bool average, void* stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 raw_grid(static_cast<unsigned>((a.s.ncells + kRawCells - 1) / kRawCells));
    if (a.grad) hipLaunchKernelGGL((curvature_raw_kernel<true>), raw_grid, dim3(kCurvBlock), 0, s, a);
    else hipLaunchKernelGGL((curvature_raw_kernel<false>), raw_grid, dim3(kCurvBlock), 0, s, a);
    if (average) {
        const dim3 grid(static_cast<unsigned>((a.s.ncells + kAvgCells - 1) / kAvgCells));
        hipLaunchKernelGGL((curvature_average_kernel<false>), grid, dim3(kAvgBlock), 0, s, a);
        hipLaunchKernelGGL((curvature_average_kernel<true>), grid, dim3(kAvgBlock), 0, s, a);
    }
    return static_cast<int>(hipGetLastError());
}

int launch_dg_average(const AvgArgs& a, bool fp32, void* stream)
{
    const dim3 grid(static_cast<unsigned>((a.s.ncells + kAvgCells - 1) / kAvgCells));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (fp32) {
        hipLaunchKernelGGL((dg_average_kernel<float, false>), grid, dim3(kAvgBlock), 0, s, a);
        hipLaunchKernelGGL((dg_average_kernel<float, true>), grid, dim3(kAvgBlock), 0, s, a);
    } else {
        hipLaunchKernelGGL((dg_average_kernel<double, false>), grid, dim3(kAvgBlock), 0, s, a);
        hipLaunchKernelGGL((dg_average_kernel<double, true>), grid, dim3(kAvgBlock), 0, s, a);
    }
    return static_cast<int>(hipGetLastError());
}

}  // namespace mpf
