// Device-side list machinery shared by the step kernels and the adjoint kernels: slot-plane
// buffer loads, count codes, the per-cell register list, the 3^d box and the neighbour merge
// (PhaseField::CalculateDerivativesSR, src/PhaseField.cpp:566-615; flags of SetFlagsSR :953-969).
// Included by one .hip file each; everything sits in an anonymous namespace.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "mpf_device.h"

namespace mpf {

namespace {


constexpr int kMainSlots = 4;    // merged-list capacity of the main kernel
constexpr int kFastSlots = 12;   // largest capacity compiled with register-resident lists
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

// Raw buffer loads over one slot plane (byte size < 2^31, checked by the host): an offset at or
// beyond the plane size returns 0 without a memory access, so predicated neighbour loads need no
// branch (CDNA buffer descriptor, word 3 = 0x00020000).
constexpr uint32_t kNoLoad = 0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}

template <typename T> __device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ uint8_t buffer_load<uint8_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}
template <> __device__ __forceinline__ uint16_t buffer_load<uint16_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}
template <> __device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// std::min / std::max semantics (first argument wins ties and unordered compares).
__host__ __device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// Fixed-point image V = v * 2^116 of a non-negative v < 2^12 as four 32-bit limbs (exact for
// v >= 2^-64; smaller parts are truncated).
__device__ __forceinline__ void value_limbs(double v, unsigned long long l[kLimbs], bool& bad)
{
    unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
    int e = static_cast<int>((bits >> 52) & 0x7ff);
    l[0] = l[1] = l[2] = l[3] = 0ull;
    if ((bits >> 63) || e == 0x7ff) { bad = true; return; }
    if (e == 0) return;  // zero or subnormal
    unsigned long long mant = (bits & 0xFFFFFFFFFFFFFull) | (1ull << 52);
    int shift = e - 959;  // (e - 1023) - 52 + 116
    if (shift > 75) { bad = true; return; }  // v >= 2^12
    unsigned long long lo, hi;
    if (shift < 0) {
        lo = shift > -64 ? (mant >> (-shift)) : 0ull;
        hi = 0ull;
    } else if (shift == 0) {
        lo = mant; hi = 0ull;
    } else if (shift < 64) {
        lo = mant << shift; hi = mant >> (64 - shift);
    } else {
        lo = 0ull; hi = mant << (shift - 64);
    }
    l[0] = lo & 0xFFFFFFFFull;
    l[1] = lo >> 32;
    l[2] = hi & 0xFFFFFFFFull;
    l[3] = hi >> 32;
}

// Adds the (grain, value) entries of all lanes of the wavefront to the grain accumulators with one
// integer atomic per distinct grain and limb. Must be reached by whole wavefronts.
// LANE_EXACT = false: the lanes' values are summed in FP64 by a butterfly (fixed lane order);
// deterministic when the cell-to-lane mapping is (the main kernel).
// LANE_EXACT = true: every value is converted to fixed point first and the limbs are summed as
// integers; deterministic for any mapping (the work-list kernels, whose order varies).
template <bool LANE_EXACT>
__device__ __forceinline__ void wave_volume_add(bool act, uint32_t g, double v,
                                                unsigned long long* acc, bool& bad)
{
    const int lane = static_cast<int>(threadIdx.x) & (warpSize - 1);
    unsigned long long lv[kLimbs] = {0ull, 0ull, 0ull, 0ull};
    if (LANE_EXACT && act) value_limbs(v, lv, bad);
    while (true) {
        unsigned long long m = __ballot(act);
        if (m == 0ull) break;
        int leader = __ffsll(static_cast<long long>(m)) - 1;
        uint32_t g0 = static_cast<uint32_t>(__shfl(static_cast<int>(g), leader));
        bool sel = act && (g == g0);
        unsigned long long l[kLimbs];
        if (LANE_EXACT) {
#pragma unroll
            for (int i = 0; i < kLimbs; ++i) {
                unsigned long long s = sel ? lv[i] : 0ull;
#pragma unroll
                for (int k = 32; k >= 1; k >>= 1) s += __shfl_xor(s, k);
                l[i] = s;
            }
        } else {
            double s = sel ? v : 0.0;
#pragma unroll
            for (int k = 32; k >= 1; k >>= 1) s += __shfl_xor(s, k);
            if (lane == leader) value_limbs(s, l, bad);
        }
        if (lane == leader) {
#pragma unroll
            for (int i = 0; i < kLimbs; ++i)
                if (l[i]) atomicAdd(&acc[static_cast<int64_t>(g0) * kLimbs + i], l[i]);
        }
        act = act && !sel;
    }
}

// InterfaceProperties::SetSR's maxima (src/InterfaceProperties.cpp:748-756): the largest energy and
// mobility over the wavefront's cells, merged into the step's words with atomicMax on the bit
// patterns (non-negative doubles order like their bits). A word already at least as large (a
// possibly stale read) needs no atomic. Must be reached by whole wavefronts.
__device__ __forceinline__ void note_maxima(unsigned long long* obs, double emax, double mmax)
{
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) {
        emax = fmax(emax, __shfl_xor(emax, k));
        mmax = fmax(mmax, __shfl_xor(mmax, k));
    }
    if ((threadIdx.x & (warpSize - 1)) == 0) {
        const unsigned long long e = static_cast<unsigned long long>(__double_as_longlong(emax));
        const unsigned long long m = static_cast<unsigned long long>(__double_as_longlong(mmax));
        if (e > __atomic_load_n(&obs[0], __ATOMIC_RELAXED)) atomicMax(&obs[0], e);
        if (m > __atomic_load_n(&obs[1], __ATOMIC_RELAXED)) atomicMax(&obs[1], m);
    }
}

// InterfaceOrientation (src/InterfaceProperties.cpp:39-45) and the cubic energy of
// InterfaceEnergyModel::Calculate (include/InterfaceProperties/InterfaceEnergyModel.h:325-345).
template <typename TC>
__host__ __device__ __forceinline__ TC cubic_energy(const double* R, TC gx, TC gy, TC gz, TC energy, TC eps1)
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

// Interface energy InterfaceProperties::SetSR gives the pair (alpha, beta) of a wide-interface
// cell (src/InterfaceProperties.cpp:620-661), k = phase pair: the phase-pair energy (ISO), or the
// cubic energy of each solid grain's interface normal weighted by phi (1 - phi), over the sum of
// the weights, MaxEnergy when neither grain weighs (CUBIC). ga/gb: grains, va/vb: values,
// da/db: gradients of the two entries.
template <typename TC>
__host__ __device__ __forceinline__ TC ip_pair_energy(const StepArgs& A, int k, uint32_t ga, uint32_t gb, TC va, TC vb,
                                             const TC (&da)[3], const TC (&db)[3])
{
    if (A.emodel[k] == 0) return static_cast<TC>(A.sigma[k]);
    const TC E = static_cast<TC>(A.sigma[k]), e1 = static_cast<TC>(A.eps[k]);
    TC en = TC(0), wa = TC(0), wb = TC(0);
    if (A.gsolid[ga]) {
        wa = fabs(va * (TC(1.0) - va));
        if (wa > TC(DBL_EPSILON))
            en += cubic_energy<TC>(A.grot + 9 * static_cast<int64_t>(ga), da[0], da[1], da[2], E, e1) * wa;
        else
            wa = TC(0);
    }
    if (A.gsolid[gb]) {
        wb = fabs(vb * (TC(1.0) - vb));
        if (wb > TC(DBL_EPSILON))
            en += cubic_energy<TC>(A.grot + 9 * static_cast<int64_t>(gb), db[0], db[1], db[2], E, e1) * wb;
        else
            wb = TC(0);
    }
    if (wa + wb > TC(DBL_EPSILON)) en /= wa + wb;
    else en = static_cast<TC>(A.max_energy[k]);
    return en;
}

enum CellResult { kCellBulk = 0, kCellWide = 1, kCellDeferred = 2, kCellOverflow = 3 };

// Count byte of a stored list. FP64 states store the entry count. FP32 states store
// count | (implicit slot << 4): the cell's own update recovers the largest entry as
// 1 - sum(other entries) in FP64, so values near 1 keep their FP64 resolution; its FP32 value is
// stored as well for the neighbours' stencils, except in single-entry lists (value 1, nothing
// stored). Implicit slot 15 means "none" (every value as given): the form mpf_upload writes.
template <typename TV> struct ListCode {
    static constexpr bool kImplicit = sizeof(TV) == 4;
    __device__ static int count(uint8_t b) { return kImplicit ? (b & 15) : b; }
    __device__ static int implicit(uint8_t b) { return kImplicit && (b >> 4) != 15 ? (b >> 4) : -1; }
};

// Per-cell list held in registers (K entries): phase-field index, value (FP64 and in the stencil
// arithmetic), Laplacian and gradient of the neighbour-merged list.
template <int K, typename TC, bool GRAD> struct CellList {
    uint32_t eid[K];
    double phi[K];
    // value in the stencil arithmetic: a separate copy only when that is not FP64
    static constexpr bool kCopy = sizeof(TC) != sizeof(double);
    TC ev_[kCopy ? K : 1];
    __device__ TC ev(int t) const
    {
        if constexpr (kCopy) return ev_[t];
        else return phi[t];
    }
    TC lap[K];
    TC gr[GRAD ? K : 1][3];
    bool keep[K];
    int n0;  // own entries (flag 2 when > 1)
    int ne;  // merged entries
    double emax, mmax;  // largest interface energy / mobility of the cell's pairs (IP.Set maxima)
};

// Neighbour cells of one cell as 32-bit offsets: nine row starts (ii, jj) plus the z coordinate.
// An offset along an inactive axis aliases the centre and is masked out. Box position
// k = 9 (ii+1) + 3 (jj+1) + (kk+1) is stencil order (x outer, z inner).
struct Box {
    uint32_t rows[9];
    uint32_t zs[3];
    __device__ uint32_t cell(int k) const { return rows[k / 3] + zs[k % 3]; }
};

__device__ __forceinline__ Box make_box(const StepArgs& A, int64_t cc)
{
    const int NY = A.ny, NZ = A.nz;
    const int x = static_cast<int>(cc / (static_cast<int64_t>(NY) * NZ));
    const int y = static_cast<int>((cc / NZ) % NY);
    const int z = static_cast<int>(cc % NZ);
    Box B;
    const int xv[3] = {A.dnx ? wrap1(x - 1, A.nx) : x, x, A.dnx ? wrap1(x + 1, A.nx) : x};
    const int yv[3] = {A.dny ? wrap1(y - 1, NY) : y, y, A.dny ? wrap1(y + 1, NY) : y};
    B.zs[0] = static_cast<uint32_t>(A.dnz ? wrap1(z - 1, NZ) : z);
    B.zs[1] = static_cast<uint32_t>(z);
    B.zs[2] = static_cast<uint32_t>(A.dnz ? wrap1(z + 1, NZ) : z);
#pragma unroll
    for (int r = 0; r < 9; ++r) B.rows[r] = static_cast<uint32_t>((xv[r / 3] * NY + yv[r % 3]) * NZ);
    return B;
}

__device__ __forceinline__ bool box_active(const StepArgs& A, int k)
{
    const int ii = k / 9 - 1, jj = (k / 3) % 3 - 1, kk = k % 3 - 1;
    return (ii == 0 || A.dnx) && (jj == 0 || A.dny) && (kk == 0 || A.dnz);
}

__device__ __forceinline__ uint8_t box_code(const uint32_t codes[7], int k)
{
    return static_cast<uint8_t>(codes[k >> 2] >> (8 * (k & 3)));
}

// box_code / Box::rows for an index known only at run time: selects over the registers
__device__ __forceinline__ uint8_t box_code_dyn(const uint32_t codes[7], int k)
{
    const int w = k >> 2;
    uint32_t v = codes[0];
#pragma unroll
    for (int t = 1; t < 7; ++t) v = w == t ? codes[t] : v;
    return static_cast<uint8_t>(v >> (8 * (k & 3)));
}

__device__ __forceinline__ uint32_t box_row_dyn(const Box& B, int r)
{
    uint32_t v = B.rows[0];
#pragma unroll
    for (int t = 1; t < 9; ++t) v = r == t ? B.rows[t] : v;
    return v;
}

// Count bytes of the whole box (27 independent loads, packed four to a word; inactive positions
// read as 0). Returns wide_interface(): flag 2 (more than one entry) or a flag-2 cell in the box.
template <typename TV>
__device__ __forceinline__ bool load_codes(const StepArgs& A, const Box& B, uint8_t code, uint32_t codes[7])
{
    const __amdgpu_buffer_rsrc_t rc = plane_rsrc(A.cnt_in, static_cast<uint32_t>(A.ncells));
#pragma unroll
    for (int w = 0; w < 7; ++w) codes[w] = 0u;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const uint8_t b = k == 13 ? code : buffer_load<uint8_t>(rc, box_active(A, k) ? B.cell(k) : kNoLoad);
        codes[k >> 2] |= static_cast<uint32_t>(b) << (8 * (k & 3));
    }
    bool wide = false;
#pragma unroll
    for (int k = 0; k < 27; ++k) wide = wide || ListCode<TV>::count(box_code(codes, k)) > 1;
    return wide;
}

// Loads the own list of cell cc into L (values in FP64; the implicit FP32 entry is 1 - sum of the
// others in slot order). Returns kCellDeferred when it has more than K entries.
template <int K, typename TV, typename TC, bool GRAD>
__device__ __forceinline__ int load_own(const StepArgs& A, int64_t cc, uint8_t code, CellList<K, TC, GRAD>& L)
{
    constexpr int U = K <= kFastSlots ? K : 1;  // full unroll only in registers
    const int64_t N = A.ncells;
    const int n0 = ListCode<TV>::count(code);
    const int imp0 = ListCode<TV>::implicit(code);
    L.n0 = n0;
    L.ne = n0;
    L.emax = 0.0;
    L.mmax = 0.0;
#pragma unroll U
    for (int t = 0; t < K; ++t) {
        L.eid[t] = 0u; L.phi[t] = 0.0; L.lap[t] = TC(0); L.keep[t] = false;
        if constexpr (CellList<K, TC, GRAD>::kCopy) L.ev_[t] = TC(0);
        if (GRAD) { L.gr[GRAD ? t : 0][0] = L.gr[GRAD ? t : 0][1] = L.gr[GRAD ? t : 0][2] = TC(0); }
    }
    if (n0 > K) return kCellDeferred;
#pragma unroll U
    for (int t = 0; t < K; ++t)
        if (t < n0) {
            L.eid[t] = A.id_in[static_cast<int64_t>(t) * N + cc];
            L.keep[t] = true;
            if (!(ListCode<TV>::kImplicit && t == imp0))
                L.phi[t] = static_cast<double>(load_val<TV>(A.val_in, static_cast<int64_t>(t) * N + cc));
        }
    if (ListCode<TV>::kImplicit) {
        double others = 0.0;
#pragma unroll U
        for (int t = 0; t < K; ++t)
            if (t < n0 && t != imp0) others += L.phi[t];
#pragma unroll U
        for (int t = 0; t < K; ++t)
            if (t == imp0) L.phi[t] = 1.0 - others;
    }
#pragma unroll U
    for (int t = 0; t < K; ++t)
        if constexpr (CellList<K, TC, GRAD>::kCopy) L.ev_[t] = static_cast<TC>(L.phi[t]);
    return kCellWide;
}

// Adds one neighbour entry (grain gid, weighted value l and gradient parts) to the merged list:
// accumulate into the matching entry or append it (CalculateDerivativesSR appends neighbour
// indices at value 0). Straight-line selects behind two tests over the active lanes of the
// wavefront (an entry no lane has, an append no lane needs); `over` records an append beyond K.
template <int K, typename TC, bool GRAD>
__device__ __forceinline__ void merge_entry(CellList<K, TC, GRAD>& L, int& ne, bool& over, bool valid,
                                            uint32_t gid, TC l, TC lx, TC ly, TC lz, bool side)
{
    constexpr int U = K <= kFastSlots ? K : 1;
    if (!__any(valid)) return;
    bool found = false;
    // adding +0 leaves a sum unchanged (the sums are never -0)
#pragma unroll U
    for (int t = 0; t < K; ++t) {
        const bool mt = valid && t < ne && L.eid[t] == gid;
        L.lap[t] += mt ? l : TC(0);
        if (GRAD && side) {
            L.gr[GRAD ? t : 0][0] = L.gr[GRAD ? t : 0][0] + (mt ? lx : TC(0));
            L.gr[GRAD ? t : 0][1] = L.gr[GRAD ? t : 0][1] + (mt ? ly : TC(0));
            L.gr[GRAD ? t : 0][2] = L.gr[GRAD ? t : 0][2] + (mt ? lz : TC(0));
        }
        found = found || mt;
    }
    const bool app = valid && !found;
    if (!__any(app)) return;
    over = over || (app && ne >= K);
#pragma unroll U
    for (int t = 0; t < K; ++t) {
        const bool at = app && t == ne;
        L.eid[t] = at ? gid : L.eid[t];
        L.lap[t] = at ? l : L.lap[t];
        if (GRAD) {
            L.gr[GRAD ? t : 0][0] = at ? lx : L.gr[GRAD ? t : 0][0];
            L.gr[GRAD ? t : 0][1] = at ? ly : L.gr[GRAD ? t : 0][1];
            L.gr[GRAD ? t : 0][2] = at ? lz : L.gr[GRAD ? t : 0][2];
        }
    }
    ne += app ? 1 : 0;
}

// Merges the neighbour lists of a wide cell into L in stencil order, accumulating Laplacians (and
// gradients); appended entries start at value 0 (CalculateDerivativesSR). One plane of nine
// neighbours per loop trip (the loop is not unrolled, which keeps the kernel within the
// instruction cache): the first two entries of the nine are loaded together with branch-free
// buffer loads, a third entry (rare) on demand. A list growing beyond K returns kCellDeferred.
// Entries beyond a neighbour's third are read one by one (MANY) or defer the cell.
template <int K, typename TV, typename TC, bool GRAD, bool MANY>
__device__ __forceinline__ int merge_neighbours(const StepArgs& A, const Box& B, const uint32_t codes[7],
                                                CellList<K, TC, GRAD>& L)
{
    const int64_t N = A.ncells;
    const uint32_t N32 = static_cast<uint32_t>(N);
    constexpr uint32_t SV = static_cast<uint32_t>(sizeof(TV));
    const __amdgpu_buffer_rsrc_t ri0 = plane_rsrc(A.id_in, N32 * 2u);
    const __amdgpu_buffer_rsrc_t ri1 = plane_rsrc(A.id_in + N, N32 * 2u);
    const __amdgpu_buffer_rsrc_t rv0 = plane_rsrc(A.val_in, N32 * SV);
    const __amdgpu_buffer_rsrc_t rv1 = plane_rsrc(static_cast<const TV*>(A.val_in) + N, N32 * SV);
    int ne = L.n0;
    bool over = false;
    const int ip_end = A.dnx ? 3 : 2;
    // Neighbour values are read as stored: FP32 lists store every value, the implicit entry's as
    // well (only the cell's own update recovers it as 1 - sum of the others in FP64).
    if constexpr (GRAD) {
        // orientation-dependent variants: a plane of nine neighbours per batch of loads
#pragma unroll 1
        for (int ip = A.dnx ? 0 : 1; ip < ip_end; ++ip) {
            // the plane's count bytes and row starts (selects: ip is a loop variable)
            uint8_t pc[9];
#pragma unroll
            for (int j = 0; j < 9; ++j)
                pc[j] = ip == 0 ? box_code(codes, j) : (ip == 1 ? box_code(codes, 9 + j) : box_code(codes, 18 + j));
            uint32_t prow[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) prow[r] = ip == 0 ? B.rows[r] : (ip == 1 ? B.rows[3 + r] : B.rows[6 + r]);
            TV v0[9], v1[9];
            uint32_t g0[9], g1[9];
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int m = ListCode<TV>::count(pc[j]), imp = ListCode<TV>::implicit(pc[j]);
                const uint32_t nb = prow[j / 3] + B.zs[j % 3];
                // a lone implicit entry (FP32) has no stored value: it is 1
                const bool lone = ListCode<TV>::kImplicit && m == 1 && imp == 0;
                g0[j] = buffer_load<uint16_t>(ri0, m > 0 ? nb * 2u : kNoLoad);
                v0[j] = lone ? TV(1) : buffer_load<TV>(rv0, m > 0 ? nb * SV : kNoLoad);
                g1[j] = buffer_load<uint16_t>(ri1, m > 1 ? nb * 2u : kNoLoad);
                v1[j] = buffer_load<TV>(rv1, m > 1 ? nb * SV : kNoLoad);
            }
            const int dx = ip - 1;
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int jj = j / 3 - 1, kk = j % 3 - 1;
                const int m = ListCode<TV>::count(pc[j]);
                const uint32_t nb = prow[j / 3] + B.zs[j % 3];
                const int nj = (jj != 0) + (kk != 0);
                const bool side = dx != 0 || nj != 0;
                const TC w = static_cast<TC>(dx != 0 ? A.wl[nj + 1] : A.wl[nj]);
                const TC gw = static_cast<TC>(dx != 0 ? A.wg[nj + 1] : A.wg[nj]);
                const TC gwx = dx == 0 ? TC(0) : (dx < 0 ? -gw : gw);
                const TC gwy = jj == 0 ? TC(0) : (jj < 0 ? -gw : gw);
                const TC gwz = kk == 0 ? TC(0) : (kk < 0 ? -gw : gw);
                // entries beyond the second (rare), loaded on demand
                TV v2 = TV(0);
                uint32_t g2 = 0u;
                if (m > 2) {
                    g2 = A.id_in[2 * N + nb];
                    v2 = load_val<TV>(A.val_in, 2 * N + nb);
                }
                const TC a0 = static_cast<TC>(v0[j]), a1 = static_cast<TC>(v1[j]);
                merge_entry<K, TC, GRAD>(L, ne, over, m > 0 && a0 != TC(0), g0[j], w * a0,
                                         gwx * a0, gwy * a0, gwz * a0, side);
                merge_entry<K, TC, GRAD>(L, ne, over, m > 1 && a1 != TC(0), g1[j], w * a1,
                                         gwx * a1, gwy * a1, gwz * a1, side);
                if (m > 2) {
                    const TC a2 = static_cast<TC>(v2);
                    merge_entry<K, TC, GRAD>(L, ne, over, a2 != TC(0), g2, w * a2, gwx * a2, gwy * a2, gwz * a2, side);
                    if (!MANY) {
                        over = over || m > 3;
                    } else {
                        for (int q = 3; q < m; ++q) {
                            const TC aq = static_cast<TC>(load_val<TV>(A.val_in, static_cast<int64_t>(q) * N + nb));
                            const uint32_t gq = A.id_in[static_cast<int64_t>(q) * N + nb];
                            merge_entry<K, TC, GRAD>(L, ne, over, aq != TC(0), gq, w * aq, gwx * aq, gwy * aq, gwz * aq, side);
                        }
                    }
                }
            }
        }
    } else {
        // plain variants: a row of three neighbours per batch of loads, which keeps the step kernel
        // within 128 VGPRs without spilling (256^3 FP64 step kernel 3.26 ms vs 3.66 ms with plane
        // batches; the orientation variants lose about 6% this way)
#pragma unroll 1
        for (int ir = (A.dnx ? 0 : 3); ir < ip_end * 3; ++ir) {
            // one row of three neighbours (plane ip, row r) per batch of loads
            const int ip = ir / 3, r = ir - 3 * (ir / 3);
            const int b = ir * 3;  // box index of the row's first neighbour
            uint8_t pc[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) pc[j] = box_code_dyn(codes, b + j);
            const uint32_t prow = box_row_dyn(B, ir);
            TV v0[3], v1[3], v2p[3], v3p[3];
            uint32_t g0[3], g1[3], g2p[3], g3p[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int m = ListCode<TV>::count(pc[j]), imp = ListCode<TV>::implicit(pc[j]);
                const uint32_t nb = prow + B.zs[j];
                // a lone implicit entry (FP32) has no stored value: it is 1
                const bool lone = ListCode<TV>::kImplicit && m == 1 && imp == 0;
                g0[j] = buffer_load<uint16_t>(ri0, m > 0 ? nb * 2u : kNoLoad);
                v0[j] = lone ? TV(1) : buffer_load<TV>(rv0, m > 0 ? nb * SV : kNoLoad);
                g1[j] = buffer_load<uint16_t>(ri1, m > 1 ? nb * 2u : kNoLoad);
                v1[j] = buffer_load<TV>(rv1, m > 1 ? nb * SV : kNoLoad);
                // the work-list kernels (MANY) load a third and fourth entry with the first two,
                // so a neighbour of four entries costs no extra round trip (256^3 deferred-cell
                // kernels 0.80 -> 0.77 ms)
                g2p[j] = 0u; v2p[j] = TV(0); g3p[j] = 0u; v3p[j] = TV(0);
                if (MANY) {
                    if (m > 2) { g2p[j] = A.id_in[2 * N + nb]; v2p[j] = load_val<TV>(A.val_in, 2 * N + nb); }
                    if (m > 3) { g3p[j] = A.id_in[3 * N + nb]; v3p[j] = load_val<TV>(A.val_in, 3 * N + nb); }
                }
            }
            const int dx = ip - 1, jj = r - 1;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int kk = j - 1;
                const int m = ListCode<TV>::count(pc[j]);
                const uint32_t nb = prow + B.zs[j];
                const int nj = (jj != 0) + (kk != 0);
                const bool side = dx != 0 || nj != 0;
                const TC w = static_cast<TC>(dx != 0 ? A.wl[nj + 1] : A.wl[nj]);
                const TC gw = static_cast<TC>(dx != 0 ? A.wg[nj + 1] : A.wg[nj]);
                const TC gwx = dx == 0 ? TC(0) : (dx < 0 ? -gw : gw);
                const TC gwy = jj == 0 ? TC(0) : (jj < 0 ? -gw : gw);
                const TC gwz = kk == 0 ? TC(0) : (kk < 0 ? -gw : gw);
                // entries beyond the second (rare), loaded on demand
                TV v2 = TV(0);
                uint32_t g2 = 0u;
                if (m > 2) {
                    g2 = MANY ? g2p[j] : A.id_in[2 * N + nb];
                    v2 = MANY ? v2p[j] : load_val<TV>(A.val_in, 2 * N + nb);
                }
                const TC a0 = static_cast<TC>(v0[j]), a1 = static_cast<TC>(v1[j]);
                merge_entry<K, TC, GRAD>(L, ne, over, m > 0 && a0 != TC(0), g0[j], w * a0,
                                         gwx * a0, gwy * a0, gwz * a0, side);
                merge_entry<K, TC, GRAD>(L, ne, over, m > 1 && a1 != TC(0), g1[j], w * a1,
                                         gwx * a1, gwy * a1, gwz * a1, side);
                if (m > 2) {
                    const TC a2 = static_cast<TC>(v2);
                    merge_entry<K, TC, GRAD>(L, ne, over, a2 != TC(0), g2, w * a2, gwx * a2, gwy * a2, gwz * a2, side);
                    if (!MANY) {
                        over = over || m > 3;
                    } else {
                        for (int q = 3; q < m; ++q) {
                            const TC aq = q == 3 ? static_cast<TC>(v3p[j])
                                                 : static_cast<TC>(load_val<TV>(A.val_in, static_cast<int64_t>(q) * N + nb));
                            const uint32_t gq = q == 3 ? g3p[j] : A.id_in[static_cast<int64_t>(q) * N + nb];
                            merge_entry<K, TC, GRAD>(L, ne, over, aq != TC(0), gq, w * aq, gwx * aq, gwy * aq, gwz * aq, side);
                        }
                    }
                }
            }
        }
    }
    L.ne = ne;
    return over ? kCellDeferred : kCellWide;
}

// Own list plus neighbour merge of cell cc (the work-list kernels). Returns kCellBulk (nothing to
// do), kCellWide, or kCellDeferred when more than K entries are needed.
template <int K, typename TV, typename TC, bool GRAD>
__device__ __forceinline__ int gather_cell(const StepArgs& A, int64_t cc, CellList<K, TC, GRAD>& L)
{
    const uint8_t code = A.cnt_in[cc];
    if (load_own<K, TV, TC, GRAD>(A, cc, code, L) == kCellDeferred) return kCellDeferred;
    const Box B = make_box(A, cc);
    uint32_t codes[7];
    if (!load_codes<TV>(A, B, code, codes)) return kCellBulk;
    return merge_neighbours<K, TV, TC, GRAD, true>(A, B, codes, L);
}

}  // namespace

}  // namespace mpf
