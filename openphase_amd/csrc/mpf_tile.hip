// MI355X (gfx950) LDS-tiled step kernel of the multiphase-field time step.
//
// Same step as step_kernel (mpf_kernels.hip): IP.Set, double-obstacle increments, driving force,
// NormalizeIncrements, MergeIncrements and finalize of every cell (src/InterfaceProperties.cpp:
// 603-762, src/DoubleObstacle.cpp:502-571, src/PhaseField.cpp:1268-1299, 1351-1567,
// include/Containers/NodePF.h:214-255) on the neighbour-merged lists of CalculateDerivativesSR
// (src/PhaseField.cpp:566-615), with a different way of reading the neighbourhood.
//
// A block owns a TY x TZ tile of the (y, z) plane and marches along x through CH planes. Each
// plane of the tile plus a one-cell halo is staged in LDS once (a ring of three planes: x-1, x,
// x+1), so every state entry is read from memory about (CH + 2) / CH x (halo area) times instead of
// 27 times. Grains are renamed to local ids 0..31 per block (an LDS hash table filled as planes are
// staged); a staged cell is a presence mask over local ids, its list order (up to three local ids)
// and its values ordered by local id. A cell's merged list is then the OR of 27 masks, walked in
// stencil order only where a new grain appears (to keep the reference's insertion order), and the
// Laplacian (and gradient) of merged entry t is a sum over the 27 neighbours of one LDS value each,
// at rank popcount(mask & below(t)) or a zero when the bit is clear: no list searches.
//
// Cells the tile cannot hold - a neighbourhood with a list of more than three entries or a grain
// that did not fit the 32 local ids, more than four merged entries, fewer than two - go to the
// deferred-cell work list of the existing kernels, which handle any list.
//
// Grain volumes (PhaseField::CalculateGrainsVolume, src/PhaseField.cpp:812-951) accumulate per
// local id in LDS as exact 128-bit fixed-point integers (entries of exactly 1 as counts) and go to
// the global accumulators once per block: integer sums, so the result depends neither on the
// schedule nor on the tile shape.

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "mpf_device.h"
#include "mpf_lists.h"
#include "mpf_step_ops.h"

namespace mpf {

namespace {

constexpr int kRanks = 4;        // entries per staged cell (longer lists make their neighbourhood deferred)
constexpr int kLocal = 32;       // local grain ids per block
constexpr int kCopies = 4;       // copies of the volume accumulators (lanes spread over them)
constexpr int kTileWaves = kBlock / 64;
static_assert(kTileHalo <= 2 * kBlock, "a thread stages at most two cells of a plane");

// info word of a staged cell: bits 0-3 entry count, bit 4 unusable (more than kRanks entries or
// an unmapped grain), bits 5-8 the FP32 implicit slot (15: none), bits 9-28 the local ids of the
// first four entries in list order (5 bits each)
constexpr int kInfoBad = 4, kInfoImp = 5, kInfoOrd = 9;

// Local id of grain g in the block's table (open addressing over 32 slots; the slot is the id),
// inserting it when absent. -1 when the table is full.
__device__ __forceinline__ int local_id(int32_t* key, uint32_t g)
{
    const uint32_t h = (g * 0x9E3779B1u) >> 27;
#pragma unroll 1
    for (int i = 0; i < kLocal; ++i) {
        const int slot = static_cast<int>((h + static_cast<uint32_t>(i)) & (kLocal - 1));
        const int k = *static_cast<volatile int32_t*>(&key[slot]);
        if (k == static_cast<int>(g)) return slot;
        if (k == -1) {
            const int prev = atomicCAS(&key[slot], -1, static_cast<int>(g));
            if (prev == -1 || prev == static_cast<int>(g)) return slot;
        }
    }
    return -1;
}

// Exact LDS accumulation of one entry's value for local id `lid` (4 limbs of V = v * 2^116) into
// the lane's copy of the accumulators (lanes of one wavefront adding to one grain spread over
// kCopies addresses instead of serialising on one).
__device__ __forceinline__ void lds_volume_add(unsigned long long (*acc)[kCopies][kLimbs], int lid, double v, bool& bad)
{
    unsigned long long l[kLimbs];
    value_limbs(v, l, bad);
    const int copy = static_cast<int>(threadIdx.x) & (kCopies - 1);
#pragma unroll
    for (int i = 0; i < kLimbs; ++i)
        if (l[i]) atomicAdd(&acc[lid][copy][i], l[i]);
}

// Entries of exactly 1.0 (single-entry cells) as counts per local id: one LDS atomic per distinct
// id of the wavefront. Must be reached by whole wavefronts.
__device__ __forceinline__ void lds_unit_add(unsigned long long (*acc)[kCopies][kLimbs], bool one, int lid)
{
    const int lane = static_cast<int>(threadIdx.x) & 63;
    while (true) {
        const unsigned long long m = __ballot(one);
        if (m == 0ull) break;
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        const int l0 = __shfl(lid, leader);
        const bool sel = one && lid == l0;
        const unsigned long long cells = static_cast<unsigned long long>(__popcll(__ballot(sel)));
        if (lane == leader) atomicAdd(&acc[l0][leader & (kCopies - 1)][3], cells << 20);
        one = one && !sel;
    }
}

// Laplacians (and gradients) of the first T merged entries of a cell in stencil order (x outer, z
// inner; inactive axes skipped): entry t of neighbour k contributes w_k * value, the value staged
// at rank popcount(mask & lm[t]) when bit bm[t] is set and a zero otherwise (adding w * 0 leaves a
// sum unchanged: the sums are never -0). pl: ring slots of the planes x-1, x, x+1; wh: the cell's
// index in its halo plane.
template <int T, int K, typename TV, typename TC, bool GRAD>
__device__ __forceinline__ void tile_derivatives(const StepArgs& A, const uint32_t (*s_mask)[kTileHalo],
                                                 const TV (*s_val)[kRanks][kTileHalo], const TV* zero, const int pl[3],
                                                 int wh, int HZ, const uint32_t bm[K], const uint32_t lm[K],
                                                 CellList<K, TC, GRAD>& L)
{
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
        if (!((dx == 0 || A.dnx) && (dy == 0 || A.dny) && (dz == 0 || A.dnz))) continue;
        const int p = pl[dx + 1];
        const int hh = wh + dy * HZ + dz;
        const uint32_t m = s_mask[p][hh];
        const int nj = (dy != 0) + (dz != 0);
        const TC w = static_cast<TC>(dx != 0 ? A.wl[nj + 1] : A.wl[nj]);
        const TC gw = static_cast<TC>(dx != 0 ? A.wg[nj + 1] : A.wg[nj]);
        const TC gwx = dx == 0 ? TC(0) : (dx < 0 ? -gw : gw);
        const TC gwy = dy == 0 ? TC(0) : (dy < 0 ? -gw : gw);
        const TC gwz = dz == 0 ? TC(0) : (dz < 0 ? -gw : gw);
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const bool pres = (m & bm[t]) != 0u;
            const TV* src = pres ? &s_val[p][__popc(m & lm[t])][hh] : zero;
            const TC a = static_cast<TC>(*src);
            L.lap[t] += w * a;
            if (GRAD && k != 13) {
                L.gr[GRAD ? t : 0][0] = L.gr[GRAD ? t : 0][0] + gwx * a;
                L.gr[GRAD ? t : 0][1] = L.gr[GRAD ? t : 0][1] + gwy * a;
                L.gr[GRAD ? t : 0][2] = L.gr[GRAD ? t : 0][2] + gwz * a;
            }
        }
    }
}

// One block of the tiled step: tile `b` of the domain A (tiles numbered z fastest, then y, then
// the x-chunk).
template <typename TV, typename TC, bool GRAD, bool DG, bool UNI>
__device__ __forceinline__ void tile_block(const StepArgs& A, const unsigned b)
{
    __shared__ uint32_t s_info[3][kTileHalo];
    __shared__ uint32_t s_mask[3][kTileHalo];
    __shared__ TV s_val[3][kRanks][kTileHalo];
    __shared__ TV s_zero;
    __shared__ int32_t s_key[kLocal];
    __shared__ unsigned long long s_acc[kLocal][kCopies][kLimbs];
    __shared__ uint16_t s_list[kBlock];
    __shared__ uint32_t s_nwide[kTileWaves];
    // cells of a plane for the deferred-cell kernels, by plane parity: appended in LDS, moved to
    // the global work list by wavefront 0 after the plane (one global atomic per plane)
    // entries: tile-local cell index (y * TZ + z) | work list << 8
    __shared__ uint16_t s_defer[2][kBlock];
    __shared__ int s_ndefer[2];

    constexpr int K = kMainSlots;
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int64_t N = A.ncells;
    const int NY = A.ny, NZ = A.nz;
    const int TY = A.tile_y, TZ = A.tile_z;
    const int HZ = TZ + 2 * A.dnz, HP = (TY + 2 * A.dny) * HZ;
    const int nty = (NY + TY - 1) / TY, ntz = (NZ + TZ - 1) / TZ;
    const int tz = static_cast<int>(b % static_cast<unsigned>(ntz));
    const int ty = static_cast<int>((b / static_cast<unsigned>(ntz)) % static_cast<unsigned>(nty));
    const int xc = static_cast<int>(b / static_cast<unsigned>(ntz * nty));
    const int y0 = ty * TY, z0 = tz * TZ;
    const int x0 = xc * A.tile_x, x1 = min(x0 + A.tile_x, A.nx);

    if (tid < kLocal) s_key[tid] = -1;
    if (tid < 2) s_ndefer[tid] = 0;
    for (int i = tid; i < kLocal * kCopies * kLimbs; i += kBlock) (&s_acc[0][0][0])[i] = 0ull;
    if (tid == 0) s_zero = TV(0);
    __syncthreads();

    bool bad_volume = false;
    double emax = 0.0, mmax = 0.0;  // IP.Set maxima of this wavefront's cells over all planes

    // ---- staging of plane xg (wrapped) into ring slot `slot`. A thread stages the halo cells h
    // and h + 256: their count bytes first, then every stored entry of both with branch-free
    // buffer loads (an offset past the plane reads 0 without a memory access), so a plane costs
    // two memory round trips; then the grains are renamed and the LDS records written.
    const uint32_t N32 = static_cast<uint32_t>(N);
    constexpr uint32_t SV = static_cast<uint32_t>(sizeof(TV));
    const __amdgpu_buffer_rsrc_t rc = plane_rsrc(A.cnt_in, N32);
    __amdgpu_buffer_rsrc_t ri[kRanks], rv[kRanks];
#pragma unroll
    for (int q = 0; q < kRanks; ++q) {
        ri[q] = plane_rsrc(A.id_in + static_cast<int64_t>(q) * N, N32 * 2u);
        rv[q] = plane_rsrc(static_cast<const TV*>(A.val_in) + static_cast<int64_t>(q) * N, N32 * SV);
    }
    // the thread's two halo cells of every plane: offsets within a plane
    uint32_t yz[2];
    bool hv[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int h = tid + j * kBlock;
        hv[j] = h < HP;
        const int hy = h / HZ, hz = h - hy * HZ;
        int yg = y0 - A.dny + hy, zg = z0 - A.dnz + hz;
        yg = yg < 0 ? yg + NY : (yg >= NY ? yg - NY : yg);
        zg = zg < 0 ? zg + NZ : (zg >= NZ ? zg - NZ : zg);
        yz[j] = hv[j] ? static_cast<uint32_t>(yg * NZ + zg) : 0u;
    }
    const uint32_t plane_cells = static_cast<uint32_t>(NY) * static_cast<uint32_t>(NZ);
    // count bytes of plane xg (issued one plane ahead of their use, so they arrive while the
    // current plane is computed)
    auto fetch = [&](int xg, uint8_t (&code)[2]) {
#pragma unroll
        for (int j = 0; j < 2; ++j)
            code[j] = buffer_load<uint8_t>(rc, hv[j] ? static_cast<uint32_t>(xg) * plane_cells + yz[j] : kNoLoad);
    };
    auto stage = [&](int slot, int xg, const uint8_t (&code)[2]) {
        uint32_t cell[2];
        int hcell[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            hcell[j] = tid + j * kBlock;
            cell[j] = static_cast<uint32_t>(xg) * plane_cells + yz[j];
        }
        uint32_t gid[2][kRanks];
        TV v[2][kRanks];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = hcell[j] < HP ? ListCode<TV>::count(code[j]) : 0;
#pragma unroll
            for (int q = 0; q < kRanks; ++q) {
                gid[j][q] = buffer_load<uint16_t>(ri[q], q < n ? cell[j] * 2u : kNoLoad);
                v[j][q] = buffer_load<TV>(rv[q], q < n ? cell[j] * SV : kNoLoad);
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int h = hcell[j];
            if (h >= HP) continue;
            const int n = ListCode<TV>::count(code[j]), imp = ListCode<TV>::implicit(code[j]);
            bool bad = n > kRanks;
            uint32_t mask = 0u, ord = 0u;
            int lid[kRanks];
#pragma unroll
            for (int q = 0; q < kRanks; ++q) {
                lid[q] = 0;
                if (q < n) {
                    const int l = local_id(s_key, gid[j][q]);
                    // a lone implicit entry (FP32) has no stored value: it is 1
                    if (ListCode<TV>::kImplicit && n == 1 && imp == 0) v[j][q] = TV(1);
                    if (l < 0) bad = true;
                    lid[q] = l < 0 ? 0 : l;
                    ord |= static_cast<uint32_t>(lid[q]) << (kInfoOrd + 5 * q);
                    if (v[j][q] != TV(0)) mask |= 1u << lid[q];
                }
            }
            if (!bad) {
#pragma unroll
                for (int q = 0; q < kRanks; ++q)
                    if (q < n && v[j][q] != TV(0)) s_val[slot][__popc(mask & ((1u << lid[q]) - 1u))][h] = v[j][q];
            }
            s_mask[slot][h] = bad ? 0u : mask;
            s_info[slot][h] = static_cast<uint32_t>(n) | (bad ? 1u << kInfoBad : 0u) |
                              (static_cast<uint32_t>(imp & 15) << kInfoImp) | ord;
        }
    };

    auto ring = [&](int x) { return (x - x0 + 1) % 3; };
    auto wrapx = [&](int x) { return x < 0 ? x + A.nx : (x >= A.nx ? x - A.nx : x); };
    uint8_t codes[2];
    if (A.dnx) {
        fetch(wrapx(x0 - 1), codes);
        stage(ring(x0 - 1), wrapx(x0 - 1), codes);
    }
    fetch(x0, codes);
    stage(ring(x0), x0, codes);
    if (A.dnx) fetch(wrapx(x0 + 1), codes);
    __syncthreads();

#pragma unroll 1
    for (int x = x0; x < x1; ++x) {
        if (A.dnx) {
            stage(ring(x + 1), wrapx(x + 1), codes);
            if (x + 1 < x1) fetch(wrapx(x + 2), codes);
        }
        __syncthreads();
        const int pl0 = A.dnx ? ring(x - 1) : ring(x), pl1 = ring(x), pl2 = A.dnx ? ring(x + 1) : ring(x);
        const int par = (x - x0) & 1;

        // ---- classify the tile's cells: bulk cells are copied, wide cells compacted for the
        // second phase, cells the tile cannot hold deferred
        const int ly = tid / TZ, lz = tid - (tid / TZ) * TZ;
        const bool own = tid < TY * TZ && y0 + ly < NY && z0 + lz < NZ;
        const int oh = (ly + A.dny) * HZ + lz + A.dnz;
        const int64_t c = own ? (static_cast<int64_t>(x) * NY + (y0 + ly)) * NZ + (z0 + lz) : 0;
        bool wide = false, single = false;
        int slid = 0;
        double sv = 0.0;
        if (own) {
            const uint32_t inf0 = s_info[pl1][oh];
            const int n0 = static_cast<int>(inf0 & 15u);
            bool multi = n0 > 1, bad = ((inf0 >> kInfoBad) & 1u) != 0u;
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                if (k == 13) continue;
                const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
                if (!((dx == 0 || A.dnx) && (dy == 0 || A.dny) && (dz == 0 || A.dnz))) continue;
                const uint32_t inf = s_info[dx < 0 ? pl0 : (dx > 0 ? pl2 : pl1)][oh + dy * HZ + dz];
                multi = multi || (inf & 15u) > 1u;
                bad = bad || ((inf >> kInfoBad) & 1u) != 0u;
            }
            if (multi) {
                if (bad || n0 > K) s_defer[par][atomicAdd(&s_ndefer[par], 1)] = static_cast<uint16_t>(tid);
                else wide = true;
            } else if (n0 == 1 && ((inf0 >> kInfoBad) & 1u) != 0u) {
                // a lone entry whose grain did not fit the local ids
                s_defer[par][atomicAdd(&s_ndefer[par], 1)] = static_cast<uint16_t>(tid);
            } else if (n0 == 1) {
                // NodePF::finalize keeps a lone entry at 1; an FP32 list stores it implicitly
                single = true;
                slid = static_cast<int>((inf0 >> kInfoOrd) & 31u);
                A.id_out[c] = static_cast<uint16_t>(s_key[slid]);
                if (ListCode<TV>::kImplicit) {
                    sv = 1.0;
                    A.cnt_out[c] = 1;
                } else {
                    sv = static_cast<double>(s_val[pl1][0][oh]);
                    static_cast<TV*>(A.val_out)[c] = static_cast<TV>(sv);
                    A.cnt_out[c] = static_cast<uint8_t>(n0);
                }
            } else {
                A.cnt_out[c] = 0;  // empty node (reported when a step empties it)
            }
        }
        lds_unit_add(s_acc, single && sv == 1.0, slid);
        if (single && sv != 1.0) lds_volume_add(s_acc, slid, sv, bad_volume);

        const unsigned long long wm = __ballot(wide);
        if (lane == 0) s_nwide[wave] = static_cast<uint32_t>(__popcll(wm));
        __syncthreads();
        int offset = __popcll(wm & ((1ull << lane) - 1ull)), nwide = 0;
#pragma unroll
        for (int w = 0; w < kTileWaves; ++w) {
            const int n = static_cast<int>(s_nwide[w]);
            offset += w < wave ? n : 0;
            nwide += n;
        }
        if (wide) s_list[offset] = static_cast<uint16_t>(tid);
        __syncthreads();

        // ---- wide cells: merged list, Laplacians (gradients) from the staged planes, then the
        // per-cell step of compute_cell
        if (wave * 64 < nwide) {
            CellList<K, TC, GRAD> L;
            int lid[K];
#pragma unroll
            for (int t = 0; t < K; ++t) {
                L.keep[t] = false; L.eid[t] = 0u; L.phi[t] = 0.0; L.lap[t] = TC(0); lid[t] = 0;
                if constexpr (CellList<K, TC, GRAD>::kCopy) L.ev_[t] = TC(0);
                if (GRAD) { L.gr[GRAD ? t : 0][0] = L.gr[GRAD ? t : 0][1] = L.gr[GRAD ? t : 0][2] = TC(0); }
            }
            L.emax = 0.0;
            L.mmax = 0.0;
            L.n0 = 0;
            L.ne = 0;
            bool empty = false;
            const int pl[3] = {pl0, pl1, pl2};
            const int local = tid < nwide ? s_list[tid] : 0;
            const int wy = local / TZ, wz = local - (local / TZ) * TZ;
            const int wh = (wy + A.dny) * HZ + wz + A.dnz;
            const int64_t cw = (static_cast<int64_t>(x) * NY + (y0 + wy)) * NZ + (z0 + wz);
            const uint32_t inf0 = s_info[pl1][wh];
            const uint32_t m0 = s_mask[pl1][wh];
            const int n0 = static_cast<int>(inf0 & 15u);
            uint32_t seen = 0u;
#pragma unroll
            for (int t = 0; t < kRanks; ++t)
                if (t < n0) {
                    lid[t] = static_cast<int>((inf0 >> (kInfoOrd + 5 * t)) & 31u);
                    seen |= 1u << lid[t];
                }
            // merged order: own entries, then grains new to the box in stencil order (x outer,
            // z inner), within one neighbour in its list order
            int ne = n0;
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
                if (k == 13 || !((dx == 0 || A.dnx) && (dy == 0 || A.dny) && (dz == 0 || A.dnz))) continue;
                const int p = pl[dx + 1];
                const int hh = wh + dy * HZ + dz;
                const uint32_t nw = s_mask[p][hh] & ~seen;
                if (nw) {
                    const uint32_t inf = s_info[p][hh];
                    const int nk = static_cast<int>(inf & 15u);
#pragma unroll
                    for (int q = 0; q < kRanks; ++q) {
                        const int l = static_cast<int>((inf >> (kInfoOrd + 5 * q)) & 31u);
                        if (q < nk && ((nw >> l) & 1u)) {
#pragma unroll
                            for (int t = 0; t < K; ++t) lid[t] = t == ne ? l : lid[t];
                            ++ne;
                        }
                    }
                    seen |= nw;
                }
            }
            const bool go = tid < nwide && ne <= K && ne >= 2;
            if (tid < nwide && !go) {
                // the deferred-cell lists by merged length: 5-6 (and fewer than 2) entries, 7-8, 9 and
                // more (levels 0, 1 and 2 of the deferred-cell kernels; lists 1 and 2 run after list 0)
                const int dest = ne <= 6 ? 0 : (ne <= 8 ? 1 : 2);
                s_defer[par][atomicAdd(&s_ndefer[par], 1)] = static_cast<uint16_t>(local | (dest << 8));
            }
            uint32_t bm[K], lm[K];
#pragma unroll
            for (int t = 0; t < K; ++t) {
                bm[t] = go && t < ne ? 1u << lid[t] : 0u;
                lm[t] = bm[t] - 1u;
            }
            // Laplacians (and gradients) for as many entries as the wavefront's longest list
            if (__any(go && ne > 3)) tile_derivatives<4>(A, s_mask, s_val, &s_zero, pl, wh, HZ, bm, lm, L);
            else if (__any(go && ne > 2)) tile_derivatives<3>(A, s_mask, s_val, &s_zero, pl, wh, HZ, bm, lm, L);
            else if (__any(go)) tile_derivatives<2>(A, s_mask, s_val, &s_zero, pl, wh, HZ, bm, lm, L);
            if (go) {
                // own values (FP64; the FP32 implicit entry as 1 - sum of the others in slot order)
                const int imp0 = static_cast<int>((inf0 >> kInfoImp) & 15u);
                double others = 0.0;
#pragma unroll
                for (int t = 0; t < kRanks; ++t)
                    if (t < n0) {
                        const bool pres = (m0 & bm[t]) != 0u;
                        L.phi[t] = pres ? static_cast<double>(s_val[pl1][__popc(m0 & lm[t])][wh]) : 0.0;
                        L.keep[t] = true;
                        if (ListCode<TV>::kImplicit && t != imp0) others += L.phi[t];
                    }
                if (ListCode<TV>::kImplicit) {
#pragma unroll
                    for (int t = 0; t < kRanks; ++t)
                        if (t == imp0 && t < n0) L.phi[t] = 1.0 - others;
                }
#pragma unroll
                for (int t = 0; t < K; ++t) {
                    if constexpr (CellList<K, TC, GRAD>::kCopy) L.ev_[t] = static_cast<TC>(L.phi[t]);
                    L.eid[t] = t < ne ? static_cast<uint32_t>(s_key[lid[t]]) : 0u;
                }
                L.n0 = n0;
                L.ne = ne;
                constexpr int PATH = UNI ? 0 : 2;
                switch (ne) {
                case 2: compute_cell<2, true, K, TC, GRAD, DG, PATH>(A, L, empty, cw); break;
                case 3: compute_cell<3, true, K, TC, GRAD, DG, PATH>(A, L, empty, cw); break;
                default: compute_cell<4, true, K, TC, GRAD, DG, PATH>(A, L, empty, cw); break;
                }
                if (!store_cell<K, TV, TC, GRAD>(A, cw, L)) report_overflow<K, TC, GRAD>(A, cw, L);
                if (empty) atomicOr(&A.status[kStatusEmpty], 1);
#pragma unroll
                for (int t = 0; t < K; ++t)
                    if (L.keep[t]) {
                        if (L.phi[t] == 1.0) atomicAdd(&s_acc[lid[t]][tid & (kCopies - 1)][3], 1ull << 20);
                        else lds_volume_add(s_acc, lid[t], L.phi[t], bad_volume);
                    }
            }
            emax = smax(emax, L.emax);
            mmax = smax(mmax, L.mmax);
        }
        __syncthreads();  // the next plane is staged over the ring slot of x - 1
        if (wave == 0) {
            // this plane's deferred cells to the global work list; the other parity's LDS list
            // is the next plane's, this one is reused two planes on (after at least one barrier)
            const int nd = s_ndefer[par];
            if (nd > 0) {
                // per list: count, reserve with one atomic, write in LDS order
                int cnt[3] = {0, 0, 0};
                for (int i0 = 0; i0 < nd; i0 += 64) {
                    const int i = i0 + lane;
                    const int d = i < nd ? (s_defer[par][i] >> 8) : 3;
#pragma unroll
                    for (int q = 0; q < 3; ++q) cnt[q] += __popcll(__ballot(d == q));
                }
                unsigned base[3] = {0u, 0u, 0u};
                if (lane == 0)
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                        if (cnt[q]) base[q] = atomicAdd(&A.work_count[q], static_cast<unsigned>(cnt[q]));
#pragma unroll
                for (int q = 0; q < 3; ++q) base[q] = static_cast<unsigned>(__shfl(static_cast<int>(base[q]), 0));
                for (int i0 = 0; i0 < nd; i0 += 64) {
                    const int i = i0 + lane;
                    const int e = i < nd ? s_defer[par][i] : 0;
                    const int d = i < nd ? (e >> 8) : 3, l = e & 255;
                    const int dy = l / TZ, dz = l - (l / TZ) * TZ;
                    const uint32_t cellv = static_cast<uint32_t>((static_cast<int64_t>(x) * NY + (y0 + dy)) * NZ + (z0 + dz));
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const unsigned long long m = __ballot(d == q);
                        if (d == q)
                            A.work_list[static_cast<int64_t>(q) * N + base[q] + __popcll(m & ((1ull << lane) - 1ull))] = cellv;
                        base[q] += static_cast<unsigned>(__popcll(m));
                    }
                }
                if (lane == 0) s_ndefer[par] = 0;
            }
        }
    }
    note_maxima(A.obs, emax, mmax);

    // ---- block's volume sums to the global accumulators (one thread per local id and limb)
    if (tid < kLocal * kLimbs) {
        const int l = tid / kLimbs, i = tid - l * kLimbs;
        const int g = s_key[l];
        unsigned long long sum = 0ull;
#pragma unroll
        for (int q = 0; q < kCopies; ++q) sum += s_acc[l][q][i];
        if (g >= 0 && sum) atomicAdd(&A.vol_acc[static_cast<int64_t>(g) * kLimbs + i], sum);
    }
    if (bad_volume) atomicOr(&A.status[kStatusVolumeRange], 1);
}

template <typename TV, typename TC, bool GRAD, bool DG, bool UNI>
__global__ __launch_bounds__(kBlock, sizeof(TV) == 4 && !GRAD ? 4 : 1) void tile_kernel(const StepArgs A)
{
    tile_block<TV, TC, GRAD, DG, UNI>(A, A.xcd_order ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x);
}

// Several domains in one launch (ensembles): block b belongs to the domain d with
// first[d] <= b < first[d+1], each with its own tile shape.
template <typename TV, typename TC, bool GRAD, bool DG, bool UNI>
__global__ __launch_bounds__(kBlock, sizeof(TV) == 4 && !GRAD ? 4 : 1) void tile_batch_kernel(
    const StepArgs* __restrict__ batch, const uint32_t* __restrict__ first, int32_t ndom)
{
    const int dom = batch_domain(first, ndom, blockIdx.x);
    tile_block<TV, TC, GRAD, DG, UNI>(batch[dom], blockIdx.x - first[dom]);
}

}  // namespace

int launch_tile(const StepArgs& a, unsigned blocks, bool fp32, bool full, bool uni, void* stream)
{
    const dim3 grid(blocks);
    hipStream_t s = static_cast<hipStream_t>(stream);
#define MPF_TILE(TV, G, D, U) hipLaunchKernelGGL((tile_kernel<TV, TV, G, D, U>), grid, dim3(kBlock), 0, s, a)
    if (fp32) {
        if (full)     MPF_TILE(float, true, true, false);
        else if (uni) MPF_TILE(float, false, false, true);
        else          MPF_TILE(float, false, false, false);
    } else {
        if (full)     MPF_TILE(double, true, true, false);
        else if (uni) MPF_TILE(double, false, false, true);
        else          MPF_TILE(double, false, false, false);
    }
#undef MPF_TILE
    return static_cast<int>(hipGetLastError());
}

int launch_tile_batch(const StepArgs* args, const uint32_t* first, int ndom, unsigned blocks, bool fp32, bool full,
                      bool uni, void* stream)
{
    const dim3 grid(blocks);
    hipStream_t s = static_cast<hipStream_t>(stream);
#define MPF_TILE(TV, G, D, U) \
    hipLaunchKernelGGL((tile_batch_kernel<TV, TV, G, D, U>), grid, dim3(kBlock), 0, s, args, first, ndom)
    if (fp32) {
        if (full)     MPF_TILE(float, true, true, false);
        else if (uni) MPF_TILE(float, false, false, true);
        else          MPF_TILE(float, false, false, false);
    } else {
        if (full)     MPF_TILE(double, true, true, false);
        else if (uni) MPF_TILE(double, false, false, true);
        else          MPF_TILE(double, false, false, false);
    }
#undef MPF_TILE
    return static_cast<int>(hipGetLastError());
}

}  // namespace mpf
