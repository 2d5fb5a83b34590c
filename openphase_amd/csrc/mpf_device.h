// Shared definitions of the MPF step kernels and the host library (no torch types).
#pragma once

#include <stdint.h>

namespace mpf {

constexpr int kMaxSlots = 16;                        // largest per-cell list capacity of the stored state
constexpr int kBlock = 256;                          // threads per block of the step kernels
constexpr int kWideBlock = 64;                       // threads per block of the deferred-cell kernel
constexpr int kWideGrid = 4096;                      // blocks of the deferred-cell kernels (grid-stride)
constexpr int kLimbs = 4;                            // 32-bit limbs of the exact per-grain volume sums

// Device status words (one int each).
enum StatusWord { kStatusOverflow = 0, kStatusEmpty = 1, kStatusVolumeRange = 2, kStatusWords = 4 };

struct StepArgs {
    int64_t ncells;
    int32_t nx, ny, nz;          // storage sizes (a reduced axis has size 1)
    int32_t dnx, dny, dnz;       // 1 if the axis is active
    int32_t slots;               // stored-list capacity (storage stride), <= kMaxSlots
    double wl[4];                // Laplacian weight by number of non-zero offset components, / dx^2
    double wg[4];                // gradient weight by number of non-zero offset components, / dx
    double prefactor;            // Pi^2 / Eta^2
    double dg_prefactor;         // Pi / Eta
    double dt;
    int32_t nphases;
    int32_t consider_nucleus;
    int32_t tj_model;            // 0 MODEL, 1 VALUE, 2 NONE
    double tj_factor;
    double tj_energy;
    // 1 when every phase pair has the same ISO energy and mobility and no triple-junction term
    // can act at VolumeRatio 1 (MODEL with factor 0, or NONE): then the pairwise gamma sums
    // (sigma_bg - sigma_ag) * (...) are exact zeros and are skipped
    int32_t uniform;
    double sigma0;               // the common energy / mobility when uniform
    double mu0;
    const double* sigma;         // nphases^2 phase-pair interface energy
    const double* mu;            // nphases^2 phase-pair mobility
    const double* eps;           // nphases^2 cubic anisotropy strength
    const double* max_energy;    // nphases^2 MaxEnergy of the energy model
    const int32_t* emodel;       // nphases^2 energy model (0 ISO, 1 CUBIC)
    const int32_t* dg_mode;      // nphases^2, 1 = constant driving force for the pair
    const double* dg_value;      // nphases^2 antisymmetric driving force
    const double* dg_allowed;    // nphases^2 tanh limiter amplitude
    // driving force averaged over the interface for this step (DrivingForce::Average), one value
    // per pair of own entries a < b at plane b(b-1)/2 + a, [pairs][ncells]; null: the phase-pair
    // value (SkipAverage)
    const double* dg_avg;
    // InterfaceRegularization (DoubleObstacle's curvature overload, in the full kernel variant):
    // reg = 1 scales a pair's increment left to right, loc_increment * mobility * norm_1 * scale;
    // reg_active (RegularizationFactor or CurvatureFactor != 1) then multiplies it by reg_factor
    // and, on interface cells, adds the curvature averaged over the interface (`curv`, one value
    // per pair of own entries, same layout as dg_avg) * mobility * sqrt(phi_a phi_b) * Pi/Eta;
    // curv is null when the curvature is not averaged (its merge weight is then 0)
    int32_t reg;
    int32_t reg_active;
    double reg_factor;
    const double* curv;
    const int32_t* gphase;       // per grain: thermodynamic phase
    const double* gvr;           // per grain: VolumeRatio
    const double* grot;          // per grain: rotation matrix, row-major
    const uint8_t* gsolid;       // per grain: 1 if solid
    const uint8_t* cnt_in;
    const uint16_t* id_in;
    const void* val_in;
    uint8_t* cnt_out;
    uint16_t* id_out;
    void* val_out;
    unsigned long long* vol_acc; // kLimbs per grain
    int32_t* status;             // kStatusWords
    unsigned long long* first_overflow;
    unsigned long long* warnings;
    // largest interface energy and mobility of the pairs this step's IP.Set evaluates (FP64 bit
    // patterns of non-negative values, atomicMax): InterfaceProperties maxEnergies/maxMobilities
    unsigned long long* obs;
    uint32_t* work_list;         // deferred cells: four lists of ncells each (mid kernel levels 0-2; wide kernel)
    unsigned* work_count;        // [4] lengths of the four lists
    // launch shape (chosen by mpf_autotune): XCD-contiguous block order of the step kernel, and
    // the number of blocks of the grid-stride deferred-cell kernels
    int32_t xcd_order;
    int32_t deferred_grid;
    // LDS-tiled step kernel (mpf_tile.hip): tile of tile_y x tile_z cells of the (y, z) plane
    // marching through tile_x planes along x; tile_x = 0 selects the per-cell step kernel
    int32_t tile_x, tile_y, tile_z;
};

struct GrainArgs {
    int32_t ngrains;
    unsigned long long* vol_acc;
    double* volume;
    double* maxvolume;
    const double* refvolume;
    double* vr;
    uint8_t* exist;
    int32_t* stage;              // 0 seed, 1 nucleus, 2 stable
    unsigned* work_count;        // [4], reset here for the next step
    unsigned long long* stats;   // [2] cumulative lengths of work lists 0 and 1
    unsigned long long* obs;     // [4] this step's maxima (StepArgs::obs, cleared here) and the last step's
};

// Adjoint of one step (mpf_adjoint.hip). `s` describes the state S_t the step starts from.
struct AdjointArgs {
    StepArgs s;
    const double* lam_next;      // [slots][ncells] cotangent of the values of S_{t+1}
    const uint8_t* next_cnt;     // count bytes of S_{t+1} (null: S_{t+1} holds the entries the
    const uint16_t* next_id;     // step keeps, in list order); [slots][ncells] its grains
    double* lam_prev;            // [slots][ncells] cotangent of the values of S_t (written)
    double* lam_own;             // [kAdjSlots][ncells] direct part per own entry
    uint8_t* adj_cnt;            // [ncells] merged entries of an interface cell (0: bulk)
    uint8_t* defer;              // [ncells] 1: the cell is on the work list
    uint32_t* list;              // [ncells] work list: cells of three or more merged entries
    unsigned* list_count;        // its length
    uint32_t* list_long;         // [ncells] second work list: cells of more than four merged entries
    unsigned* long_count;        // its length
    double* cellg;               // [ncells][3 nphases^2] table cotangents of listed cells
    uint16_t* adj_id;            // [kAdjSlots][ncells]
    double* adj_lap;             // [kAdjSlots][ncells] cotangent of the Laplacians
    double* adj_grad;            // [3 kAdjSlots][ncells] cotangent of the gradients (may be null)
    double* partial;             // [blocks][3 nphases^2] per-block table cotangents
    double* grad;                // [3][nphases^2] cotangents of sigma, mu, driving force (accumulated)
    double dg_limit;
    double regularization;
    int32_t grad_energy;         // orientation-dependent energies present
    int32_t driving_force;       // constant driving force present
};
constexpr int kAdjSlots = 8;     // merged entries per cell the adjoint handles
constexpr int kAdjPhases = 4;    // thermodynamic phases the adjoint handles

int64_t adjoint_blocks(int64_t ncells);
int launch_adjoint(const AdjointArgs& a, bool fp32, void* stream);

constexpr int kGrainBlock = 256;                     // threads per block of the grain kernel
constexpr int kTileHalo = 396;                       // cells of a tile plane with its halo, at most (6 x 66)

// LDS-tiled step kernel over one domain (mpf_tile.hip); `blocks` = tiles x x-chunks.
int launch_tile(const StepArgs& a, unsigned blocks, bool fp32, bool full, bool uni, void* stream);
// The same over several domains (device array `args`, first[d] = the first block of domain d,
// ndom + 1 prefix sums of the domains' tile counts), each with its own tile shape.
int launch_tile_batch(const StepArgs* args, const uint32_t* first, int ndom, unsigned blocks, bool fp32, bool full,
                      bool uni, void* stream);

// Driving-force averaging (mpf_dgavg.hip): CollectAverage into `tmp`, DistributeAverage into
// `avg` (= s.dg_avg), both [pairs][ncells] with pair a < b of the own entries at b(b-1)/2 + a.
struct AvgArgs {
    StepArgs s;                  // state of the step start (cnt_in, id_in, val_in), grains, tables
    double* tmp;
    double* avg;
    // offsets of the averaging sphere in the reference's loop order (x outer, z inner; each axis
    // within min(Range, N - 1), 0 on reduced axes) that have Range - |offset| > 0: 3 ints each,
    // and that distance weight
    const int32_t* off;
    const double* off_w;
    int32_t noff;
    int32_t weights;             // 0 PHASEFIELDS, 1 RANGE, 2 COUNTER
    double threshold;            // sqrt(PhiThreshold * (1 - PhiThreshold))
};
constexpr int kAvgBlock = 256;
int launch_dg_average(const AvgArgs& a, bool fp32, void* stream);

// Interface regularization (mpf_dgavg.hip), FP64 state: the raw curvature of every own pair of
// the interface cells into `raw`, then InterfaceRegularization's CollectAverage into `tmp` and
// DistributeAverage into `avg` (= s.curv), all [pairs][ncells] like AvgArgs.
struct CurvArgs {
    StepArgs s;                  // state of the step start, grains, tables
    double* raw;
    double* tmp;
    double* avg;
    const int32_t* off;          // sphere offsets within Range (3 ints each), x outer, z inner
    const double* off_w;         // Range - |offset|
    int32_t noff;
    double threshold;            // sqrt(PhiThreshold * (1 - PhiThreshold))
    double factor;               // CurvatureFactor - RegularizationFactor
    double prefactor2;           // Eta / Pi
    int32_t grad;                // orientation-dependent energies present (gradients needed)
};
int launch_curvature(const CurvArgs& a, bool average, void* stream);

// Host-side launchers (mpf_kernels.hip); return a hipError_t value. The step, deferred-cell and
// grain kernels run over one or several domains of the same kernel variant in one launch: `args`
// (device memory) holds one argument block per domain and `first` (device, ndom + 1 entries) the
// prefix sums of the domains' block counts (the step kernel takes a single domain's arguments by
// value in `one` when ndom is 1, and ignores `args` then); the deferred-cell kernels give every domain
// `per_domain` blocks. `full`: orientation-dependent energies and/or constant driving force;
// `uni`: the equal-energy form.
int launch_step(const StepArgs& one, const StepArgs* args, const uint32_t* first, int ndom, unsigned blocks, bool fp32,
                bool full, bool uni, void* stream);
int launch_deferred(const StepArgs* args, int ndom, int per_domain, bool fp32, bool full, void* stream);
// the variants behind launch_step / launch_deferred, one translation unit each
#define MPF_VARIANT_LAUNCHERS(NS)                                                                            \
    namespace NS {                                                                                           \
    int launch_step(const StepArgs& one, const StepArgs* args, const uint32_t* first, int ndom, unsigned blocks, \
                    bool uni, void* stream);                                                                 \
    int launch_deferred(const StepArgs* args, int ndom, int per_domain, void* stream);                       \
    }
MPF_VARIANT_LAUNCHERS(v64_plain)
MPF_VARIANT_LAUNCHERS(v64_full)
MPF_VARIANT_LAUNCHERS(v32_plain)
MPF_VARIANT_LAUNCHERS(v32_full)
#undef MPF_VARIANT_LAUNCHERS
int launch_finalize_init(const StepArgs& a, bool fp32, void* stream);

// Interface energy (mpf_energy.hip): `cur` the state summed over (cnt_in, id_in, val_in), grains
// and tables; ip_* the state the last IP.Set saw (null: none, every energy 0).
struct EnergyArgs {
    StepArgs cur;
    const uint8_t* ip_cnt;
    const uint16_t* ip_id;
    const void* ip_val;
    double eta;                  // interface width
    double pref;                 // Eta^2 / Pi^2
    double* partial;             // energy_blocks(ncells) per-block sums
    int32_t* status;             // set to 1 when a merged list exceeds the step's capacity
};
int64_t energy_blocks(int64_t ncells);
// Writes the sum over all cells to `out` (device) after the per-block sums.
int launch_energy(const EnergyArgs& a, bool fp32, double* out, void* stream);
// Per-cell phase fractions of the state (cnt_in, id_in, val_in): out[phase * ncells + cell].
int launch_fractions(const StepArgs& a, bool fp32, double* out, void* stream);
int launch_grains(const GrainArgs* args, const uint32_t* first, int ndom, unsigned blocks, void* stream);
int launch_copy(const void* src, void* dst, int64_t bytes, void* stream);

}  // namespace mpf
