// Host side of the C ABI declared in include/mpf.h.
//
// Owns the device state (two slot-major SoA buffers, swapped every step), the grain table and the
// phase-pair tables, and enqueues the kernels of mpf_kernels.hip on the context's stream.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mpf.h"
#include "mpf_device.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

template <typename T> hipError_t dalloc(T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
}

template <typename T> void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Host-to-device copy of n elements; the first failure of a sequence is kept in `e`.
template <typename T> void h2d(hipError_t& e, T* dst, const T* src, size_t n)
{
    if (e == hipSuccess) e = hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace

struct mpf_ctx {
    mpf_grid grid{};
    int nx = 1, ny = 1, nz = 1, dnx = 0, dny = 0, dnz = 0, active = 0;
    int64_t ncells = 0;
    int slots = 8;
    bool fp32 = false, grad = false, dg = false;
    int nphases = 1;
    std::vector<double> sigma, mu, eps, maxe, dgval, dgallowed;
    std::vector<int32_t> emodel, dgmode;
    double tj_factor = 0.0, tj_energy = 0.0;
    bool uniform = false;            // see StepArgs::uniform
    int tj_model = 0;
    int consider_nucleus = 1;
    double nucleus_volume_factor = 1.0, regularization = 1.0, dg_limit = 0.0;
    double wl[4] = {0, 0, 0, 0}, wg[4] = {0, 0, 0, 0};
    double prefactor = 0.0, dg_prefactor = 0.0;
    // driving-force averaging (DrivingForce::Average); tmp and averaged value per pair of own entries
    bool dg_avg = false;
    int avg_range = 0, avg_weights = 0;
    double avg_threshold = 0.0;
    std::vector<int32_t> avg_off;    // sphere offsets (3 each) and their distance weights
    std::vector<double> avg_off_w;
    double* d_avg_tmp = nullptr;
    double* d_avg = nullptr;
    int32_t* d_avg_off = nullptr;
    double* d_avg_off_w = nullptr;
    // interface regularization (InterfaceRegularization): raw, tmp and averaged curvature per pair
    // of own entries, and the averaging sphere
    bool reg = false, reg_active = false, reg_average = true;
    double curvature_factor = 1.0, reg_threshold = 0.0;
    int reg_range = 0;
    std::vector<int32_t> reg_off;
    std::vector<double> reg_off_w;
    double* d_curv_raw = nullptr;
    double* d_curv_tmp = nullptr;
    double* d_curv = nullptr;
    int32_t* d_reg_off = nullptr;
    double* d_reg_off_w = nullptr;
    // per-step terms (mpf_set_step_terms): the driving force acts / is averaged in the next steps
    bool dg_on = true, avg_on = true;
    int32_t* d_dgmode_off = nullptr;  // all-zero mode table: no pair takes the driving force

    hipStream_t stream = nullptr;
    uint8_t* cnt[2] = {nullptr, nullptr};
    uint16_t* id[2] = {nullptr, nullptr};
    void* val[2] = {nullptr, nullptr};
    int cur = 0;

    int ngrains = 0;
    int32_t* g_phase = nullptr;
    double* g_vr = nullptr;
    double* g_rot = nullptr;
    uint8_t* g_solid = nullptr;
    double* g_volume = nullptr;
    double* g_maxvolume = nullptr;
    double* g_refvolume = nullptr;
    uint8_t* g_exist = nullptr;
    int32_t* g_stage = nullptr;
    unsigned long long* g_acc = nullptr;

    double* d_sigma = nullptr;
    double* d_mu = nullptr;
    double* d_eps = nullptr;
    double* d_maxe = nullptr;
    int32_t* d_emodel = nullptr;
    int32_t* d_dgmode = nullptr;
    double* d_dgval = nullptr;
    double* d_dgallowed = nullptr;

    // adjoint work arrays (allocated on first use)
    double* a_lam_own = nullptr;
    uint8_t* a_cnt = nullptr;
    uint16_t* a_id = nullptr;
    double* a_lap = nullptr;
    double* a_grad = nullptr;
    double* a_partial = nullptr;      // per-block table cotangents
    uint8_t* a_defer = nullptr;       // cells on the work list
    uint32_t* a_list = nullptr;       // work lists of cells with longer merged lists (two of N)
    unsigned* a_list_count = nullptr; // their lengths
    double* a_cellg = nullptr;        // per listed cell table cotangents

    int32_t* d_status = nullptr;
    unsigned long long* d_first_overflow = nullptr;
    unsigned long long* d_warnings = nullptr;
    uint32_t* d_work_list = nullptr;
    unsigned* d_work_count = nullptr;
    unsigned long long* d_stats = nullptr;
    unsigned long long* d_obs = nullptr;  // IP.Set maxima: this step's (2 words), the last step's (2)
    bool stepped = false;                 // a step ran since creation (IP.Set maxima exist)
    // the other buffer still holds the state the last step started from (interface energies of
    // that step's IP.Set)
    bool prev_valid = false;
    double* d_energy = nullptr;           // per-block partial sums of the interface energy, then the sum
    size_t energy_cap = 0;
    double* d_fractions = nullptr;        // phase fractions for host downloads
    size_t fractions_cap = 0;

    int xcd_order = 1;               // launch shape, see StepArgs
    int deferred_grid = mpf::kWideGrid;
    int tile_x = 0, tile_y = 0, tile_z = 0;  // LDS-tiled step kernel shape (tile_x = 0: per-cell kernel)

    // device copies of the kernel arguments (both buffer parities), uploaded when they change
    mpf::StepArgs h_args[2]{};
    mpf::GrainArgs h_gargs{};
    bool args_ok = false;
    mpf::StepArgs* d_args = nullptr;
    uint32_t* d_first = nullptr;
    mpf::GrainArgs* d_gargs = nullptr;
    uint32_t* d_gfirst = nullptr;

    bool have_state = false, have_grains = false, initialized = false;
    bool timing = false;
    std::vector<hipEvent_t> events;
    double last_step_ms = 0.0, last_grain_ms = 0.0;
    std::string err;

    mpf_status fail(mpf_status s, const std::string& msg)
    {
        err = msg;
        return s;
    }
    mpf_status hip(hipError_t e, const char* what)
    {
        if (e == hipSuccess) return MPF_OK;
        err = std::string(what) + ": " + hipGetErrorString(e);
        return MPF_ERR_HIP;
    }
    size_t vsize() const { return fp32 ? sizeof(float) : sizeof(double); }
};

// Members of one kernel variant (precision, orientation/driving-force terms, equal-energy form)
// share every launch: device arrays of their arguments for the two captured steps.
struct mpf_group {
    bool fp32 = false, full = false, uni = false, tiled = false;
    std::vector<int> members;
    std::vector<mpf::StepArgs> h_a1, h_a2;
    std::vector<mpf::GrainArgs> h_g;
    std::vector<uint32_t> h_first, h_gfirst;
    mpf::StepArgs* d_a1 = nullptr;
    mpf::StepArgs* d_a2 = nullptr;
    mpf::GrainArgs* d_g = nullptr;
    uint32_t* d_first = nullptr;
    uint32_t* d_gfirst = nullptr;
    int per_domain = 256;
};

struct mpf_ensemble {
    std::vector<mpf_ctx*> members;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<mpf_group> groups;
    std::vector<hipEvent_t> join;
    std::string err;
};

namespace {

// The driving force acts in the next step: the context has one and mpf_set_step_terms left it on.
bool dg_active(const mpf_ctx* c) { return c->dg && c->dg_on; }
// DrivingForce::Average averages it over the interface in the next step.
bool avg_active(const mpf_ctx* c) { return c->dg_avg && c->dg_on && c->avg_on; }
// Kernel variant of the next step: orientation-dependent energies and/or an acting driving force.
bool full_variant(const mpf_ctx* c) { return c->grad || dg_active(c) || c->reg; }

void free_grains(mpf_ctx* c)
{
    dfree(c->g_phase); dfree(c->g_vr); dfree(c->g_rot); dfree(c->g_solid);
    dfree(c->g_volume); dfree(c->g_maxvolume); dfree(c->g_refvolume);
    dfree(c->g_exist); dfree(c->g_stage); dfree(c->g_acc);
}

mpf::StepArgs step_args(const mpf_ctx* c, double dt, int src)
{
    mpf::StepArgs a{};
    a.ncells = c->ncells;
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz;
    a.dnx = c->dnx; a.dny = c->dny; a.dnz = c->dnz;
    a.slots = c->slots;
    for (int i = 0; i < 4; ++i) { a.wl[i] = c->wl[i]; a.wg[i] = c->wg[i]; }
    a.prefactor = c->prefactor;
    a.dg_prefactor = c->dg_prefactor;
    a.dt = dt;
    a.nphases = c->nphases;
    a.consider_nucleus = c->consider_nucleus;
    a.tj_model = c->tj_model;
    a.tj_factor = c->tj_factor;
    a.tj_energy = c->tj_energy;
    a.uniform = c->uniform ? 1 : 0;
    a.sigma0 = c->sigma[0];
    a.mu0 = c->mu[0];
    a.sigma = c->d_sigma; a.mu = c->d_mu; a.eps = c->d_eps; a.max_energy = c->d_maxe;
    a.emodel = c->d_emodel; a.dg_value = c->d_dgval; a.dg_allowed = c->d_dgallowed;
    a.dg_mode = dg_active(c) ? c->d_dgmode : c->d_dgmode_off;
    a.dg_avg = avg_active(c) ? c->d_avg : nullptr;
    a.reg = c->reg ? 1 : 0;
    a.reg_active = c->reg_active ? 1 : 0;
    a.reg_factor = c->regularization;
    a.curv = c->reg_active && c->reg_average ? c->d_curv : nullptr;
    a.gphase = c->g_phase; a.gvr = c->g_vr; a.grot = c->g_rot; a.gsolid = c->g_solid;
    a.cnt_in = c->cnt[src]; a.id_in = c->id[src]; a.val_in = c->val[src];
    a.cnt_out = c->cnt[1 - src]; a.id_out = c->id[1 - src]; a.val_out = c->val[1 - src];
    a.vol_acc = c->g_acc;
    a.status = c->d_status;
    a.first_overflow = c->d_first_overflow;
    a.warnings = c->d_warnings;
    a.obs = c->d_obs;
    a.work_list = c->d_work_list;
    a.work_count = c->d_work_count;
    a.xcd_order = c->xcd_order;
    a.deferred_grid = c->deferred_grid;
    a.tile_x = c->tile_x;
    a.tile_y = c->tile_y;
    a.tile_z = c->tile_z;
    return a;
}

mpf::GrainArgs grain_args(const mpf_ctx* c)
{
    mpf::GrainArgs g{};
    g.ngrains = c->ngrains;
    g.vol_acc = c->g_acc;
    g.volume = c->g_volume;
    g.maxvolume = c->g_maxvolume;
    g.refvolume = c->g_refvolume;
    g.vr = c->g_vr;
    g.exist = c->g_exist;
    g.stage = c->g_stage;
    g.work_count = c->d_work_count;
    g.stats = c->d_stats;
    g.obs = c->d_obs;
    return g;
}

unsigned step_blocks(const mpf_ctx* c) { return static_cast<unsigned>((c->ncells + mpf::kBlock - 1) / mpf::kBlock); }

// Tile shape of the LDS-tiled step kernel for this grid: z (the contiguous axis) up to 64 cells
// wide in 3D (256 when y is reduced), y filling 256 threads, the halo plane within kTileHalo; x
// chunks sized for about 8192 blocks (256^3 FP64: 2.38 ms at 8 planes per block, 2.42 at 4, 2.45
// at 16, 2.61 at 32). 1D grids keep the per-cell kernel (tile_x = 0).
void tile_shape(const mpf_ctx* c, int tile_x, int* ty, int* tz, int* tx)
{
    *ty = *tz = *tx = 0;
    if (c->active < 2) return;
    // 16 x 16 in 3D (the tile the autotuner picks at 256^3: a smaller halo than 4 x 64), 64 cells
    // along z in 2D (y, z) planes, whole z rows otherwise
    int z = c->dnz ? std::min(c->nz, c->dny ? (c->dnx ? 16 : 64) : 256) : 1;
    int y = c->dny ? std::min(c->ny, std::max(1, mpf::kBlock / z)) : 1;
    while ((y + 2 * c->dny) * (z + 2 * c->dnz) > mpf::kTileHalo && y > 1) --y;
    while ((y + 2 * c->dny) * (z + 2 * c->dnz) > mpf::kTileHalo && z > 1) --z;
    int x = tile_x;
    if (x <= 0) {
        const int64_t per = static_cast<int64_t>(y) * z * 8192;
        x = static_cast<int>(std::min<int64_t>(32, std::max<int64_t>(1, c->ncells / per)));
    }
    if (!c->dnx) x = 1;
    *ty = y;
    *tz = z;
    *tx = std::min(x, c->nx);
}

unsigned tile_blocks(const mpf_ctx* c)
{
    const int64_t nty = (c->ny + c->tile_y - 1) / c->tile_y, ntz = (c->nz + c->tile_z - 1) / c->tile_z;
    const int64_t nxc = (c->nx + c->tile_x - 1) / c->tile_x;
    return static_cast<unsigned>(nxc * nty * ntz);
}

// The main step kernel of one domain: the LDS-tiled kernel when a tile shape is set, else the
// per-cell kernel.
hipError_t launch_main(const mpf_ctx* c, const mpf::StepArgs& a, hipStream_t stream)
{
    const bool full = full_variant(c);
    if (c->tile_x > 0) return static_cast<hipError_t>(mpf::launch_tile(a, tile_blocks(c), c->fp32, full, c->uniform, stream));
    return static_cast<hipError_t>(mpf::launch_step(a, nullptr, nullptr, 1, step_blocks(c), c->fp32, full, c->uniform, stream));
}
unsigned grain_blocks(const mpf_ctx* c)
{
    const int n = c->ngrains > 0 ? c->ngrains : 1;
    return static_cast<unsigned>((n + mpf::kGrainBlock - 1) / mpf::kGrainBlock);
}

// Makes the device copies of the kernel arguments current for time step dt (both parities).
// Uploads only when something changed (dt, tables, launch shape, grains); then it waits for the
// copies, so the host arrays can be reused.
mpf_status sync_args(mpf_ctx* c, double dt, hipStream_t stream)
{
    const mpf::StepArgs a[2] = {step_args(c, dt, 0), step_args(c, dt, 1)};
    const mpf::GrainArgs g = grain_args(c);
    if (c->args_ok && std::memcmp(a, c->h_args, sizeof(a)) == 0 && std::memcmp(&g, &c->h_gargs, sizeof(g)) == 0)
        return MPF_OK;
    if (!c->d_args) {
        if (c->hip(dalloc(&c->d_args, 2), "argument blocks") || c->hip(dalloc(&c->d_first, 2), "argument blocks") ||
            c->hip(dalloc(&c->d_gargs, 1), "argument blocks") || c->hip(dalloc(&c->d_gfirst, 2), "argument blocks"))
            return MPF_ERR_NOMEM;
    }
    std::memcpy(c->h_args, a, sizeof(a));
    c->h_gargs = g;
    const uint32_t first[2] = {0u, step_blocks(c)}, gfirst[2] = {0u, grain_blocks(c)};
    hipError_t e = hipMemcpyAsync(c->d_args, c->h_args, sizeof(a), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_first, first, sizeof(first), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_gargs, &c->h_gargs, sizeof(g), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_gfirst, gfirst, sizeof(gfirst), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    c->args_ok = e == hipSuccess;
    return c->hip(e, "argument upload");
}

mpf::AvgArgs avg_args(const mpf_ctx* c, const mpf::StepArgs& s)
{
    mpf::AvgArgs a{};
    a.s = s;
    a.tmp = c->d_avg_tmp;
    a.avg = c->d_avg;
    a.off = c->d_avg_off;
    a.off_w = c->d_avg_off_w;
    a.noff = static_cast<int32_t>(c->avg_off_w.size());
    a.weights = c->avg_weights;
    a.threshold = c->avg_threshold;
    return a;
}

mpf::CurvArgs curv_args(const mpf_ctx* c, const mpf::StepArgs& s)
{
    mpf::CurvArgs a{};
    a.s = s;
    a.raw = c->d_curv_raw;
    a.tmp = c->d_curv_tmp;
    a.avg = c->d_curv;
    a.off = c->d_reg_off;
    a.off_w = c->d_reg_off_w;
    a.noff = static_cast<int32_t>(c->reg_off_w.size());
    a.threshold = std::sqrt(c->reg_threshold * (1.0 - c->reg_threshold));
    a.factor = c->curvature_factor - c->regularization;
    a.prefactor2 = c->grid.iwidth * c->grid.dx / kPi;
    a.grad = c->grad ? 1 : 0;
    return a;
}

mpf_status launch_step_part(mpf_ctx* c, hipStream_t stream)
{
    if (avg_active(c)) {
        mpf_status st = c->hip(static_cast<hipError_t>(mpf::launch_dg_average(avg_args(c, c->h_args[c->cur]), c->fp32, stream)),
                               "driving-force averaging kernels");
        if (st) return st;
    }
    if (c->reg_active && c->reg_average) {
        // the curvature only enters through its average: unaveraged, its merge weight is 0
        mpf_status st = c->hip(static_cast<hipError_t>(mpf::launch_curvature(curv_args(c, c->h_args[c->cur]), true, stream)),
                               "interface regularization kernels");
        if (st) return st;
    }
    return c->hip(launch_main(c, c->h_args[c->cur], stream), "step kernel");
}

mpf_status launch_rest_part(mpf_ctx* c, hipStream_t stream)
{
    const bool full = full_variant(c);
    mpf_status st = c->hip(static_cast<hipError_t>(mpf::launch_deferred(c->d_args + c->cur, 1, c->deferred_grid, c->fp32,
                                                                        full, stream)), "deferred-cell kernels");
    if (!st) st = c->hip(static_cast<hipError_t>(mpf::launch_grains(c->d_gargs, c->d_gfirst, 1, grain_blocks(c), stream)),
                         "grain kernel");
    return st;
}

// Enqueues one step (step kernel, deferred-cell kernels, grain kernel) on `stream` and flips the
// double buffer.
mpf_status enqueue_step(mpf_ctx* c, double dt, hipStream_t stream)
{
    mpf_status st = sync_args(c, dt, stream);
    if (!st) st = launch_step_part(c, stream);
    if (!st) st = launch_rest_part(c, stream);
    if (!st) c->cur = 1 - c->cur;
    return st;
}

mpf_status read_status(mpf_ctx* c)
{
    int32_t st[mpf::kStatusWords];
    mpf_status s = c->hip(hipMemcpyAsync(st, c->d_status, sizeof(st), hipMemcpyDeviceToHost, c->stream),
                          "status readback");
    if (s) return s;
    s = c->hip(hipStreamSynchronize(c->stream), "stream synchronize");
    if (s) return s;
    if (st[mpf::kStatusOverflow]) return c->fail(MPF_SLOT_OVERFLOW, "a cell needed more phase-field slots than the capacity");
    if (st[mpf::kStatusEmpty]) return c->fail(MPF_EMPTY_NODE, "finalize erased every entry of a cell");
    if (st[mpf::kStatusVolumeRange]) return c->fail(MPF_ERR_STATE, "phase-field value outside [0, 2) in the volume sum");
    return MPF_OK;
}

}  // namespace

extern "C" {

mpf_status mpf_create(const mpf_grid* grid, const mpf_params* params, mpf_ctx** out)
{
    if (!grid || !params || !out) return MPF_ERR_ARG;
    *out = nullptr;
    mpf_ctx* c = new (std::nothrow) mpf_ctx();
    if (!c) return MPF_ERR_NOMEM;
    c->grid = *grid;
    auto bail = [&](mpf_status s, const char* msg) {
        std::fprintf(stderr, "mpf_create: %s\n", msg);
        mpf_destroy(c);
        return s;
    };
    if (grid->nx < 0 || grid->ny < 0 || grid->nz < 0 || grid->nx + grid->ny + grid->nz == 0)
        return bail(MPF_ERR_ARG, "grid sizes must be >= 0 and not all zero");
    if (!(grid->dx > 0.0) || !(grid->iwidth > 0.0)) return bail(MPF_ERR_ARG, "dx and iwidth must be positive");
    if (grid->slots < 2 || grid->slots > mpf::kMaxSlots) return bail(MPF_ERR_ARG, "slots must be in [2, 16]");
    if (grid->precision == MPF_FP32 && grid->slots > 15) return bail(MPF_ERR_ARG, "FP32 states hold at most 15 slots");
    if (grid->precision != MPF_FP64 && grid->precision != MPF_FP32) return bail(MPF_ERR_ARG, "unknown precision");
    for (int i = 0; i < 6; ++i)
        if (grid->bc[i] != MPF_BC_PERIODIC) return bail(MPF_ERR_UNSUPPORTED, "only periodic boundary conditions are implemented");
    if (params->nphases < 1 || !params->sigma || !params->mu) return bail(MPF_ERR_ARG, "nphases, sigma and mu are required");
    if (params->tj_model < 0 || params->tj_model > 2) return bail(MPF_ERR_ARG, "unknown triple-junction model");

    c->dnx = grid->nx != 0; c->dny = grid->ny != 0; c->dnz = grid->nz != 0;
    c->nx = grid->nx ? grid->nx : 1;
    c->ny = grid->ny ? grid->ny : 1;
    c->nz = grid->nz ? grid->nz : 1;
    c->active = c->dnx + c->dny + c->dnz;
    c->ncells = static_cast<int64_t>(c->nx) * c->ny * c->nz;
    tile_shape(c, 0, &c->tile_y, &c->tile_z, &c->tile_x);
    // the step kernels address one slot plane of FP64 values with 31-bit byte offsets
    if (c->ncells >= (int64_t{1} << 28))
        return bail(MPF_ERR_UNSUPPORTED, "at most 2^28 - 1 cells per device (e.g. 645^3)");
    c->slots = grid->slots;
    c->fp32 = grid->precision == MPF_FP32;
    // deferred-cell grid: 16384 blocks from 4M cells on (256^3 FP64 step 3.98 ms at 4096 blocks,
    // 3.92 at 8192, 3.85 at 16384: the level-0 list then needs one pass), 4096 below
    c->deferred_grid = c->ncells >= (int64_t{1} << 22) ? 4 * mpf::kWideGrid : mpf::kWideGrid;
    c->nphases = params->nphases;
    const size_t np2 = static_cast<size_t>(c->nphases) * c->nphases;
    c->sigma.assign(params->sigma, params->sigma + np2);
    c->mu.assign(params->mu, params->mu + np2);
    c->eps.assign(np2, 0.0);
    if (params->epsilon) c->eps.assign(params->epsilon, params->epsilon + np2);
    c->emodel.assign(np2, MPF_ENERGY_ISO);
    if (params->energy_model) c->emodel.assign(params->energy_model, params->energy_model + np2);
    c->grad = false;
    for (size_t i = 0; i < np2; ++i) {
        if (c->emodel[i] != MPF_ENERGY_ISO && c->emodel[i] != MPF_ENERGY_CUBIC)
            return bail(MPF_ERR_UNSUPPORTED, "energy model must be ISO or CUBIC");
        if (c->emodel[i] == MPF_ENERGY_CUBIC) c->grad = true;
    }
    // InterfaceEnergyModel::ReadInputIso/ReadInputCubic MaxEnergy (InterfaceEnergyModel.h:78-94)
    c->maxe.resize(np2);
    for (size_t i = 0; i < np2; ++i)
        c->maxe[i] = c->emodel[i] == MPF_ENERGY_CUBIC ? c->sigma[i] * (1.0 + c->eps[i] * 1.41) : c->sigma[i];
    c->tj_model = params->tj_model;
    c->tj_factor = params->tj_factor;
    c->tj_energy = params->tj_energy;
    c->uniform = !c->grad && ((c->tj_model == MPF_TJ_MODEL && c->tj_factor == 0.0) || c->tj_model == MPF_TJ_NONE);
    for (size_t i = 0; i < np2; ++i)
        if (c->sigma[i] != c->sigma[0] || c->mu[i] != c->mu[0]) c->uniform = false;
    c->consider_nucleus = params->consider_nucleus_volume;
    c->nucleus_volume_factor = params->nucleus_volume_factor;
    c->regularization = params->regularization_factor;
    // InterfaceRegularization::Initialize / ReadInput (src/InterfaceRegularization.cpp:46-118)
    c->reg = params->regularization != 0;
    if (c->reg) {
        if (c->fp32) return bail(MPF_ERR_UNSUPPORTED, "interface regularization needs FP64 storage");
        c->curvature_factor = params->curvature_factor;
        c->reg_active = params->regularization_factor != 1.0 || params->curvature_factor != 1.0;
        c->reg_average = params->reg_average != 0;
        c->reg_range = params->reg_range > 0 ? params->reg_range : static_cast<int>((grid->iwidth + 1) / 2);
        c->reg_threshold = params->reg_threshold > 0.0 ? params->reg_threshold : DBL_EPSILON;
        c->uniform = false;
        if ((c->dnx && c->nx < c->reg_range) || (c->dny && c->ny < c->reg_range) || (c->dnz && c->nz < c->reg_range))
            return bail(MPF_ERR_UNSUPPORTED, "interface regularization needs every active axis >= Range cells");
        // CollectAverageSR / DistributeAverageSR loop bounds and distance weights (:198-213)
        const int xr = std::min(c->reg_range, c->nx - 1) * c->dnx;
        const int yr = std::min(c->reg_range, c->ny - 1) * c->dny;
        const int zr = std::min(c->reg_range, c->nz - 1) * c->dnz;
        for (int ii = -xr; ii <= xr; ii++)
            for (int jj = -yr; jj <= yr; jj++)
                for (int kk = -zr; kk <= zr; kk++) {
                    const double wd = c->reg_range - std::sqrt(static_cast<double>(ii * ii + jj * jj + kk * kk));
                    if (!(wd > 0.0)) continue;
                    c->reg_off.insert(c->reg_off.end(), {ii, jj, kk});
                    c->reg_off_w.push_back(wd);
                }
    }

    const double eta = grid->iwidth * grid->dx;
    c->prefactor = kPi * kPi / (eta * eta);
    c->dg_prefactor = kPi / eta;

    c->dg = params->dg_mode != nullptr && params->dg_value != nullptr;
    c->dg_limit = params->dg_limit;
    c->dgmode.assign(np2, 0);
    c->dgval.assign(np2, 0.0);
    c->dgallowed.assign(np2, 1.0);
    if (c->dg) {
        c->dgmode.assign(params->dg_mode, params->dg_mode + np2);
        c->dgval.assign(params->dg_value, params->dg_value + np2);
        for (int a = 0; a < c->nphases; ++a)
            for (int b = 0; b < c->nphases; ++b) {
                const size_t k = static_cast<size_t>(a) * c->nphases + b;
                if (a == b) c->dgmode[k] = 0;  // UserDrivingForce::ReadInput: same-phase pairs are NONE
                // DrivingForce::MergePhaseFieldIncrementsSR limiter amplitude (src/DrivingForce.cpp:548-558)
                double allowed = params->dg_limit * c->dg_prefactor * c->maxe[k];
                allowed *= c->regularization;
                c->dgallowed[k] = allowed;
            }
    }
    // DrivingForce::Initialize / ReadInput (src/DrivingForce.cpp:45-90, 126-152)
    c->dg_avg = c->dg && params->dg_average != 0;
    if (c->dg_avg) {
        c->avg_range = params->dg_range > 0 ? params->dg_range : static_cast<int>(grid->iwidth);
        const double thr = params->dg_threshold > 0.0 ? params->dg_threshold
                                                      : (grid->iwidth < 5.0 ? grid->iwidth / 15.0 : 1.0 / 3.0);
        c->avg_threshold = std::sqrt(thr * (1.0 - thr));
        c->avg_weights = params->dg_weights;
        if (c->avg_weights < MPF_DG_WEIGHTS_PHASEFIELDS || c->avg_weights > MPF_DG_WEIGHTS_COUNTER)
            return bail(MPF_ERR_ARG, "unknown driving-force weights mode");
        // the reference's periodic halo of Range cells is a wrapped index only on axes of >= Range cells
        if ((c->dnx && c->nx < c->avg_range) || (c->dny && c->ny < c->avg_range) || (c->dnz && c->nz < c->avg_range))
            return bail(MPF_ERR_UNSUPPORTED, "driving-force averaging needs every active axis >= Range cells");
        // CollectAverage / DistributeAverage loop bounds and distance weights (src/DrivingForce.cpp:295-324)
        const int xr = std::min(c->avg_range, c->nx - 1) * c->dnx;
        const int yr = std::min(c->avg_range, c->ny - 1) * c->dny;
        const int zr = std::min(c->avg_range, c->nz - 1) * c->dnz;
        for (int ii = -xr; ii <= xr; ii++)
            for (int jj = -yr; jj <= yr; jj++)
                for (int kk = -zr; kk <= zr; kk++) {
                    const double wd = c->avg_range - std::sqrt(static_cast<double>(ii * ii + jj * jj + kk * kk));
                    if (!(wd > 0.0)) continue;
                    c->avg_off.insert(c->avg_off.end(), {ii, jj, kk});
                    c->avg_off_w.push_back(wd);
                }
    }

    // Stencil weights by number of non-zero offset components: Laplacian / dx^2, gradient / dx
    // (LaplacianStencil.h:85-119, GradientStencil.h:58-92; 1D/2D use the middle slices).
    const double dx = grid->dx;
    double lt[4], gt[4];
    if (c->active == 1) {
        lt[0] = -2.0; lt[1] = 1.0; lt[2] = 0.0; lt[3] = 0.0;
        gt[0] = 0.0; gt[1] = 0.5; gt[2] = 0.0; gt[3] = 0.0;
    } else if (c->active == 2) {
        lt[0] = -10.0 / 3.0; lt[1] = 2.0 / 3.0; lt[2] = 1.0 / 6.0; lt[3] = 0.0;
        gt[0] = 0.0; gt[1] = 1.0 / 3.0; gt[2] = 0.25 / 3.0; gt[3] = 0.0;
    } else {
        lt[0] = -64.0 / 15.0; lt[1] = 7.0 / 15.0; lt[2] = 1.0 / 10.0; lt[3] = 1.0 / 30.0;
        gt[0] = 0.0; gt[1] = 1.0 / 4.64; gt[2] = 0.245 / 4.64; gt[3] = 0.085 / 4.64;
    }
    for (int i = 0; i < 4; ++i) {
        c->wl[i] = lt[i] / (dx * dx);
        c->wg[i] = gt[i] / dx;
    }

    const size_t N = static_cast<size_t>(c->ncells);
    mpf_status s = MPF_OK;
    if ((s = c->hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate"))) return bail(s, c->err.c_str());
    for (int b = 0; b < 2; ++b) {
        if ((s = c->hip(dalloc(&c->cnt[b], N), "hipMalloc counts"))) return bail(MPF_ERR_NOMEM, c->err.c_str());
        if ((s = c->hip(dalloc(&c->id[b], N * c->slots), "hipMalloc indices"))) return bail(MPF_ERR_NOMEM, c->err.c_str());
        if ((s = c->hip(hipMalloc(&c->val[b], N * c->slots * c->vsize()), "hipMalloc values"))) return bail(MPF_ERR_NOMEM, c->err.c_str());
    }
    if (c->hip(dalloc(&c->d_sigma, np2), "tables") || c->hip(dalloc(&c->d_mu, np2), "tables") ||
        c->hip(dalloc(&c->d_eps, np2), "tables") || c->hip(dalloc(&c->d_maxe, np2), "tables") ||
        c->hip(dalloc(&c->d_emodel, np2), "tables") || c->hip(dalloc(&c->d_dgmode, np2), "tables") ||
        c->hip(dalloc(&c->d_dgval, np2), "tables") || c->hip(dalloc(&c->d_dgallowed, np2), "tables") ||
        c->hip(dalloc(&c->d_dgmode_off, np2), "tables") ||
        c->hip(dalloc(&c->d_status, mpf::kStatusWords), "status") ||
        c->hip(dalloc(&c->d_first_overflow, 1), "status") || c->hip(dalloc(&c->d_warnings, 1), "status") ||
        c->hip(dalloc(&c->d_work_list, 4 * N), "work list") || c->hip(dalloc(&c->d_work_count, 4), "work list") ||
        c->hip(dalloc(&c->d_stats, 2), "statistics") || c->hip(dalloc(&c->d_obs, 4), "statistics"))
        return bail(MPF_ERR_NOMEM, c->err.c_str());
    if (c->reg_active && c->reg_average) {
        const size_t pairs = static_cast<size_t>(c->slots) * (c->slots - 1) / 2;
        if (c->hip(dalloc(&c->d_curv_raw, pairs * N), "regularization arrays") ||
            c->hip(dalloc(&c->d_curv_tmp, pairs * N), "regularization arrays") ||
            c->hip(dalloc(&c->d_curv, pairs * N), "regularization arrays") ||
            c->hip(dalloc(&c->d_reg_off, c->reg_off.size()), "regularization arrays") ||
            c->hip(dalloc(&c->d_reg_off_w, c->reg_off_w.size()), "regularization arrays"))
            return bail(MPF_ERR_NOMEM, c->err.c_str());
    }
    if (c->dg_avg) {
        const size_t pairs = static_cast<size_t>(c->slots) * (c->slots - 1) / 2;
        if (c->hip(dalloc(&c->d_avg_tmp, pairs * N), "averaging arrays") || c->hip(dalloc(&c->d_avg, pairs * N), "averaging arrays") ||
            c->hip(dalloc(&c->d_avg_off, c->avg_off.size()), "averaging arrays") ||
            c->hip(dalloc(&c->d_avg_off_w, c->avg_off_w.size()), "averaging arrays"))
            return bail(MPF_ERR_NOMEM, c->err.c_str());
    }
    hipError_t e = hipSuccess;
    h2d(e, c->d_sigma, c->sigma.data(), np2);
    h2d(e, c->d_mu, c->mu.data(), np2);
    h2d(e, c->d_eps, c->eps.data(), np2);
    h2d(e, c->d_maxe, c->maxe.data(), np2);
    h2d(e, c->d_emodel, c->emodel.data(), np2);
    h2d(e, c->d_dgmode, c->dgmode.data(), np2);
    const std::vector<int32_t> no_mode(np2, 0);
    h2d(e, c->d_dgmode_off, no_mode.data(), np2);
    h2d(e, c->d_dgval, c->dgval.data(), np2);
    h2d(e, c->d_dgallowed, c->dgallowed.data(), np2);
    if (c->reg_active && c->reg_average) {
        h2d(e, c->d_reg_off, c->reg_off.data(), c->reg_off.size());
        h2d(e, c->d_reg_off_w, c->reg_off_w.data(), c->reg_off_w.size());
    }
    if (c->dg_avg) {
        h2d(e, c->d_avg_off, c->avg_off.data(), c->avg_off.size());
        h2d(e, c->d_avg_off_w, c->avg_off_w.data(), c->avg_off_w.size());
    }
    const int32_t zeros[mpf::kStatusWords] = {0, 0, 0, 0};
    const unsigned long long none = ~0ull, zero = 0ull;
    h2d(e, c->d_status, zeros, mpf::kStatusWords);
    h2d(e, c->d_first_overflow, &none, 1);
    h2d(e, c->d_warnings, &zero, 1);
    const unsigned zero_u[4] = {0u, 0u, 0u, 0u};
    h2d(e, c->d_work_count, zero_u, 4);
    const unsigned long long zero_s[4] = {0ull, 0ull, 0ull, 0ull};
    h2d(e, c->d_stats, zero_s, 2);
    h2d(e, c->d_obs, zero_s, 4);
    if ((s = c->hip(e, "init copies"))) return bail(s, c->err.c_str());
    *out = c;
    return MPF_OK;
}

mpf_status mpf_set_grains(mpf_ctx* c, int32_t ngrains, const int32_t* phase, const double* rotation,
                          const uint8_t* solid)
{
    if (!c) return MPF_ERR_ARG;
    if (ngrains < 1 || ngrains > 65536) return c->fail(MPF_ERR_ARG, "ngrains must be in [1, 65536]");
    // every argument is checked and every device array allocated before the current grains are
    // replaced: a failed call leaves the context as it was
    const size_t G = static_cast<size_t>(ngrains);
    std::vector<int32_t> ph(G, 0);
    std::vector<double> rot(9 * G, 0.0), ref(G), zero(G, 0.0);
    std::vector<uint8_t> sol(G, 1), ex(G, 1);
    std::vector<int32_t> stage(G, 2);
    for (size_t g = 0; g < G; ++g) {
        if (phase) {
            if (phase[g] < 0 || phase[g] >= c->nphases) return c->fail(MPF_ERR_ARG, "grain phase out of range");
            ph[g] = phase[g];
        }
        for (int q = 0; q < 9; ++q) rot[9 * g + q] = rotation ? rotation[9 * g + q] : (q % 4 == 0 ? 1.0 : 0.0);
        if (solid) sol[g] = solid[g] ? 1 : 0;
    }
    // PhaseField::CalculateReferenceVolume(iWidth) (src/PhaseField.cpp:367-391)
    const double r = c->grid.iwidth;
    double rv = 0.0;
    if (c->active == 1) rv = r;
    else if (c->active == 2) rv = kPi * r * r;
    else rv = (4.0 / 3.0) * kPi * r * r * r;
    rv = rv * c->nucleus_volume_factor;
    std::fill(ref.begin(), ref.end(), rv);

    mpf_ctx n;  // the new arrays; moved into c only when complete
    auto drop = [&](mpf_status s) {
        free_grains(&n);
        return s;
    };
    if (c->hip(dalloc(&n.g_phase, G), "grains") || c->hip(dalloc(&n.g_vr, G), "grains") ||
        c->hip(dalloc(&n.g_rot, 9 * G), "grains") || c->hip(dalloc(&n.g_solid, G), "grains") ||
        c->hip(dalloc(&n.g_volume, G), "grains") || c->hip(dalloc(&n.g_maxvolume, G), "grains") ||
        c->hip(dalloc(&n.g_refvolume, G), "grains") || c->hip(dalloc(&n.g_exist, G), "grains") ||
        c->hip(dalloc(&n.g_stage, G), "grains") || c->hip(dalloc(&n.g_acc, G * mpf::kLimbs), "grains"))
        return drop(MPF_ERR_NOMEM);
    std::vector<unsigned long long> acc(G * mpf::kLimbs, 0ull);
    hipError_t e = hipSuccess;
    h2d(e, n.g_phase, ph.data(), G);
    h2d(e, n.g_vr, zero.data(), G);
    h2d(e, n.g_rot, rot.data(), 9 * G);
    h2d(e, n.g_solid, sol.data(), G);
    h2d(e, n.g_volume, zero.data(), G);
    h2d(e, n.g_maxvolume, zero.data(), G);
    h2d(e, n.g_refvolume, ref.data(), G);
    h2d(e, n.g_exist, ex.data(), G);
    h2d(e, n.g_stage, stage.data(), G);
    h2d(e, n.g_acc, acc.data(), G * mpf::kLimbs);
    if (mpf_status s = c->hip(e, "grain upload")) return drop(s);
    // work queued on the stream may still read the old arrays
    if (mpf_status s = c->hip(hipStreamSynchronize(c->stream), "stream synchronize")) return drop(s);
    free_grains(c);
    c->g_phase = n.g_phase; c->g_vr = n.g_vr; c->g_rot = n.g_rot; c->g_solid = n.g_solid;
    c->g_volume = n.g_volume; c->g_maxvolume = n.g_maxvolume; c->g_refvolume = n.g_refvolume;
    c->g_exist = n.g_exist; c->g_stage = n.g_stage; c->g_acc = n.g_acc;
    n.g_phase = nullptr; n.g_vr = nullptr; n.g_rot = nullptr; n.g_solid = nullptr;
    n.g_volume = nullptr; n.g_maxvolume = nullptr; n.g_refvolume = nullptr;
    n.g_exist = nullptr; n.g_stage = nullptr; n.g_acc = nullptr;
    // an uploaded state was checked against the previous grain count: with fewer grains its
    // indices may be out of range, so it has to be uploaded again
    if (ngrains < c->ngrains) c->have_state = false;
    c->prev_valid = false;
    c->ngrains = ngrains;
    c->have_grains = true;
    c->initialized = false;
    c->args_ok = false;
    return MPF_OK;
}

mpf_status mpf_upload(mpf_ctx* c, const int32_t* count, const int32_t* index, const double* value)
{
    if (!c || !count || !index || !value) return c ? c->fail(MPF_ERR_ARG, "null buffer") : MPF_ERR_ARG;
    if (!c->have_grains) return c->fail(MPF_ERR_STATE, "mpf_set_grains must precede mpf_upload");
    const size_t N = static_cast<size_t>(c->ncells);
    const int K = c->slots;
    std::vector<uint8_t> cnt(N);
    std::vector<uint16_t> ids(N * K, 0);
    std::vector<double> vd;
    std::vector<float> vf;
    if (c->fp32) vf.assign(N * K, 0.0f);
    else vd.assign(N * K, 0.0);
    size_t off = 0;
    for (size_t i = 0; i < N; ++i) {
        if (count[i] < 1 || count[i] > K)
            return c->fail(count[i] < 1 ? MPF_ERR_ARG : MPF_SLOT_OVERFLOW, "cell " + std::to_string(i) + " has " +
                           std::to_string(count[i]) + " entries (allowed 1.." + std::to_string(K) + ")");
        // FP32 states: implicit slot 15 = every value stored (FinalizeInitialization re-encodes)
        cnt[i] = static_cast<uint8_t>(c->fp32 ? (count[i] | (15 << 4)) : count[i]);
        for (int s = 0; s < count[i]; ++s, ++off) {
            if (index[off] < 0 || index[off] >= c->ngrains)
                return c->fail(MPF_ERR_ARG, "phase-field index out of range in cell " + std::to_string(i));
            ids[static_cast<size_t>(s) * N + i] = static_cast<uint16_t>(index[off]);
            if (c->fp32) vf[static_cast<size_t>(s) * N + i] = static_cast<float>(value[off]);
            else vd[static_cast<size_t>(s) * N + i] = value[off];
        }
    }
    const int b = c->cur;
    mpf_status s = c->hip(hipMemcpy(c->cnt[b], cnt.data(), N, hipMemcpyHostToDevice), "upload counts");
    if (!s) s = c->hip(hipMemcpy(c->id[b], ids.data(), N * K * sizeof(uint16_t), hipMemcpyHostToDevice), "upload indices");
    if (!s) s = c->hip(hipMemcpy(c->val[b], c->fp32 ? static_cast<const void*>(vf.data()) : static_cast<const void*>(vd.data()),
                                 N * K * c->vsize(), hipMemcpyHostToDevice), "upload values");
    if (s) return s;
    c->have_state = true;
    c->initialized = false;
    c->prev_valid = false;
    return MPF_OK;
}

mpf_status mpf_finalize_initialization(mpf_ctx* c)
{
    if (!c) return MPF_ERR_ARG;
    if (!c->have_state || !c->have_grains) return c->fail(MPF_ERR_STATE, "upload state and grains first");
    mpf::StepArgs a = step_args(c, 0.0, c->cur);
    mpf_status s = c->hip(static_cast<hipError_t>(mpf::launch_finalize_init(a, c->fp32, c->stream)), "finalize kernel");
    if (s) return s;
    s = sync_args(c, 1.0, c->stream);
    if (s) return s;
    s = c->hip(static_cast<hipError_t>(mpf::launch_grains(c->d_gargs, c->d_gfirst, 1, grain_blocks(c), c->stream)),
               "grain kernel");
    if (s) return s;
    c->cur = 1 - c->cur;
    c->prev_valid = false;
    s = read_status(c);
    if (s) return s;
    c->initialized = true;
    return MPF_OK;
}

mpf_status mpf_step_async(mpf_ctx* c, double dt, int32_t nsteps)
{
    if (!c) return MPF_ERR_ARG;
    if (!c->initialized) return c->fail(MPF_ERR_STATE, "mpf_finalize_initialization must precede mpf_step");
    if (nsteps < 0 || !(dt > 0.0)) return c->fail(MPF_ERR_ARG, "dt must be > 0 and nsteps >= 0");
    // Per-step events (timing on): before the register kernel, after it, after the deferred-cell
    // and grain kernels.
    const bool timing = c->timing && nsteps > 0;
    if (timing && c->events.size() < static_cast<size_t>(3 * nsteps)) {
        for (auto e : c->events) (void)hipEventDestroy(e);
        c->events.assign(3 * static_cast<size_t>(nsteps), nullptr);
        for (auto& e : c->events)
            if (mpf_status st = c->hip(hipEventCreate(&e), "hipEventCreate")) return st;
    }
    auto record = [&](int i) -> mpf_status {
        return timing ? c->hip(hipEventRecord(c->events[i], c->stream), "hipEventRecord") : MPF_OK;
    };
    if (nsteps > 0)
        if (mpf_status st = sync_args(c, dt, c->stream)) return st;
    for (int s = 0; s < nsteps; ++s) {
        mpf_status st = record(3 * s);
        if (!st) st = launch_step_part(c, c->stream);
        if (!st) st = record(3 * s + 1);
        if (!st) st = launch_rest_part(c, c->stream);
        if (!st) st = record(3 * s + 2);
        if (st) return st;
        c->cur = 1 - c->cur;
        c->stepped = true;
        c->prev_valid = true;
    }
    if (timing) {
        if (mpf_status st = c->hip(hipEventSynchronize(c->events[3 * (nsteps - 1) + 2]), "hipEventSynchronize")) return st;
        double ks = 0.0, kg = 0.0;
        for (int s = 0; s < nsteps; ++s) {
            float a = 0.f, b = 0.f;
            if (mpf_status st = c->hip(hipEventElapsedTime(&a, c->events[3 * s], c->events[3 * s + 1]), "hipEventElapsedTime")) return st;
            if (mpf_status st = c->hip(hipEventElapsedTime(&b, c->events[3 * s + 1], c->events[3 * s + 2]), "hipEventElapsedTime")) return st;
            ks += a; kg += b;
        }
        c->last_step_ms = ks / nsteps;
        c->last_grain_ms = kg / nsteps;
    }
    return MPF_OK;
}

mpf_status mpf_step(mpf_ctx* c, double dt, int32_t nsteps)
{
    mpf_status s = mpf_step_async(c, dt, nsteps);
    if (s) return s;
    return read_status(c);
}

mpf_status mpf_synchronize(mpf_ctx* c)
{
    if (!c) return MPF_ERR_ARG;
    return read_status(c);
}

mpf_status mpf_total_entries(mpf_ctx* c, int64_t* total)
{
    if (!c || !total) return MPF_ERR_ARG;
    if (!c->have_state) return c->fail(MPF_ERR_STATE, "no state");
    std::vector<uint8_t> cnt(static_cast<size_t>(c->ncells));
    mpf_status s = c->hip(hipMemcpyAsync(cnt.data(), c->cnt[c->cur], cnt.size(), hipMemcpyDeviceToHost, c->stream), "download counts");
    if (!s) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s) return s;
    int64_t t = 0;
    for (uint8_t v : cnt) t += c->fp32 ? (v & 15) : v;
    *total = t;
    return MPF_OK;
}

mpf_status mpf_download(mpf_ctx* c, int32_t* count, int32_t* index, double* value, int8_t* flag)
{
    if (!c) return MPF_ERR_ARG;
    if (!c->have_state) return c->fail(MPF_ERR_STATE, "no state");
    const size_t N = static_cast<size_t>(c->ncells);
    const int K = c->slots;
    std::vector<uint8_t> cnt(N);
    std::vector<uint16_t> ids(N * K);
    std::vector<double> vd;
    std::vector<float> vf;
    if (c->fp32) vf.resize(N * K); else vd.resize(N * K);
    const int b = c->cur;
    mpf_status s = c->hip(hipMemcpyAsync(cnt.data(), c->cnt[b], N, hipMemcpyDeviceToHost, c->stream), "download");
    if (!s) s = c->hip(hipMemcpyAsync(ids.data(), c->id[b], N * K * sizeof(uint16_t), hipMemcpyDeviceToHost, c->stream), "download");
    if (!s) s = c->hip(hipMemcpyAsync(c->fp32 ? static_cast<void*>(vf.data()) : static_cast<void*>(vd.data()), c->val[b],
                                      N * K * c->vsize(), hipMemcpyDeviceToHost, c->stream), "download");
    if (!s) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s) return s;
    // decode FP32 states: count | (implicit slot << 4); the implicit value is 1 - sum(others)
    std::vector<uint8_t> n(N);
    for (size_t i = 0; i < N; ++i) n[i] = c->fp32 ? (cnt[i] & 15) : cnt[i];
    size_t off = 0;
    for (size_t i = 0; i < N; ++i) {
        if (count) count[i] = n[i];
        const int imp = c->fp32 && (cnt[i] >> 4) != 15 ? (cnt[i] >> 4) : -1;
        double others = 0.0;
        for (int q = 0; q < n[i]; ++q)
            if (q != imp) others += c->fp32 ? static_cast<double>(vf[static_cast<size_t>(q) * N + i]) : vd[static_cast<size_t>(q) * N + i];
        for (int q = 0; q < n[i]; ++q, ++off) {
            if (index) index[off] = ids[static_cast<size_t>(q) * N + i];
            if (!value) continue;
            if (q == imp) value[off] = 1.0 - others;
            else value[off] = c->fp32 ? static_cast<double>(vf[static_cast<size_t>(q) * N + i]) : vd[static_cast<size_t>(q) * N + i];
        }
    }
    if (flag) {
        for (size_t i = 0; i < N; ++i) {
            if (n[i] > 1) { flag[i] = 2; continue; }
            const int x = static_cast<int>(i / (static_cast<size_t>(c->ny) * c->nz));
            const int y = static_cast<int>((i / c->nz) % c->ny);
            const int z = static_cast<int>(i % c->nz);
            int8_t f = 0;
            for (int ii = -c->dnx; ii <= c->dnx && !f; ++ii)
                for (int jj = -c->dny; jj <= c->dny && !f; ++jj)
                    for (int kk = -c->dnz; kk <= c->dnz && !f; ++kk) {
                        const int xx = (x + ii + c->nx) % c->nx, yy = (y + jj + c->ny) % c->ny, zz = (z + kk + c->nz) % c->nz;
                        const size_t nb = (static_cast<size_t>(xx) * c->ny + yy) * c->nz + zz;
                        if (n[nb] > 1) f = 1;
                    }
            flag[i] = f;
        }
    }
    return MPF_OK;
}

mpf_status mpf_get_grains(mpf_ctx* c, double* volume, double* volume_ratio)
{
    if (!c) return MPF_ERR_ARG;
    if (!c->have_grains) return c->fail(MPF_ERR_STATE, "no grains");
    const size_t G = static_cast<size_t>(c->ngrains);
    mpf_status s = MPF_OK;
    if (volume) s = c->hip(hipMemcpyAsync(volume, c->g_volume, G * sizeof(double), hipMemcpyDeviceToHost, c->stream), "grains");
    if (!s && volume_ratio) s = c->hip(hipMemcpyAsync(volume_ratio, c->g_vr, G * sizeof(double), hipMemcpyDeviceToHost, c->stream), "grains");
    if (!s) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    return s;
}

mpf_status mpf_get_grain_records(mpf_ctx* c, double* maxvolume, double* refvolume, uint8_t* exist, int32_t* stage)
{
    if (!c) return MPF_ERR_ARG;
    if (!c->have_grains) return c->fail(MPF_ERR_STATE, "no grains");
    const size_t G = static_cast<size_t>(c->ngrains);
    mpf_status s = MPF_OK;
    auto get = [&](void* dst, const void* src, size_t bytes) {
        if (dst && !s) s = c->hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "grains");
    };
    get(maxvolume, c->g_maxvolume, G * sizeof(double));
    get(refvolume, c->g_refvolume, G * sizeof(double));
    get(exist, c->g_exist, G);
    get(stage, c->g_stage, G * sizeof(int32_t));
    if (!s) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    return s;
}

mpf_status mpf_max_time_step(mpf_ctx* c, double* dt_max)
{
    if (!c || !dt_max) return MPF_ERR_ARG;
    // maxEnergies / maxMobilities as the last step's IP.Set left them; 0 before any step
    double me = 0.0, mm = 0.0;
    if (c->stepped) {
        unsigned long long w[2] = {0ull, 0ull};
        mpf_status s = c->hip(hipMemcpyAsync(w, c->d_obs + 2, sizeof(w), hipMemcpyDeviceToHost, c->stream), "maxima");
        if (!s) s = c->hip(hipStreamSynchronize(c->stream), "sync");
        if (s) return s;
        std::memcpy(&me, &w[0], sizeof(double));
        std::memcpy(&mm, &w[1], sizeof(double));
    }
    double dt = DBL_MAX;
    if (me > DBL_EPSILON && mm > 0.0) {
        const double f = c->active == 1 ? 0.50 : (c->active == 2 ? 0.25 : 0.16);
        dt = f * c->grid.dx * c->grid.dx / (mm * me * c->regularization);
    }
    *dt_max = dt;
    return MPF_OK;
}

mpf_status mpf_interface_energy(mpf_ctx* c, int32_t ip_state, double* sum)
{
    if (!c || !sum) return MPF_ERR_ARG;
    if (!c->initialized) return c->fail(MPF_ERR_STATE, "mpf_finalize_initialization first");
    if (ip_state < MPF_IP_NONE || ip_state > MPF_IP_PREVIOUS) return c->fail(MPF_ERR_ARG, "unknown ip_state");
    if (ip_state == MPF_IP_PREVIOUS && !c->prev_valid)
        return c->fail(MPF_ERR_STATE, "the state the last IP.Set saw is held only after a step (none since the last "
                                      "upload, restore or finalize)");
    const size_t nb = static_cast<size_t>(mpf::energy_blocks(c->ncells));
    if (c->energy_cap < nb + 1) {
        dfree(c->d_energy);
        c->energy_cap = 0;
        if (c->hip(dalloc(&c->d_energy, nb + 1), "energy partial sums")) return MPF_ERR_NOMEM;
        c->energy_cap = nb + 1;
    }
    mpf::EnergyArgs a{};
    a.cur = step_args(c, 1.0, c->cur);
    if (ip_state != MPF_IP_NONE) {
        const int ip = ip_state == MPF_IP_CURRENT ? c->cur : 1 - c->cur;
        a.ip_cnt = c->cnt[ip];
        a.ip_id = c->id[ip];
        a.ip_val = c->val[ip];
    }
    const double eta = c->grid.iwidth * c->grid.dx;
    a.eta = eta;
    a.pref = eta * eta / (kPi * kPi);
    a.partial = c->d_energy;
    a.status = c->d_status + 3;
    const int32_t zero = 0;
    int32_t flag = 0;
    mpf_status st = c->hip(hipMemcpyAsync(a.status, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream), "status reset");
    if (!st) st = c->hip(static_cast<hipError_t>(mpf::launch_energy(a, c->fp32, c->d_energy + nb, c->stream)), "energy kernels");
    if (!st) st = c->hip(hipMemcpyAsync(sum, c->d_energy + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream), "energy readback");
    if (!st) st = c->hip(hipMemcpyAsync(&flag, a.status, sizeof(flag), hipMemcpyDeviceToHost, c->stream), "status readback");
    if (!st) st = c->hip(hipStreamSynchronize(c->stream), "stream synchronize");
    if (!st && flag) st = c->fail(MPF_SLOT_OVERFLOW, "a cell's merged list exceeds 16 entries");
    return st;
}

mpf_status mpf_phase_fractions(mpf_ctx* c, double* dst, int32_t dst_on_host)
{
    if (!c || !dst) return MPF_ERR_ARG;
    if (!c->initialized) return c->fail(MPF_ERR_STATE, "mpf_finalize_initialization first");
    const size_t n = static_cast<size_t>(c->nphases) * static_cast<size_t>(c->ncells);
    double* out = dst;
    if (dst_on_host) {
        if (c->fractions_cap < n) {
            dfree(c->d_fractions);
            c->fractions_cap = 0;
            if (c->hip(dalloc(&c->d_fractions, n), "phase fractions")) return MPF_ERR_NOMEM;
            c->fractions_cap = n;
        }
        out = c->d_fractions;
    }
    const mpf::StepArgs a = step_args(c, 1.0, c->cur);
    mpf_status st = c->hip(static_cast<hipError_t>(mpf::launch_fractions(a, c->fp32, out, c->stream)), "fractions kernel");
    if (!st && dst_on_host)
        st = c->hip(hipMemcpyAsync(dst, out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "fractions download");
    if (!st) st = c->hip(hipStreamSynchronize(c->stream), "stream synchronize");
    return st;
}

mpf_status mpf_set_step_terms(mpf_ctx* c, int32_t driving_force, int32_t average)
{
    if (!c) return MPF_ERR_ARG;
    if (driving_force && !c->dg) return c->fail(MPF_ERR_STATE, "the context was created without a driving force");
    c->dg_on = driving_force != 0;
    c->avg_on = average != 0;
    return MPF_OK;
}

mpf_status mpf_limiter_warnings(mpf_ctx* c, int64_t* count)
{
    if (!c || !count) return MPF_ERR_ARG;
    unsigned long long w = 0;
    mpf_status s = c->hip(hipMemcpyAsync(&w, c->d_warnings, sizeof(w), hipMemcpyDeviceToHost, c->stream), "warnings");
    if (!s) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    *count = static_cast<int64_t>(w);
    return s;
}

mpf_status mpf_overflow_cell(mpf_ctx* c, int64_t* cell)
{
    if (!c || !cell) return MPF_ERR_ARG;
    unsigned long long w = 0;
    mpf_status s = c->hip(hipMemcpyAsync(&w, c->d_first_overflow, sizeof(w), hipMemcpyDeviceToHost, c->stream), "overflow");
    if (!s) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    *cell = w == ~0ull ? -1 : static_cast<int64_t>(w);
    return s;
}

mpf_status mpf_device_state(mpf_ctx* c, void** count, void** index, void** value, void** stream)
{
    if (!c) return MPF_ERR_ARG;
    if (count) *count = c->cnt[c->cur];
    if (index) *index = c->id[c->cur];
    if (value) *value = c->val[c->cur];
    if (stream) *stream = c->stream;
    return MPF_OK;
}

mpf_status mpf_last_timing(mpf_ctx* c, double* step_kernel_ms, double* grain_kernel_ms)
{
    if (!c) return MPF_ERR_ARG;
    if (step_kernel_ms) *step_kernel_ms = c->last_step_ms;
    if (grain_kernel_ms) *grain_kernel_ms = c->last_grain_ms;
    return MPF_OK;
}

mpf_status mpf_enable_timing(mpf_ctx* c, int32_t on)
{
    if (!c) return MPF_ERR_ARG;
    c->timing = on != 0;
    return MPF_OK;
}

mpf_status mpf_step_stats(mpf_ctx* c, int64_t* deferred_mid, int64_t* deferred_wide)
{
    if (!c) return MPF_ERR_ARG;
    unsigned long long s[2] = {0ull, 0ull};
    mpf_status st = c->hip(hipMemcpyAsync(s, c->d_stats, sizeof(s), hipMemcpyDeviceToHost, c->stream), "statistics");
    if (!st) st = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (deferred_mid) *deferred_mid = static_cast<int64_t>(s[0]);
    if (deferred_wide) *deferred_wide = static_cast<int64_t>(s[1]);
    return st;
}

mpf_status mpf_copy_bandwidth(mpf_ctx* c, int64_t bytes, int32_t repeats, double* ms)
{
    if (!c || !ms) return MPF_ERR_ARG;
    if (bytes <= 0 || bytes % 16 || repeats < 1) return c->fail(MPF_ERR_ARG, "bytes must be a positive multiple of 16, repeats >= 1");
    void* src = nullptr;
    void* dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    mpf_status s = c->hip(hipMalloc(&src, static_cast<size_t>(bytes)), "hipMalloc copy source");
    if (!s) s = c->hip(hipMalloc(&dst, static_cast<size_t>(bytes)), "hipMalloc copy target");
    if (!s) s = c->hip(hipMemsetAsync(src, 1, static_cast<size_t>(bytes), c->stream), "hipMemsetAsync");
    if (!s) s = c->hip(hipEventCreate(&e0), "hipEventCreate");
    if (!s) s = c->hip(hipEventCreate(&e1), "hipEventCreate");
    if (!s) s = c->hip(static_cast<hipError_t>(mpf::launch_copy(src, dst, bytes, c->stream)), "copy kernel");  // warm-up
    if (!s) s = c->hip(hipEventRecord(e0, c->stream), "hipEventRecord");
    for (int r = 0; r < repeats && !s; ++r)
        s = c->hip(static_cast<hipError_t>(mpf::launch_copy(src, dst, bytes, c->stream)), "copy kernel");
    if (!s) s = c->hip(hipEventRecord(e1, c->stream), "hipEventRecord");
    if (!s) s = c->hip(hipEventSynchronize(e1), "hipEventSynchronize");
    float t = 0.f;
    if (!s) s = c->hip(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
    if (!s) *ms = static_cast<double>(t) / repeats;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(src);
    dfree(dst);
    return s;
}

const char* mpf_error_message(mpf_ctx* c)
{
    return c ? c->err.c_str() : "null context";
}

void mpf_destroy(mpf_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto e : c->events) (void)hipEventDestroy(e);
    for (int b = 0; b < 2; ++b) { dfree(c->cnt[b]); dfree(c->id[b]); dfree(c->val[b]); }
    free_grains(c);
    dfree(c->a_lam_own); dfree(c->a_cnt); dfree(c->a_id); dfree(c->a_lap); dfree(c->a_grad); dfree(c->a_partial); dfree(c->a_defer);
    dfree(c->a_list); dfree(c->a_list_count); dfree(c->a_cellg);
    dfree(c->d_args); dfree(c->d_first); dfree(c->d_gargs); dfree(c->d_gfirst);
    dfree(c->d_sigma); dfree(c->d_mu); dfree(c->d_eps); dfree(c->d_maxe);
    dfree(c->d_emodel); dfree(c->d_dgmode); dfree(c->d_dgval); dfree(c->d_dgallowed); dfree(c->d_dgmode_off);
    dfree(c->d_energy); dfree(c->d_fractions);
    dfree(c->d_status); dfree(c->d_first_overflow); dfree(c->d_warnings);
    dfree(c->d_work_list); dfree(c->d_work_count); dfree(c->d_stats); dfree(c->d_obs);
    dfree(c->d_avg_tmp); dfree(c->d_avg); dfree(c->d_avg_off); dfree(c->d_avg_off_w);
    dfree(c->d_curv_raw); dfree(c->d_curv_tmp); dfree(c->d_curv); dfree(c->d_reg_off); dfree(c->d_reg_off_w);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// ------------------------------------------------------------------ parameters, checkpoints, adjoint

mpf_status mpf_set_pair_params(mpf_ctx* c, const double* sigma, const double* mu, const double* dg_value)
{
    if (!c) return MPF_ERR_ARG;
    const size_t np2 = static_cast<size_t>(c->nphases) * c->nphases;
    if (dg_value && !c->dg) return c->fail(MPF_ERR_ARG, "the context was created without a driving force");
    if (sigma) c->sigma.assign(sigma, sigma + np2);
    if (mu) c->mu.assign(mu, mu + np2);
    if (dg_value) c->dgval.assign(dg_value, dg_value + np2);
    // derived tables as in mpf_create: MaxEnergy, the tanh amplitude, the equal-energy flag
    for (size_t i = 0; i < np2; ++i)
        c->maxe[i] = c->emodel[i] == MPF_ENERGY_CUBIC ? c->sigma[i] * (1.0 + c->eps[i] * 1.41) : c->sigma[i];
    c->uniform = !c->grad && !c->reg && ((c->tj_model == MPF_TJ_MODEL && c->tj_factor == 0.0) || c->tj_model == MPF_TJ_NONE);
    for (size_t i = 0; i < np2; ++i)
        if (c->sigma[i] != c->sigma[0] || c->mu[i] != c->mu[0]) c->uniform = false;
    if (c->dg)
        for (size_t k = 0; k < np2; ++k) {
            double allowed = c->dg_limit * c->dg_prefactor * c->maxe[k];
            allowed *= c->regularization;
            c->dgallowed[k] = allowed;
        }
    hipError_t e = hipStreamSynchronize(c->stream);
    h2d(e, c->d_sigma, c->sigma.data(), np2);
    h2d(e, c->d_mu, c->mu.data(), np2);
    h2d(e, c->d_maxe, c->maxe.data(), np2);
    h2d(e, c->d_dgval, c->dgval.data(), np2);
    h2d(e, c->d_dgallowed, c->dgallowed.data(), np2);
    return c->hip(e, "pair tables");
}

namespace {

// Byte offsets of the parts of a saved state: count, index, value planes, then the grain arrays
// (volume, MAXVolume, VolumeRatio, Exist, Stage), each part 256-byte aligned.
struct StateLayout {
    size_t cnt, id, val, volume, maxvolume, vr, exist, stage, total;
};

StateLayout state_layout(const mpf_ctx* c)
{
    auto up = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
    const size_t N = static_cast<size_t>(c->ncells), S = static_cast<size_t>(c->slots), G = static_cast<size_t>(c->ngrains);
    StateLayout L{};
    L.cnt = 0;
    L.id = up(L.cnt + N);
    L.val = up(L.id + S * N * sizeof(uint16_t));
    L.volume = up(L.val + S * N * c->vsize());
    L.maxvolume = up(L.volume + G * sizeof(double));
    L.vr = up(L.maxvolume + G * sizeof(double));
    L.exist = up(L.vr + G * sizeof(double));
    L.stage = up(L.exist + G);
    L.total = up(L.stage + G * sizeof(int32_t));
    return L;
}

}  // namespace

mpf_status mpf_state_bytes(mpf_ctx* c, int64_t* bytes)
{
    if (!c || !bytes) return MPF_ERR_ARG;
    if (!c->have_grains) return c->fail(MPF_ERR_STATE, "mpf_set_grains first");
    *bytes = static_cast<int64_t>(state_layout(c).total);
    return MPF_OK;
}

mpf_status mpf_save_state(mpf_ctx* c, void* dst)
{
    if (!c || !dst) return MPF_ERR_ARG;
    if (!c->initialized) return c->fail(MPF_ERR_STATE, "mpf_finalize_initialization first");
    const StateLayout L = state_layout(c);
    const size_t N = static_cast<size_t>(c->ncells), S = static_cast<size_t>(c->slots), G = static_cast<size_t>(c->ngrains);
    char* d = static_cast<char*>(dst);
    hipError_t e = hipSuccess;
    auto cp = [&](size_t off, const void* src, size_t n) {
        if (e == hipSuccess) e = hipMemcpyAsync(d + off, src, n, hipMemcpyDeviceToDevice, c->stream);
    };
    cp(L.cnt, c->cnt[c->cur], N);
    cp(L.id, c->id[c->cur], S * N * sizeof(uint16_t));
    cp(L.val, c->val[c->cur], S * N * c->vsize());
    cp(L.volume, c->g_volume, G * sizeof(double));
    cp(L.maxvolume, c->g_maxvolume, G * sizeof(double));
    cp(L.vr, c->g_vr, G * sizeof(double));
    cp(L.exist, c->g_exist, G);
    cp(L.stage, c->g_stage, G * sizeof(int32_t));
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return c->hip(e, "state save");
}

mpf_status mpf_restore_state(mpf_ctx* c, const void* src)
{
    if (!c || !src) return MPF_ERR_ARG;
    if (!c->initialized) return c->fail(MPF_ERR_STATE, "mpf_finalize_initialization first");
    const StateLayout L = state_layout(c);
    const size_t N = static_cast<size_t>(c->ncells), S = static_cast<size_t>(c->slots), G = static_cast<size_t>(c->ngrains);
    const char* s = static_cast<const char*>(src);
    hipError_t e = hipSuccess;
    auto cp = [&](void* dstp, size_t off, size_t n) {
        if (e == hipSuccess) e = hipMemcpyAsync(dstp, s + off, n, hipMemcpyDeviceToDevice, c->stream);
    };
    c->prev_valid = false;
    cp(c->cnt[c->cur], L.cnt, N);
    cp(c->id[c->cur], L.id, S * N * sizeof(uint16_t));
    cp(c->val[c->cur], L.val, S * N * c->vsize());
    cp(c->g_volume, L.volume, G * sizeof(double));
    cp(c->g_maxvolume, L.maxvolume, G * sizeof(double));
    cp(c->g_vr, L.vr, G * sizeof(double));
    cp(c->g_exist, L.exist, G);
    cp(c->g_stage, L.stage, G * sizeof(int32_t));
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return c->hip(e, "state restore");
}

mpf_status mpf_adjoint_step(mpf_ctx* c, double dt, const double* lam_next, double* lam_prev, double* grad,
                            const void* next_count, const void* next_index)
{
    if (!c || !lam_next || !lam_prev || !grad) return MPF_ERR_ARG;
    if (!c->initialized) return c->fail(MPF_ERR_STATE, "mpf_finalize_initialization first");
    if (!(dt > 0.0)) return c->fail(MPF_ERR_ARG, "dt must be > 0");
    if ((next_count == nullptr) != (next_index == nullptr)) return c->fail(MPF_ERR_ARG, "give both next_count and next_index");
    if (c->fp32 && !next_count) return c->fail(MPF_ERR_ARG, "FP32 states need the next state's lists (next_count, next_index)");
    if (avg_active(c)) return c->fail(MPF_ERR_UNSUPPORTED, "the adjoint takes the driving force without averaging");
    if (c->reg) return c->fail(MPF_ERR_UNSUPPORTED, "the adjoint does not differentiate interface regularization");
    if (c->nphases > mpf::kAdjPhases) return c->fail(MPF_ERR_UNSUPPORTED, "the adjoint handles at most 4 phases");
    const size_t N = static_cast<size_t>(c->ncells);
    constexpr size_t KA = mpf::kAdjSlots;
    if (!c->a_cnt) {
        const size_t G = 3 * static_cast<size_t>(c->nphases) * c->nphases;
        if (c->hip(dalloc(&c->a_lam_own, KA * N), "adjoint arrays") || c->hip(dalloc(&c->a_cnt, N), "adjoint arrays") ||
            c->hip(dalloc(&c->a_id, KA * N), "adjoint arrays") || c->hip(dalloc(&c->a_lap, KA * N), "adjoint arrays") ||
            c->hip(dalloc(&c->a_defer, N), "adjoint arrays") || c->hip(dalloc(&c->a_list, 2 * N), "adjoint arrays") ||
            c->hip(dalloc(&c->a_list_count, 2), "adjoint arrays") || c->hip(dalloc(&c->a_cellg, N * G), "adjoint arrays") ||
            c->hip(dalloc(&c->a_partial, static_cast<size_t>(mpf::adjoint_blocks(c->ncells)) * G), "adjoint arrays"))
            return MPF_ERR_NOMEM;
        if (c->grad && c->hip(dalloc(&c->a_grad, 3 * KA * N), "adjoint arrays")) return MPF_ERR_NOMEM;
    }
    mpf::AdjointArgs a{};
    a.s = step_args(c, dt, c->cur);
    a.lam_next = lam_next;
    a.next_cnt = static_cast<const uint8_t*>(next_count);
    a.next_id = static_cast<const uint16_t*>(next_index);
    a.lam_prev = lam_prev;
    a.lam_own = c->a_lam_own;
    a.adj_cnt = c->a_cnt;
    a.defer = c->a_defer;
    a.list = c->a_list;
    a.list_count = c->a_list_count;
    a.list_long = c->a_list + N;
    a.long_count = c->a_list_count + 1;
    a.cellg = c->a_cellg;
    a.adj_id = c->a_id;
    a.adj_lap = c->a_lap;
    a.adj_grad = c->a_grad;
    a.partial = c->a_partial;
    a.grad = grad;
    a.dg_limit = c->dg_limit;
    a.regularization = c->regularization;
    a.grad_energy = c->grad ? 1 : 0;
    a.driving_force = dg_active(c) ? 1 : 0;
    const int32_t zeros[mpf::kStatusWords] = {0, 0, 0, 0};
    mpf_status st = c->hip(hipMemcpyAsync(c->d_status, zeros, sizeof(zeros), hipMemcpyHostToDevice, c->stream), "status reset");
    if (!st) st = c->hip(static_cast<hipError_t>(mpf::launch_adjoint(a, c->fp32, c->stream)), "adjoint kernels");
    int32_t words[mpf::kStatusWords];
    if (!st) st = c->hip(hipMemcpyAsync(words, c->d_status, sizeof(words), hipMemcpyDeviceToHost, c->stream), "status readback");
    if (!st) st = c->hip(hipStreamSynchronize(c->stream), "stream synchronize");
    if (!st && words[3]) st = c->fail(MPF_ERR_UNSUPPORTED, "a cell's merged list exceeds the adjoint's 8 entries");
    return st;
}

// ------------------------------------------------------------------ launch-shape autotuner

mpf_status mpf_autotune(mpf_ctx* c, double dt, int32_t steps, int32_t* xcd_order, int32_t* deferred_grid)
{
    if (!c || steps < 1 || !(dt > 0.0)) return MPF_ERR_ARG;
    if (!c->initialized) return c->fail(MPF_ERR_STATE, "mpf_finalize_initialization first");
    // time each candidate on the context's own state, restoring it before every trial
    int64_t nbytes = 0;
    if (mpf_status st = mpf_state_bytes(c, &nbytes)) return st;
    void* saved = nullptr;
    if (c->hip(hipMalloc(&saved, static_cast<size_t>(nbytes)), "autotune checkpoint")) return MPF_ERR_NOMEM;
    const int cur0 = c->cur;
    const int x0 = c->xcd_order, g0 = c->deferred_grid;
    // the trials add their steps to the statistics and counters; those are put back afterwards
    int32_t status0[mpf::kStatusWords];
    unsigned long long first0 = 0, warn0 = 0, stats0[2] = {0, 0}, obs0[4] = {0, 0, 0, 0};
    const bool stepped0 = c->stepped;
    mpf_status st = c->hip(hipMemcpyAsync(status0, c->d_status, sizeof(status0), hipMemcpyDeviceToHost, c->stream), "autotune snapshot");
    if (!st) st = c->hip(hipMemcpyAsync(&first0, c->d_first_overflow, sizeof(first0), hipMemcpyDeviceToHost, c->stream), "autotune snapshot");
    if (!st) st = c->hip(hipMemcpyAsync(&warn0, c->d_warnings, sizeof(warn0), hipMemcpyDeviceToHost, c->stream), "autotune snapshot");
    if (!st) st = c->hip(hipMemcpyAsync(stats0, c->d_stats, sizeof(stats0), hipMemcpyDeviceToHost, c->stream), "autotune snapshot");
    if (!st) st = c->hip(hipMemcpyAsync(obs0, c->d_obs, sizeof(obs0), hipMemcpyDeviceToHost, c->stream), "autotune snapshot");
    if (!st) st = c->hip(hipStreamSynchronize(c->stream), "autotune snapshot");
    if (!st) st = mpf_save_state(c, saved);
    const bool have_saved = st == MPF_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!st) st = c->hip(hipEventCreate(&e0), "hipEventCreate");
    if (!st) st = c->hip(hipEventCreate(&e1), "hipEventCreate");
    // one trial: the saved state, a warm-up step, then `steps` timed steps
    auto trial = [&](float* ms) -> mpf_status {
        c->cur = cur0;
        mpf_status t = mpf_restore_state(c, saved);
        if (!t) t = enqueue_step(c, dt, c->stream);
        if (!t) t = c->hip(hipEventRecord(e0, c->stream), "hipEventRecord");
        for (int k = 0; k < steps && !t; ++k) t = enqueue_step(c, dt, c->stream);
        if (!t) t = c->hip(hipEventRecord(e1, c->stream), "hipEventRecord");
        if (!t) t = c->hip(hipEventSynchronize(e1), "hipEventSynchronize");
        if (!t) t = c->hip(hipEventElapsedTime(ms, e0, e1), "hipEventElapsedTime");
        if (!t) t = read_status(c);
        return t;
    };
    // 1. step-kernel tile: plane tiles of 64-256 cells x 1-16 planes per block (or the default
    //    count), and the per-cell kernel
    const int tx0 = c->tile_x, ty0 = c->tile_y, tz0 = c->tile_z;
    float best = 1e30f;
    int best_tx = tx0, best_ty = ty0, best_tz = tz0;
    if (c->active >= 2 && !st) {
        std::vector<std::array<int, 2>> planes;
        if (c->dny && c->dnz) planes = {{4, 64}, {8, 32}, {16, 16}, {4, 32}};
        else if (c->dny) planes = {{64, 1}, {128, 1}, {256, 1}};
        else planes = {{1, 64}, {1, 128}, {1, 256}};
        std::vector<std::array<int, 3>> cands = {{0, 0, 0}};  // tile_x 0: the per-cell kernel
        std::vector<std::array<int, 2>> seen;
        for (auto pl : planes) {
            // a tile wider than the grid is the grid's width
            const int ty = std::min(pl[0], c->ny), tz = std::min(pl[1], c->nz);
            if ((ty + 2 * c->dny) * (tz + 2 * c->dnz) > mpf::kTileHalo) continue;
            if (std::find(seen.begin(), seen.end(), std::array<int, 2>{ty, tz}) != seen.end()) continue;
            seen.push_back({ty, tz});
            for (int xs : {0, 1, 4, 16}) {
                int dy = 0, dz = 0, tx = 0;
                tile_shape(c, xs, &dy, &dz, &tx);
                cands.push_back({tx, ty, tz});
            }
        }
        for (const auto& cd : cands) {
            if (st) break;
            c->tile_x = cd[0];
            c->tile_y = cd[1];
            c->tile_z = cd[2];
            float ms = 0.f;
            st = trial(&ms);
            if (!st && ms < best) { best = ms; best_tx = cd[0]; best_ty = cd[1]; best_tz = cd[2]; }
        }
        if (st) { best_tx = tx0; best_ty = ty0; best_tz = tz0; }
        c->tile_x = best_tx;
        c->tile_y = best_ty;
        c->tile_z = best_tz;
    }
    // 2. block order of the step kernel x grid of the deferred-cell kernels
    const int grids[4] = {256, 1024, mpf::kWideGrid, 4 * mpf::kWideGrid};
    best = 1e30f;
    int best_x = x0, best_g = g0;
    for (int x = 0; x < 2 && !st; ++x)
        for (int gi = 0; gi < 4 && !st; ++gi) {
            c->xcd_order = x;
            c->deferred_grid = grids[gi];
            float ms = 0.f;
            st = trial(&ms);
            if (!st && ms < best) { best = ms; best_x = x; best_g = grids[gi]; }
        }
    if (st) {  // a failed trial keeps the previous shape
        best_x = x0;
        best_g = g0;
        c->tile_x = tx0;
        c->tile_y = ty0;
        c->tile_z = tz0;
    }
    c->xcd_order = best_x;
    c->deferred_grid = best_g;
    c->cur = cur0;
    // the state and the counters are restored whether or not a trial failed
    if (have_saved) {
        (void)hipStreamSynchronize(c->stream);
        mpf_status rs = mpf_restore_state(c, saved);
        if (!rs) rs = c->hip(hipMemcpyAsync(c->d_status, status0, sizeof(status0), hipMemcpyHostToDevice, c->stream), "autotune restore");
        if (!rs) rs = c->hip(hipMemcpyAsync(c->d_first_overflow, &first0, sizeof(first0), hipMemcpyHostToDevice, c->stream), "autotune restore");
        if (!rs) rs = c->hip(hipMemcpyAsync(c->d_warnings, &warn0, sizeof(warn0), hipMemcpyHostToDevice, c->stream), "autotune restore");
        if (!rs) rs = c->hip(hipMemcpyAsync(c->d_stats, stats0, sizeof(stats0), hipMemcpyHostToDevice, c->stream), "autotune restore");
        if (!rs) rs = c->hip(hipMemcpyAsync(c->d_obs, obs0, sizeof(obs0), hipMemcpyHostToDevice, c->stream), "autotune restore");
        c->stepped = stepped0;
        if (!rs) rs = c->hip(hipStreamSynchronize(c->stream), "autotune restore");
        if (!st) st = rs;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(saved);
    if (xcd_order) *xcd_order = best_x;
    if (deferred_grid) *deferred_grid = best_g;
    return st;
}

mpf_status mpf_set_tile_shape(mpf_ctx* c, int32_t tile_x, int32_t tile_y, int32_t tile_z)
{
    if (!c) return MPF_ERR_ARG;
    if (tile_x < 0) {
        c->tile_x = c->tile_y = c->tile_z = 0;
        return MPF_OK;
    }
    if (c->active < 2) return c->fail(MPF_ERR_UNSUPPORTED, "the tiled step kernel needs two or three active axes");
    int ty = 0, tz = 0, tx = 0;
    tile_shape(c, tile_x, &ty, &tz, &tx);
    if (tile_y > 0) ty = tile_y;
    if (tile_z > 0) tz = tile_z;
    if ((!c->dny && ty != 1) || (!c->dnz && tz != 1) || ty > c->ny || tz > c->nz || ty * tz > mpf::kBlock ||
        (ty + 2 * c->dny) * (tz + 2 * c->dnz) > mpf::kTileHalo)
        return c->fail(MPF_ERR_ARG, "tile shape out of range (tile_y * tile_z <= 256 cells, halo plane <= " +
                                        std::to_string(mpf::kTileHalo) + " cells, 1 on reduced axes)");
    c->tile_x = tx;
    c->tile_y = ty;
    c->tile_z = tz;
    return MPF_OK;
}

mpf_status mpf_get_tile_shape(mpf_ctx* c, int32_t* tile_x, int32_t* tile_y, int32_t* tile_z)
{
    if (!c) return MPF_ERR_ARG;
    if (tile_x) *tile_x = c->tile_x;
    if (tile_y) *tile_y = c->tile_y;
    if (tile_z) *tile_z = c->tile_z;
    return MPF_OK;
}

mpf_status mpf_set_launch_shape(mpf_ctx* c, int32_t xcd_order, int32_t deferred_grid)
{
    if (!c || deferred_grid < 1) return MPF_ERR_ARG;
    c->xcd_order = xcd_order ? 1 : 0;
    c->deferred_grid = deferred_grid;
    return MPF_OK;
}

// ------------------------------------------------------------------ ensembles

namespace {

void drop_graph(mpf_ensemble* e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

void free_group(mpf_group& g)
{
    dfree(g.d_a1); dfree(g.d_a2); dfree(g.d_g); dfree(g.d_first); dfree(g.d_gfirst);
}

// Host argument arrays of every group for the next two steps (members at their current parity).
// Returns true when they differ from what the device arrays hold.
bool build_args(mpf_ensemble* e, const double* dt)
{
    bool changed = false;
    for (auto& g : e->groups) {
        const size_t n = g.members.size();
        std::vector<mpf::StepArgs> a1(n), a2(n);
        std::vector<mpf::GrainArgs> ga(n);
        std::vector<uint32_t> first(n + 1, 0u), gfirst(n + 1, 0u);
        int per = 1;
        for (size_t k = 0; k < n; ++k) {
            const mpf_ctx* c = e->members[g.members[k]];
            const double d = dt[g.members[k]];
            a1[k] = step_args(c, d, c->cur);
            a2[k] = step_args(c, d, 1 - c->cur);
            a1[k].xcd_order = a2[k].xcd_order = 1;
            ga[k] = grain_args(c);
            first[k + 1] = first[k] + (g.tiled ? tile_blocks(c) : step_blocks(c));
            gfirst[k + 1] = gfirst[k] + grain_blocks(c);
            // deferred-cell blocks per member: at most 256 (a 64-domain 2D ensemble: 7.0e9
            // cell-updates/s at 256, 6.8e9 at 512, 6.5e9 at 1024, 6.2e9 at 128; most blocks of a
            // small domain find no deferred cell)
            per = std::max(per, std::min(c->deferred_grid, 256));
        }
        auto differs = [](const void* x, const void* y, size_t bytes) { return std::memcmp(x, y, bytes) != 0; };
        if (g.h_a1.size() != n || differs(a1.data(), g.h_a1.data(), n * sizeof(mpf::StepArgs)) ||
            differs(a2.data(), g.h_a2.data(), n * sizeof(mpf::StepArgs)) ||
            differs(ga.data(), g.h_g.data(), n * sizeof(mpf::GrainArgs)) || g.per_domain != per) {
            changed = true;
            g.h_a1 = std::move(a1);
            g.h_a2 = std::move(a2);
            g.h_g = std::move(ga);
            g.h_first = std::move(first);
            g.h_gfirst = std::move(gfirst);
            g.per_domain = per;
        }
    }
    return changed;
}

// Group of a member: kernel variant (precision; a member with a driving force runs the full
// variant whether mpf_set_step_terms switched it on or off - off: an all-zero mode table - so the
// grouping holds across switches; uniform tables) and step kernel (tiled or per-cell).
std::array<bool, 4> group_key(const mpf_ctx* c)
{
    return {c->fp32, c->grad || c->dg, c->uniform, c->tile_x > 0};
}

// Sorts the members into groups by group_key; drops the graph.
void regroup(mpf_ensemble* e)
{
    drop_graph(e);
    for (auto& g : e->groups) free_group(g);
    e->groups.clear();
    for (size_t i = 0; i < e->members.size(); ++i) {
        const std::array<bool, 4> k = group_key(e->members[i]);
        mpf_group* g = nullptr;
        for (auto& q : e->groups)
            if (std::array<bool, 4>{q.fp32, q.full, q.uni, q.tiled} == k) g = &q;
        if (!g) {
            e->groups.emplace_back();
            g = &e->groups.back();
            g->fp32 = k[0];
            g->full = k[1];
            g->uni = k[2];
            g->tiled = k[3];
        }
        g->members.push_back(static_cast<int>(i));
    }
}

// Whether a member's group key changed since the grouping (new tables, tile shape).
bool regroup_needed(const mpf_ensemble* e)
{
    for (const auto& g : e->groups)
        for (int m : g.members)
            if (group_key(e->members[m]) != std::array<bool, 4>{g.fp32, g.full, g.uni, g.tiled}) return true;
    return false;
}

// Uploads the groups' argument arrays and captures two steps of every group into one graph:
// per step and group one launch of each kernel, over all its members.
mpf_status capture(mpf_ensemble* e)
{
    drop_graph(e);
    for (auto& g : e->groups) {
        const size_t n = g.members.size();
        free_group(g);
        hipError_t h = dalloc(&g.d_a1, n);
        if (h == hipSuccess) h = dalloc(&g.d_a2, n);
        if (h == hipSuccess) h = dalloc(&g.d_g, n);
        if (h == hipSuccess) h = dalloc(&g.d_first, n + 1);
        if (h == hipSuccess) h = dalloc(&g.d_gfirst, n + 1);
        if (h == hipSuccess) h = hipMemcpy(g.d_a1, g.h_a1.data(), n * sizeof(mpf::StepArgs), hipMemcpyHostToDevice);
        if (h == hipSuccess) h = hipMemcpy(g.d_a2, g.h_a2.data(), n * sizeof(mpf::StepArgs), hipMemcpyHostToDevice);
        if (h == hipSuccess) h = hipMemcpy(g.d_g, g.h_g.data(), n * sizeof(mpf::GrainArgs), hipMemcpyHostToDevice);
        if (h == hipSuccess) h = hipMemcpy(g.d_first, g.h_first.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (h == hipSuccess) h = hipMemcpy(g.d_gfirst, g.h_gfirst.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (h != hipSuccess) { e->err = "ensemble argument upload failed"; return MPF_ERR_HIP; }
    }
    if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        e->err = "hipStreamBeginCapture failed";
        return MPF_ERR_HIP;
    }
    int rc = 0;
    for (int half = 0; half < 2 && rc == 0; ++half)
        for (auto& g : e->groups) {
            const int n = static_cast<int>(g.members.size());
            mpf::StepArgs* args = half == 0 ? g.d_a1 : g.d_a2;
            const unsigned blocks = g.h_first[n], gblocks = g.h_gfirst[n];
            for (int k = 0; k < n && rc == 0; ++k) {
                const mpf_ctx* c = e->members[g.members[k]];
                if ((half == 0 ? g.h_a1[k] : g.h_a2[k]).dg_avg) rc = mpf::launch_dg_average(avg_args(c, half == 0 ? g.h_a1[k] : g.h_a2[k]), g.fp32, e->stream);
            }
            if (!rc)
                rc = g.tiled ? mpf::launch_tile_batch(args, g.d_first, n, blocks, g.fp32, g.full, g.uni, e->stream)
                             : mpf::launch_step(half == 0 ? g.h_a1[0] : g.h_a2[0], args, g.d_first, n, blocks, g.fp32,
                                                g.full, g.uni, e->stream);
            if (!rc) rc = mpf::launch_deferred(args, n, g.per_domain, g.fp32, g.full, e->stream);
            if (!rc) rc = mpf::launch_grains(g.d_g, g.d_gfirst, n, gblocks, e->stream);
            if (rc) break;
        }
    hipGraph_t graph = nullptr;
    const hipError_t h = hipStreamEndCapture(e->stream, &graph);
    if (rc || h != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        e->err = "ensemble capture failed";
        return MPF_ERR_HIP;
    }
    e->graph = graph;
    if (hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        drop_graph(e);
        e->err = "hipGraphInstantiate failed";
        return MPF_ERR_HIP;
    }
    return MPF_OK;
}

}  // namespace

mpf_status mpf_ensemble_create(mpf_ctx* const* members, int32_t n, mpf_ensemble** out)
{
    if (!members || n < 1 || !out) return MPF_ERR_ARG;
    *out = nullptr;
    for (int32_t i = 0; i < n; ++i) {
        if (!members[i] || !members[i]->initialized) return MPF_ERR_STATE;
        if (members[i]->reg) return MPF_ERR_UNSUPPORTED;  // members step without interface regularization
        for (int32_t j = 0; j < i; ++j)
            if (members[j] == members[i]) return MPF_ERR_ARG;
    }
    mpf_ensemble* e = new (std::nothrow) mpf_ensemble();
    if (!e) return MPF_ERR_NOMEM;
    e->members.assign(members, members + n);
    regroup(e);
    bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess;
    e->join.assign(n, nullptr);
    for (int32_t i = 0; i < n && ok; ++i) ok = hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        mpf_ensemble_destroy(e);
        return MPF_ERR_HIP;
    }
    *out = e;
    return MPF_OK;
}

mpf_status mpf_ensemble_step(mpf_ensemble* e, const double* dt, int32_t nsteps)
{
    if (!e || !dt || nsteps < 0) return MPF_ERR_ARG;
    const size_t n = e->members.size();
    for (size_t i = 0; i < n; ++i)
        if (!(dt[i] > 0.0)) return MPF_ERR_ARG;
    // members' earlier work on their own streams precedes the graph
    for (size_t i = 0; i < n; ++i)
        if (hipEventRecord(e->join[i], e->members[i]->stream) != hipSuccess || hipStreamWaitEvent(e->stream, e->join[i], 0) != hipSuccess)
            return MPF_ERR_HIP;
    mpf_status st = MPF_OK;
    if (regroup_needed(e)) regroup(e);
    if (nsteps >= 2 && (build_args(e, dt) || !e->exec)) st = capture(e);
    for (int32_t r = 0; r < nsteps / 2 && !st; ++r)
        if (hipGraphLaunch(e->exec, e->stream) != hipSuccess) { e->err = "hipGraphLaunch failed"; st = MPF_ERR_HIP; }
    if (!st && hipStreamSynchronize(e->stream) != hipSuccess) st = MPF_ERR_HIP;
    if (!st && (nsteps & 1))
        for (size_t i = 0; i < n && !st; ++i) st = enqueue_step(e->members[i], dt[i], e->members[i]->stream);
    for (size_t i = 0; i < n && !st; ++i) st = read_status(e->members[i]);
    if (!st && nsteps > 0)
        for (auto* m : e->members) m->stepped = m->prev_valid = true;
    return st;
}

const char* mpf_ensemble_error_message(mpf_ensemble* e)
{
    return e ? e->err.c_str() : "null ensemble";
}

void mpf_ensemble_destroy(mpf_ensemble* e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    for (auto& g : e->groups) free_group(g);
    for (auto ev : e->join) if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

}  // extern "C"

// ------------------------------------------------------------------ step without allocation
//
// mpf_step_into keeps the per-step scratch of a step - deferred-cell work lists and their
// counters, per-grain volume accumulators, status words and the argument blocks the deferred-cell
// and grain kernels read - in one file-scope device array instead of context allocations, and it
// reads and writes the caller's state buffers. Each GPU holds its own copy of the array from
// module load until the process ends; its device address is looked up once per device.

namespace {

constexpr int64_t kScratchCells = int64_t{1} << 19;  // cells per call (2D 724^2, 3D 80^3)
constexpr int kScratchGrains = 65536;                  // every grain count the library allows
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t{255}; }
// layout of the scratch array (byte offsets, 256-byte aligned parts)
constexpr size_t kOffStepArgs = 0;
constexpr size_t kOffGrainArgs = align256(kOffStepArgs + sizeof(mpf::StepArgs));
constexpr size_t kOffFirst = align256(kOffGrainArgs + sizeof(mpf::GrainArgs));      // 2 x uint32
constexpr size_t kOffCount = align256(kOffFirst + 2 * sizeof(uint32_t));            // 4 x uint32
constexpr size_t kOffStatus = align256(kOffCount + 4 * sizeof(unsigned));           // 4 x int32
constexpr size_t kOffWords = align256(kOffStatus + mpf::kStatusWords * sizeof(int32_t));  // 8 x u64
constexpr size_t kOffAcc = align256(kOffWords + 8 * sizeof(unsigned long long));
constexpr size_t kOffList = align256(kOffAcc + size_t{kScratchGrains} * mpf::kLimbs * sizeof(unsigned long long));
constexpr size_t kScratchBytes = align256(kOffList + 4 * size_t{kScratchCells} * sizeof(uint32_t));

// 10 MiB + 2.2 KiB per GPU
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device, 0 = not looked up

// Device address of this device's `scratch`: looked up on the first call on a device, with that
// device current (it is the current device of the calling thread).
mpf_status scratch_address(mpf_ctx* c, unsigned char** out)
{
    int dev = 0;
    if (mpf_status s = c->hip(hipGetDevice(&dev), "hipGetDevice")) return s;
    if (dev < 0 || dev >= kMaxDevices) return c->fail(MPF_ERR_UNSUPPORTED, "device index beyond the scratch table");
    void* p = __atomic_load_n(&g_scratch_ptr[dev], __ATOMIC_ACQUIRE);
    if (!p) {
        if (mpf_status s = c->hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return s;
        if (!p) return c->fail(MPF_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        __atomic_store_n(&g_scratch_ptr[dev], p, __ATOMIC_RELEASE);
    }
    *out = static_cast<unsigned char*>(p);
    return MPF_OK;
}

// Clears what a step expects at zero (volume accumulators of `ngrains`, list counters, status,
// warnings, IP.Set maxima, statistics; the first overflowing cell at its "none" value) and stores
// the argument blocks the deferred-cell and grain kernels read. A kernel rather than memset and
// copy calls, so the sequence records into a HIP graph as kernels only.
__global__ void scratch_setup_kernel(unsigned char* base, int32_t ngrains, mpf::StepArgs a, mpf::GrainArgs g,
                                     uint32_t grain_blocks)
{
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    auto* acc = reinterpret_cast<unsigned long long*>(base + kOffAcc);
    for (int64_t i = tid; i < static_cast<int64_t>(ngrains) * mpf::kLimbs; i += stride) acc[i] = 0ull;
    if (tid == 0) {
        *reinterpret_cast<mpf::StepArgs*>(base + kOffStepArgs) = a;
        *reinterpret_cast<mpf::GrainArgs*>(base + kOffGrainArgs) = g;
        auto* first = reinterpret_cast<uint32_t*>(base + kOffFirst);
        first[0] = 0u;
        first[1] = grain_blocks;
        auto* cnt = reinterpret_cast<unsigned*>(base + kOffCount);
        auto* st = reinterpret_cast<int32_t*>(base + kOffStatus);
        auto* w = reinterpret_cast<unsigned long long*>(base + kOffWords);
        for (int i = 0; i < 4; ++i) cnt[i] = 0u;
        for (int i = 0; i < mpf::kStatusWords; ++i) st[i] = 0;
        w[0] = ~0ull;                       // first overflowing cell: none
        for (int i = 1; i < 8; ++i) w[i] = 0ull;  // warnings, IP.Set maxima (4), statistics (2)
    }
}

}  // namespace

extern "C" {

mpf_status mpf_step_into(mpf_ctx* c, double dt, const void* count_in, const void* index_in, const void* value_in,
                         void* count_out, void* index_out, void* value_out)
{
    if (!c) return MPF_ERR_ARG;
    if (!c->initialized) return c->fail(MPF_ERR_STATE, "mpf_finalize_initialization must precede mpf_step_into");
    if (!(dt > 0.0)) return c->fail(MPF_ERR_ARG, "dt must be > 0");
    if (!count_in || !index_in || !value_in || !count_out || !index_out || !value_out)
        return c->fail(MPF_ERR_ARG, "null state buffer");
    if (count_in == count_out || index_in == index_out || value_in == value_out)
        return c->fail(MPF_ERR_ARG, "input and output state buffers must differ");
    if (c->ncells > kScratchCells)
        return c->fail(MPF_ERR_UNSUPPORTED, "mpf_step_into: " + std::to_string(c->ncells) + " cells exceed its scratch limit of " +
                                                std::to_string(kScratchCells) + " cells per call (use mpf_step)");
    if (c->ngrains > kScratchGrains) return c->fail(MPF_ERR_UNSUPPORTED, "mpf_step_into: too many grains");
    if (c->reg) return c->fail(MPF_ERR_UNSUPPORTED, "mpf_step_into: interface regularization needs mpf_step");
    if (avg_active(c))
        return c->fail(MPF_ERR_UNSUPPORTED, "mpf_step_into: driving-force averaging keeps per-pair arrays of the context (use mpf_step)");
    unsigned char* base = nullptr;
    if (mpf_status s = scratch_address(c, &base)) return s;

    mpf::StepArgs a = step_args(c, dt, c->cur);
    a.cnt_in = static_cast<const uint8_t*>(count_in);
    a.id_in = static_cast<const uint16_t*>(index_in);
    a.val_in = value_in;
    a.cnt_out = static_cast<uint8_t*>(count_out);
    a.id_out = static_cast<uint16_t*>(index_out);
    a.val_out = value_out;
    a.vol_acc = reinterpret_cast<unsigned long long*>(base + kOffAcc);
    a.status = reinterpret_cast<int32_t*>(base + kOffStatus);
    auto* words = reinterpret_cast<unsigned long long*>(base + kOffWords);
    a.first_overflow = words;
    a.warnings = words + 1;
    a.obs = words + 2;
    a.work_list = reinterpret_cast<uint32_t*>(base + kOffList);
    a.work_count = reinterpret_cast<unsigned*>(base + kOffCount);
    mpf::GrainArgs g = grain_args(c);
    g.vol_acc = a.vol_acc;
    g.work_count = a.work_count;
    g.obs = words + 2;
    g.stats = words + 6;
    const unsigned gblocks = grain_blocks(c);

    hipLaunchKernelGGL(scratch_setup_kernel, dim3(64), dim3(256), 0, c->stream, base, c->ngrains, a, g, gblocks);
    if (mpf_status s = c->hip(hipGetLastError(), "scratch set-up kernel")) return s;
    const bool full = full_variant(c);
    mpf_status st = c->hip(launch_main(c, a, c->stream), "step kernel");
    if (!st)
        st = c->hip(static_cast<hipError_t>(mpf::launch_deferred(reinterpret_cast<const mpf::StepArgs*>(base + kOffStepArgs), 1,
                                                                 c->deferred_grid, c->fp32, full, c->stream)),
                    "deferred-cell kernels");
    if (!st)
        st = c->hip(static_cast<hipError_t>(mpf::launch_grains(reinterpret_cast<const mpf::GrainArgs*>(base + kOffGrainArgs),
                                                               reinterpret_cast<const uint32_t*>(base + kOffFirst), 1, gblocks,
                                                               c->stream)), "grain kernel");
    if (st) return st;
    int32_t status[mpf::kStatusWords];
    st = c->hip(hipMemcpyAsync(status, base + kOffStatus, sizeof(status), hipMemcpyDeviceToHost, c->stream), "status readback");
    if (!st) st = c->hip(hipStreamSynchronize(c->stream), "stream synchronize");
    if (st) return st;
    if (status[mpf::kStatusOverflow]) return c->fail(MPF_SLOT_OVERFLOW, "a cell needed more phase-field slots than the capacity");
    if (status[mpf::kStatusEmpty]) return c->fail(MPF_EMPTY_NODE, "finalize erased every entry of a cell");
    if (status[mpf::kStatusVolumeRange]) return c->fail(MPF_ERR_STATE, "phase-field value outside [0, 2) in the volume sum");
    return MPF_OK;
}

}  // extern "C"
