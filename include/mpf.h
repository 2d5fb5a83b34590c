/*
 * mpf.h -- C ABI of the MI355X multiphase-field (MPF) time step.
 *
 * This is the boundary behind which OpenPhase's single-resolution grain-growth step runs on one
 * MI355X (gfx950). Each entry point replaces one reference call (cited below); the C++ classes in
 * include/openphase_amd/OpenPhaseAMD.h keep the reference class/method names on top of it.
 *
 * State: per cell a fixed-capacity, insertion-ordered list of at most `slots` (phase-field index,
 * value) entries, stored slot-major (structure of arrays) in device memory owned by the context.
 * Host buffers are owned by the caller. Calls on one context are not thread-safe; all device work
 * is ordered on the context's stream. No entry point ever exits the process: every failure is a
 * status code, with details from mpf_error_message().
 */
#ifndef OPENPHASE_AMD_MPF_H
#define OPENPHASE_AMD_MPF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mpf_status {
    MPF_OK = 0,
    MPF_ERR_ARG = 1,               /* invalid argument (message says which) */
    MPF_ERR_UNSUPPORTED = 2,       /* configuration outside what the library implements */
    MPF_ERR_HIP = 3,               /* HIP runtime error */
    MPF_ERR_NOMEM = 4,             /* device or host allocation failed */
    MPF_SLOT_OVERFLOW = 5,         /* a cell needed more than `slots` entries; see mpf_overflow_cell */
    MPF_EMPTY_NODE = 6,            /* finalize erased every entry of a cell (reference: undefined) */
    MPF_ERR_STATE = 7              /* call out of order (e.g. step before upload) */
} mpf_status;

typedef enum mpf_precision {
    MPF_FP64 = 0,                  /* FP64 storage and arithmetic: the reference's arithmetic */
    MPF_FP32 = 1                   /* FP32 storage, stencils and increments; FP64 limiter,
                                      normalization, finalize and grain volumes */
} mpf_precision;

typedef enum mpf_bc {              /* BoundaryConditionTypes, include/BoundaryConditions.h:37-45 */
    MPF_BC_PERIODIC = 0,
    MPF_BC_NOFLUX = 1,
    MPF_BC_FREE = 2,
    MPF_BC_FIXED = 3,
    MPF_BC_MIRROR = 4
} mpf_bc;

typedef enum mpf_tj_model {        /* TripleJunctionModels, src/InterfaceProperties.cpp:274-290 */
    MPF_TJ_MODEL = 0,
    MPF_TJ_VALUE = 1,
    MPF_TJ_NONE = 2
} mpf_tj_model;

typedef enum mpf_energy_model {    /* InterfaceEnergyModels subset, InterfaceEnergyModel.h:36-49 */
    MPF_ENERGY_ISO = 0,
    MPF_ENERGY_CUBIC = 1
} mpf_energy_model;

typedef enum mpf_dg_weights {      /* AveragingWeightsModes, include/DrivingForce.h:46-51 */
    MPF_DG_WEIGHTS_PHASEFIELDS = 0,
    MPF_DG_WEIGHTS_RANGE = 1,
    MPF_DG_WEIGHTS_COUNTER = 2
} mpf_dg_weights;

/* GridParameters (include/GridParameters.h:37-172) + BoundaryConditions + storage options. */
typedef struct mpf_grid {
    int32_t nx, ny, nz;            /* cells per axis; 0 = reduced axis (GridParameters.cpp:645-706) */
    double dx;                     /* grid spacing [m] */
    double iwidth;                 /* interface width in cells; Eta = iwidth*dx */
    int32_t bc[6];                 /* BC0X, BCNX, BC0Y, BCNY, BC0Z, BCNZ (mpf_bc) */
    int32_t precision;             /* mpf_precision */
    int32_t slots;                 /* per-cell stored-list capacity K, 2..16 (8 for the benchmarks) */
} mpf_grid;

/* InterfaceProperties / PhaseField / DrivingForce parameters for Nphases thermodynamic phases.
 * Tables are nphases*nphases, row-major, indexed [phaseA*nphases + phaseB]. */
typedef struct mpf_params {
    int32_t nphases;
    const double* sigma;           /* Sigma_n_m   interface energy [J/m^2] */
    const double* mu;              /* Mu_n_m      interface mobility [m^4/(J s)] */
    const int32_t* energy_model;   /* EnergyModel_n_m (mpf_energy_model), NULL = all ISO */
    const double* epsilon;         /* EpsilonE_n_m for CUBIC, NULL = 0 */
    int32_t tj_model;              /* mpf_tj_model */
    double tj_factor;              /* TripleJunctionFactor */
    double tj_energy;              /* TripleJunctionEnergy */
    int32_t consider_nucleus_volume; /* PhaseField ConsiderNucleusVolume (default 1) */
    double nucleus_volume_factor;  /* PhaseField NucleusVolumeFactor (default 1.0) */
    const int32_t* dg_mode;        /* UDF_Mode_n_m: 1 = VALUE, 0 = NONE; NULL = no driving force */
    const double* dg_value;        /* UDF_Value_n_m, antisymmetric [J/m^3] */
    double dg_limit;               /* DrivingForce Limit_n_m (default 0.95) */
    double regularization_factor;  /* InterfaceProperties RegularizationFactor (default 1.0) */
    /* DrivingForce::Average (src/DrivingForce.cpp:242-255, 293-478), read by ReadInput (:126-152).
     * The reference's default is Average = Yes; 0 here selects SkipAverage (:257-274). */
    int32_t dg_average;            /* Average: 1 = CollectAverage + DistributeAverage */
    int32_t dg_range;              /* Range in cells; <= 0: (int)iwidth */
    double dg_threshold;           /* Threshold (PhiThreshold); <= 0: iwidth < 5 ? iwidth/15 : 1/3 */
    int32_t dg_weights;            /* WeightsMode (mpf_dg_weights) */
    /* InterfaceRegularization (src/InterfaceRegularization.cpp): regularization = 1 makes every step
     * DoubleObstacle::CalculatePhaseFieldIncrements(Phase, IP, Kappa) (src/DoubleObstacle.cpp:644-770:
     * increments times RegularizationFactor, curvature (CurvatureFactor - RegularizationFactor) per
     * pair), then Kappa.Average (:156-176) and Kappa.MergePhaseFieldIncrements (:403-456), before the
     * driving force. ClearingMode AUTOMATIC. FP64 storage, single contexts (no adjoint, ensemble or
     * mpf_step_into). */
    int32_t regularization;
    double curvature_factor;       /* InterfaceProperties CurvatureFactor (default 1.0) */
    int32_t reg_average;           /* Average (default Yes): SetWeights + CollectAverage + DistributeAverage */
    int32_t reg_range;             /* Range in cells; <= 0: (int)((iwidth + 1)/2) */
    double reg_threshold;          /* Threshold (PhiThreshold); <= 0: DBL_EPSILON */
} mpf_params;

typedef struct mpf_ctx mpf_ctx;

/* Construction: PhaseField/InterfaceProperties/BoundaryConditions/DoubleObstacle constructors
 * (Initialize + ReadInput, include/OPObject.h:64-77). Allocates the device state. */
mpf_status mpf_create(const mpf_grid* grid, const mpf_params* params, mpf_ctx** out);

/* Grains: PhaseField::AddGrainInfo x n (src/PhaseField.cpp:2970-2979) + per-grain orientation
 * (GrainsProperties Orientation, rotation matrix row-major 9 doubles, NULL = identity) and
 * aggregate state (1 = solid, NULL = all solid). Resets Volume/MAXVolume, VolumeRatio = 0. */
mpf_status mpf_set_grains(mpf_ctx* ctx, int32_t ngrains, const int32_t* phase,
                          const double* rotation, const uint8_t* solid);

/* PhaseField::Read(FileName) (src/PhaseField.cpp:2487-2543, NodePF::read NodePF.h:751-763):
 * count[ncells] entries per cell, then the concatenated (index, value) lists, cells ordered
 * x outer, z inner (the raw .dat order). Host buffers; copied to the device. */
mpf_status mpf_upload(mpf_ctx* ctx, const int32_t* count, const int32_t* index, const double* value);

/* PhaseField::FinalizeInitialization (src/PhaseField.cpp:1068-1077). */
mpf_status mpf_finalize_initialization(mpf_ctx* ctx);

/* One full step, nsteps times, at fixed dt, on the device-resident state:
 *   InterfaceProperties::Set (src/InterfaceProperties.cpp:557-576)
 *   DoubleObstacle::CalculatePhaseFieldIncrements (src/DoubleObstacle.cpp:461-500)
 *   [UserDrivingForce::SetDrivingForce + DrivingForce::Average (no averaging) +
 *    DrivingForce::MergePhaseFieldIncrements when dg_mode is set]
 *   PhaseField::NormalizeIncrements (src/PhaseField.cpp:1334-1349)
 *   PhaseField::MergeIncrements + Finalize (src/PhaseField.cpp:1248-1266, 989-1033)
 * Enqueued on the context's stream; returns after the last step completed and its status was
 * read back (one synchronisation per call). */
mpf_status mpf_step(mpf_ctx* ctx, double dt, int32_t nsteps);

/* Same as mpf_step but does not synchronise; errors surface at the next synchronising call. */
mpf_status mpf_step_async(mpf_ctx* ctx, double dt, int32_t nsteps);

/* Blocks until all work on the context's stream is done; returns the accumulated status. */
mpf_status mpf_synchronize(mpf_ctx* ctx);

/* Total entries over all cells (size of the index/value download buffers). */
mpf_status mpf_total_entries(mpf_ctx* ctx, int64_t* total);

/* PhaseField::Write (src/PhaseField.cpp:2340-2388) content: finalized lists in the same layout
 * as mpf_upload; flag (0/1/2) per cell may be NULL. */
mpf_status mpf_download(mpf_ctx* ctx, int32_t* count, int32_t* index, double* value, int8_t* flag);

/* GrainsProperties Volume and VolumeRatio per grain (either may be NULL). */
mpf_status mpf_get_grains(mpf_ctx* ctx, double* volume, double* volume_ratio);

/* GrainsProperties MAXVolume, RefVolume, Exist and Stage (0 Seed, 1 Nucleus, 2 Stable) per grain,
 * any may be NULL: the Grain fields GrainsProperties::Write stores (include/GrainsProperties.h:196-230)
 * besides Phase and Orientation, which the caller set. */
mpf_status mpf_get_grain_records(mpf_ctx* ctx, double* maxvolume, double* refvolume, uint8_t* exist, int32_t* stage);

/* InterfaceProperties::ReportMaximumTimeStep (src/InterfaceProperties.cpp:523-556): c dx^2 /
 * (maxMobility maxEnergy RegularizationFactor) from the largest interface energy and mobility the
 * last step's IP.Set evaluated over the pairs of its interface cells (SetSR, :603-756; CUBIC
 * energies as computed from the orientations). DBL_MAX before the first step, as the reference
 * before its first IP.Set. */
mpf_status mpf_max_time_step(mpf_ctx* ctx, double* dt_max);

/* State the interface energies of InterfaceProperties::Properties come from (what the last
 * IP.Set evaluated): no IP.Set yet, the current state, or the state the last step started from
 * (the usual IP.Set -> step -> energy order). */
typedef enum mpf_ip_state {
    MPF_IP_NONE = 0,
    MPF_IP_CURRENT = 1,
    MPF_IP_PREVIOUS = 2
} mpf_ip_state;

/* DoubleObstacle::PointEnergy (src/DoubleObstacle.cpp:903-923) summed over every cell of the
 * current state: DoubleObstacle::AverageEnergyDensity (:989-1005) is sum / cells, Energy
 * (:953-967) sum * cell volume. Per wide-interface cell and pair of its neighbour-merged list,
 * 4 E (phi_a phi_b - Eta^2/Pi^2 grad_a . grad_b) / Eta, E the pair's energy from the IP.Set state
 * named by ip_state (0 for a pair absent there). Summed on the device in a fixed order (the same
 * from run to run; the reference's sequential sum differs in rounding only). MPF_IP_PREVIOUS needs
 * a step since the last upload / restore / finalize. */
mpf_status mpf_interface_energy(mpf_ctx* ctx, int32_t ip_state, double* sum);

/* PhaseField::CalculateFractions (src/PhaseField.cpp:411-423), which the reference runs in every
 * Finalize: per cell and thermodynamic phase the sum of its grains' values, computed on the
 * device from the current state into dst[phase * cells + cell] (cells in x-outer, z-inner
 * order): caller device memory of nphases * cells doubles, or host memory when dst_on_host. */
mpf_status mpf_phase_fractions(mpf_ctx* ctx, double* dst, int32_t dst_on_host);

/* Terms of the following steps (default: as created). driving_force 0 leaves the constant
 * driving force out: the reference applies it only in steps whose driver calls
 * UserDrivingForce::SetDrivingForce (src/UserDrivingForce.cpp:152-175), DrivingForce::Average
 * (src/DrivingForce.cpp:242-256) and DrivingForce::MergePhaseFieldIncrements (:480-495); 1 needs
 * a context created with a driving force. average 0 takes the raw value (DrivingForce::SkipAverage,
 * :258-274) in a context created with averaging; without averaging it has no effect. */
mpf_status mpf_set_step_terms(mpf_ctx* ctx, int32_t driving_force, int32_t average);

/* Number of limiter plausibility warnings (PhaseField.cpp:1536-1561) since creation. */
mpf_status mpf_limiter_warnings(mpf_ctx* ctx, int64_t* count);

/* First cell (linear index) that overflowed the slot capacity, -1 if none. */
mpf_status mpf_overflow_cell(mpf_ctx* ctx, int64_t* cell);

/* Device pointers of the current state (for device-resident consumers): counts (uint8),
 * indices (uint16, slot-major [slot*ncells + cell]), values (double or float by precision),
 * and the HIP stream (hipStream_t) the context runs on. Any output pointer may be NULL. */
mpf_status mpf_device_state(mpf_ctx* ctx, void** count, void** index, void** value, void** stream);

/* Kernel timing of the last mpf_step call (timing on), average per step in ms, measured with HIP
 * events on the context's stream: the main step kernel (with the driving-force averaging
 * kernels when they run), and the rest of the step - the deferred-cell kernels and the grain
 * kernel - in `rest_ms`. */
mpf_status mpf_last_timing(mpf_ctx* ctx, double* step_kernel_ms, double* rest_ms);

/* Enables/disables per-kernel HIP event timing in mpf_step (off by default). */
mpf_status mpf_enable_timing(mpf_ctx* ctx, int32_t on);

/* Cumulative number of cell-steps finished outside the main kernel since creation: merged lists
 * of 5-6 entries (or more than 8 own entries) and lists of 1, 7-16 entries (no reference
 * counterpart: the reference has one code path). */
mpf_status mpf_step_stats(mpf_ctx* ctx, int64_t* deferred_mid, int64_t* deferred_wide);

/* Measurement helper (no reference counterpart): copies `bytes` (multiple of 16) between two
 * device buffers owned by the context with a plain streaming copy kernel `repeats` times and
 * returns the average kernel time in ms; the HBM bandwidth reference for the roofline and the
 * calibration kernel for rocprof FETCH_SIZE/WRITE_SIZE. */
mpf_status mpf_copy_bandwidth(mpf_ctx* ctx, int64_t bytes, int32_t repeats, double* ms);

/* Replaces the phase-pair tables (nphases x nphases, row-major; any pointer may be NULL to keep
 * that table): interface energy (InterfaceProperties' Energy / EnergyModel, src/InterfaceProperties.cpp
 * :229-291), mobility (MaxMobility), constant driving force (UserDrivingForce Value,
 * src/UserDrivingForce.cpp:60-98; the context must have been created with one). MaxEnergy and the
 * driving-force limiter amplitude are rederived as at creation. For calibration loops. */
mpf_status mpf_set_pair_params(mpf_ctx* ctx, const double* sigma, const double* mu, const double* dg_value);

/* Checkpoints (no reference counterpart; segment checkpointing of adjoint rollouts): the size of
 * a saved state, and save / restore of the current state - lists and grain volumes, MAXVolume,
 * VolumeRatio, Exist, Stage - to / from caller-owned DEVICE memory of that size. */
mpf_status mpf_state_bytes(mpf_ctx* ctx, int64_t* bytes);
mpf_status mpf_save_state(mpf_ctx* ctx, void* device_dst);
mpf_status mpf_restore_state(mpf_ctx* ctx, const void* device_src);

/* Adjoint of one step from the current state S_t with time step dt (no reference counterpart:
 * the reference has no derivatives). All pointers are DEVICE memory:
 *   lam_next    [slots * ncells]  cotangent of the values of S_{t+1}, slot-major like the state;
 *   lam_prev    [slots * ncells]  written: cotangent of the values of S_t (0 beyond each count);
 *   grad        [3 * nphases^2]   accumulated: cotangents of the sigma, mu and driving-force
 *                                 tables (entries as read, before any symmetrization), summed over
 *                                 the cells in a fixed order (the same from run to run);
 *   next_count, next_index        S_{t+1}'s count bytes and grain indices (the layout of
 *                                 mpf_device_state / mpf_save_state), mapping lam_next to the
 *                                 entries by grain; both NULL: S_{t+1} is the step's own result
 *                                 from S_t (FP64 only).
 * Derivatives follow the branches the forward step takes (clamps, erasures, limiter factors);
 * VolumeRatios are held fixed. FP32 states are differentiated at their values in FP64. At most 4
 * phases, merged lists of at most 8 entries and no driving-force averaging (MPF_ERR_UNSUPPORTED
 * otherwise). */
mpf_status mpf_adjoint_step(mpf_ctx* ctx, double dt, const double* lam_next, double* lam_prev, double* grad,
                            const void* next_count, const void* next_index);

/* Launch-shape autotuner (no reference counterpart): times `steps` steps of each candidate on the
 * context's own state (saved and restored around every trial, so the state is unchanged
 * afterwards) and keeps the fastest, in two rounds: (1) the step kernel's tile (grids with two or
 * three active axes) - plane tiles of 64-256 cells (2D: 64 / 128 / 256 along the active plane
 * axis; 3D: 4x64, 8x32, 16x16, 4x32) times 1 / 4 / 16 / the default number of planes per block,
 * and the per-cell kernel; (2) the XCD-contiguous block order of the step kernel on/off times
 * 256 / 1024 / 4096 / 16384 blocks for the deferred-cell kernels. Results do not depend on the
 * shape. The chosen launch shape is returned through the optional out pointers, the tile by
 * mpf_get_tile_shape. */
mpf_status mpf_autotune(mpf_ctx* ctx, double dt, int32_t steps, int32_t* xcd_order, int32_t* deferred_grid);
mpf_status mpf_set_launch_shape(mpf_ctx* ctx, int32_t xcd_order, int32_t deferred_grid);

/* Shape of the LDS-tiled step kernel (no reference counterpart): a block holds a tile_y x tile_z
 * tile of the (y, z) plane (at most 256 cells, its halo plane at most 396 cells; 1 on reduced
 * axes; 0 = the default for that axis) and marches through tile_x planes along x (0 = about 8192
 * blocks). tile_x < 0 selects the per-cell step kernel. Contexts with two or three active axes
 * start with the default tile; 1D grids always use the per-cell kernel. Results do not depend on
 * the shape. */
mpf_status mpf_set_tile_shape(mpf_ctx* ctx, int32_t tile_x, int32_t tile_y, int32_t tile_z);
mpf_status mpf_get_tile_shape(mpf_ctx* ctx, int32_t* tile_x, int32_t* tile_y, int32_t* tile_z);

/* Ensembles of independent domains (no reference counterpart; calibration runs over many uneven
 * microstructures): the members (initialized contexts, each with its own grid, tables and dt)
 * advance together. Two steps of every member are captured into one HIP graph (one branch per
 * member) and replayed nsteps/2 times; an odd last step runs directly. A changed dt or member
 * parity recaptures the graph. mpf_ensemble_step returns after all members finished and reports
 * the first member error. Members must not be stepped elsewhere while an ensemble step runs. */
typedef struct mpf_ensemble mpf_ensemble;
mpf_status mpf_ensemble_create(mpf_ctx* const* members, int32_t n, mpf_ensemble** out);
mpf_status mpf_ensemble_step(mpf_ensemble* ens, const double* dt, int32_t nsteps);
/* Description of the last error of the ensemble itself (graph capture / launch); member errors
 * are reported by mpf_error_message of the member. */
const char* mpf_ensemble_error_message(mpf_ensemble* ens);
void mpf_ensemble_destroy(mpf_ensemble* ens);

/* One step without any allocation (no reference counterpart; for callers that cannot allocate
 * device memory: code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Advances the state in the caller's DEVICE buffers count_in / index_in / value_in (the layout of
 * mpf_device_state: uint8 counts, uint16 indices and FP64 or FP32 values slot-major with the
 * context's slot capacity, as this library wrote them) by one mpf_step into count_out / index_out /
 * value_out (same layout, distinct buffers) with the context's grains, tables and launch shape;
 * the context's own state is not touched, its grain records (Volume, VolumeRatio, ...) become
 * those of the output state. The step's scratch (deferred-cell work lists, volume accumulators,
 * status words) is one static device array per GPU, cleared at the start of every call: at most
 * 2^19 cells per call (MPF_ERR_UNSUPPORTED beyond, and with driving-force averaging), 10 MiB per
 * GPU held until the process ends. Calls that overlap on one GPU (other threads, other streams)
 * race on that array. Synchronises on the context's stream and returns the step's status. */
mpf_status mpf_step_into(mpf_ctx* ctx, double dt, const void* count_in, const void* index_in, const void* value_in,
                         void* count_out, void* index_out, void* value_out);

/* Human-readable description of the last error on this context (static storage per context). */
const char* mpf_error_message(mpf_ctx* ctx);

void mpf_destroy(mpf_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
