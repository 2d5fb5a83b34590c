/*
 * CPU oracle for the multiphase-field (MPF) time step -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch FP64 restatement, in C, of OpenPhase's single-resolution step
 *   InterfaceProperties::Set -> DoubleObstacle::CalculatePhaseFieldIncrements ->
 *   PhaseField::NormalizeIncrements -> PhaseField::MergeIncrements (-> Finalize)
 * following the reference loops pass by pass (list insertion order, summation order and the
 * DBL_EPSILON thresholds). Only tests, the smoke check and the CPU baseline leg of the benchmark
 * may load it; the product library never links it.
 *
 * Per-node phase lists are fixed-capacity arrays (ORC_CAP entries); exceeding the capacity is
 * reported as an error, never truncated.
 */
#ifndef MPF_ORACLE_H
#define MPF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_CAP 16

enum orc_status {
    ORC_OK = 0,
    ORC_ERR_ARG = 1,
    ORC_ERR_CAPACITY = 2,      /* a node list exceeded ORC_CAP */
    ORC_ERR_EMPTY_NODE = 3,    /* finalize erased every entry of a node */
    ORC_ERR_NOMEM = 4
};

enum orc_tj_model { ORC_TJ_MODEL = 0, ORC_TJ_VALUE = 1, ORC_TJ_NONE = 2 };
enum orc_energy_model { ORC_EN_ISO = 0, ORC_EN_CUBIC = 1 };

typedef struct orc_params {
    int32_t nx, ny, nz;             /* grid size; 0 marks a reduced (inactive) axis */
    double dx;                      /* grid spacing */
    double iwidth;                  /* interface width in cells (Eta = iwidth*dx) */
    int32_t nphases;
    const double* sigma;            /* nphases*nphases interface energy table (phase pairs) */
    const double* mu;               /* nphases*nphases interface mobility table */
    const int32_t* energy_model;    /* nphases*nphases, orc_energy_model (NULL = all ISO) */
    const double* epsilon;          /* nphases*nphases cubic anisotropy strength (NULL = 0) */
    int32_t tj_model;               /* orc_tj_model */
    double tj_factor;               /* TripleJunctionFactor (MODEL) */
    double tj_energy;               /* TripleJunctionEnergy (VALUE) */
    int32_t consider_nucleus_volume;
    double nucleus_volume_factor;
    /* constant driving force between phase pairs (UserDrivingForce VALUE mode), NULL = none */
    const int32_t* dg_mode;         /* nphases*nphases, 1 = VALUE mode for the pair (UDF_Mode_n_m) */
    const double* dg_value;         /* nphases*nphases, antisymmetric (UDF_Value_n_m) */
    double dg_limit;                /* DrivingForce::Limit (0.95 default) */
    double regularization_factor;   /* InterfaceProperties::RegularizationFactor (1.0) */
    /* DrivingForce::Average (src/DrivingForce.cpp:242-255, 293-478): 0 = SkipAverage */
    int32_t dg_average;
    int32_t dg_range;               /* Range in cells; <= 0: (int)iwidth (DrivingForce.cpp:64) */
    double dg_threshold;            /* PhiThreshold; <= 0: iwidth < 5 ? iwidth/15 : 1/3 (:66-73) */
    int32_t dg_weights;             /* WeightsMode: 0 PHASEFIELDS, 1 RANGE, 2 COUNTER (:135-152) */
    /* InterfaceRegularization: 1 = the steps use DoubleObstacle's curvature overload
     * (src/DoubleObstacle.cpp:686-770) + InterfaceRegularization::Average / MergePhaseFieldIncrements */
    int32_t regularization;
    double curvature_factor;        /* InterfaceProperties::CurvatureFactor (1.0) */
    int32_t reg_average;            /* Average (Yes) */
    int32_t reg_range;              /* Range; <= 0: (int)((iwidth + 1)/2) (InterfaceRegularization.cpp:61) */
    double reg_threshold;           /* Threshold; <= 0: DBL_EPSILON (:59) */
} orc_params;

typedef struct orc_ctx orc_ctx;

/* grains: phase index per grain (ngrains), rotation matrices row-major 9 doubles per grain (NULL = identity),
 * solid flag per grain (NULL = all solid). */
int orc_create(const orc_params* p, int32_t ngrains, const int32_t* grain_phase,
               const double* grain_rot, const uint8_t* grain_solid, orc_ctx** out);
void orc_destroy(orc_ctx* c);

/* Load raw lists (PhaseField::Read semantics, src/PhaseField.cpp:2487-2543 + NodePF::read):
 * count[ncells], then concatenated (index, value) pairs in cell order x-outer, z-inner. */
int orc_load(orc_ctx* c, const int32_t* count, const int32_t* index, const double* value);
/* PhaseField::FinalizeInitialization (src/PhaseField.cpp:1068-1077): finalize + flags + derivatives + volumes */
int orc_finalize_initialization(orc_ctx* c);
/* nsteps full steps at fixed dt */
int orc_step(orc_ctx* c, double dt, int32_t nsteps);
/* Number of entries with value != 0 over all cells (size of the download buffers). */
int64_t orc_total_entries(const orc_ctx* c);
/* Finalized lists (entries with value != 0, in list order), flags (0/1/2) per cell. */
int orc_download(const orc_ctx* c, int32_t* count, int32_t* index, double* value, int8_t* flag);
/* Per-grain Volume and VolumeRatio. */
int orc_grains(const orc_ctx* c, double* volume, double* volume_ratio);
/* GrainsProperties MAXVolume, RefVolume, Exist and Stage per grain (any may be NULL). */
int orc_grain_records(const orc_ctx* c, double* maxvolume, double* refvolume, uint8_t* exist, int32_t* stage);
/* Limiter plausibility warnings seen so far (PhaseField.cpp:1536-1561). */
int64_t orc_limiter_warnings(const orc_ctx* c);

/* PhaseField::NormalizeIncrementsSR on one interface node given directly: its list (n entries,
 * phase-field index and value, in list order) and its pair increments (npairs in: the pairs added
 * with NodeAB::add_asym1 in that order; out: the pairs left, in storage order, with their limited
 * value1). Counts the plausibility warnings the reference would print. */
int orc_normalize_pairs(int32_t n, const int32_t* idx, const double* val, int32_t* npairs, int32_t* pa, int32_t* pb,
                        double* v1, double dt, int64_t* warnings);
/* InterfaceProperties::ReportMaximumTimeStep for the last IP.Set (src/InterfaceProperties.cpp:523-556). */
double orc_max_time_step(const orc_ctx* c);
void orc_set_threads(int32_t n);

/* Per-step switches of the driving force. The reference applies it in the steps whose driver
 * calls UserDrivingForce::SetDrivingForce, DrivingForce::Average and
 * DrivingForce::MergePhaseFieldIncrements: driving_force 0 leaves it out of the following steps;
 * average 0 takes the raw value (DrivingForce::SkipAverage) in a context created with averaging. */
int orc_set_step_terms(orc_ctx* c, int32_t driving_force, int32_t average);
/* Keep a copy of the lists every following step starts from (what that step's IP.Set sees), for
 * orc_interface_energy(c, ORC_IP_PREVIOUS, ...). */
int orc_keep_step_start(orc_ctx* c, int32_t on);
/* State the interface energies of InterfaceProperties come from: no IP.Set yet, the current
 * lists, or the lists the last step started from (IP.Set -> step -> energy). */
enum orc_ip_state { ORC_IP_NONE = 0, ORC_IP_CURRENT = 1, ORC_IP_PREVIOUS = 2 };
/* Sum over all cells of DoubleObstacle::PointEnergy (src/DoubleObstacle.cpp:903-923), in storage
 * order: DoubleObstacle::AverageEnergyDensity = sum / cells (:989-1005), Energy = sum * cell
 * volume (:953-967). */
int orc_interface_energy(const orc_ctx* c, int32_t ip_state, double* sum);
/* PhaseField::CalculateFractions (src/PhaseField.cpp:411-423): out[phase * cells + cell]. */
int orc_phase_fractions(const orc_ctx* c, double* out);
/* Longest list seen after the neighbour merge (incl. zero-valued entries) and after finalize. */
void orc_list_stats(const orc_ctx* c, int32_t* max_merged, int32_t* max_finalized);

#ifdef __cplusplus
}
#endif
#endif
