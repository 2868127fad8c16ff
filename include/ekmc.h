/*
 * ekmc.h -- C ABI of the MI355X electron-kinetics Monte Carlo engine.
 *
 * The reference (LoKI-MC v1.1.0) has no FFI: its Monte Carlo solver is the C++
 * class BoltzmannMC, constructed from the setup (BoltzmannMC.h:244-380), driven
 * by solve() once per job (BoltzmannMC.C:273-297, called from lokimc.C:31), and
 * read back through ~50 public members by Output<T>/GUI<T> (Output.h:117-149,
 * 784-823; BoltzmannMC.h:35-238).  This header is the seam that replaces that
 * class: a host program (the reference's Setup<T>, or any other driver) fills
 * the plain structs below from its parsed setup and LXCat data, calls
 * ekmc_create + ekmc_solve, and reads the results with the ekmc_get_* calls.
 *
 * Conventions
 *  - Every entry point returns an ekmc_status; 0 is success.  The reference's
 *    fatal Message::error (Message.C:6-16: errorLog.txt + exit(1)) becomes a
 *    distinct non-zero code plus a message from ekmc_last_error().
 *  - All pointers are caller-owned host memory, read or written only during
 *    the call.  No pointer crosses the ABI after return; an engine handle owns
 *    all device memory until ekmc_destroy.
 *  - One engine runs on one HIP stream; calls on one handle must not overlap.
 *  - Units are SI except energies (eV), E/N (Td) and B/N (Hx), as in the
 *    reference's setup files.
 */
#ifndef EKMC_H
#define EKMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EKMC_ABI_VERSION 2

/* ----------------------------------------------------------------------------
 * Status codes (replace Message::error, Message.C:6-16)
 * ------------------------------------------------------------------------- */
typedef enum {
    EKMC_OK = 0,
    EKMC_ERR_INVALID_ARGUMENT = 1,   /* bad input; Setup.h:576-946 checks */
    EKMC_ERR_HIP = 2,                /* a HIP runtime call failed */
    EKMC_ERR_ENERGY_OUT_OF_TABLE = 3,/* electron energy beyond the elastic
                                        cross-section range, BoltzmannMC.C:1426-1430 */
    EKMC_ERR_ZERO_VELOCITY = 4,      /* cart2sph of a null vector, MathFunctions.C:150-152 */
    EKMC_ERR_CAPACITY = 5,           /* newborn-electron buffer exhausted */
    EKMC_ERR_NO_DEVICE = 6,          /* no gfx950 device visible */
    EKMC_ERR_STATE = 7,              /* call out of order (e.g. results before solve) */
    EKMC_ERR_STATISTICS = 8          /* fewer integration points than the 50 error
                                        bins, MathFunctions.C:219-221 */
} ekmc_status;

/* Process types, GeneralDefinitions.h:33-36 */
enum { EKMC_CONSERVATIVE = 0, EKMC_IONIZATION = 1, EKMC_ATTACHMENT = 2 };

/* Special process ids written per electron, GeneralDefinitions.h:38-40 and
 * Constant.h:22 (NON_DEF marks an electron already at the synchronization time) */
enum { EKMC_NULL_COLLISION = -1, EKMC_PARTIAL_FLIGHT = -2, EKMC_NON_DEF = -123456789 };

/* Angular scattering models, AngularScatteringFunctions.h:27-74,77-114 */
enum {
    EKMC_SCAT_ISOTROPIC = 0,
    EKMC_SCAT_FORWARD = 1,
    EKMC_SCAT_BORN_DIPOLE = 2,
    EKMC_SCAT_SURENDRA = 3,
    EKMC_SCAT_COULOMB_SCREEN = 4,                 /* params: {option, screening energy [eV]} */
    EKMC_SCAT_MOMENTUM_CONSERVATION_IONIZATION = 5/* Boeuf coplanar, BoltzmannMC.C:1228-1241 */
};

/* Energy sharing in ionization, GeneralDefinitions.h:42-46 */
enum {
    EKMC_IONIZ_EQUAL_SHARING = 0,
    EKMC_IONIZ_ONE_TAKES_ALL = 1,
    EKMC_IONIZ_USING_SDCS = 2,
    EKMC_IONIZ_RANDOM_UNIFORM = 3
};

/* gasTemperatureEffect, GeneralDefinitions.h:48-51 */
enum { EKMC_GASTEMP_FALSE = 0, EKMC_GASTEMP_TRUE = 1, EKMC_GASTEMP_SMART = 2 };

/* Arithmetic of the flight/collision kernels */
enum {
    EKMC_PRECISION_FP64 = 0,  /* plain FP64: the drop-in form */
    EKMC_PRECISION_MIXED = 1  /* FP32 flight/scatter/tables; FP64 clock, positions,
                                 post-threshold energy and all reductions */
};

/* ----------------------------------------------------------------------------
 * Gas mixture: the flattened process list of BoltzmannMC.C:29-271.
 * Inelastic and superelastic directions are separate processes; a superelastic
 * entry immediately follows its inelastic partner (BoltzmannMC.C:209-266) and
 * has no table of its own (Klein-Rosseland, BoltzmannMC.C:584-595).
 * "Effective" collisions must not be passed (BoltzmannMC.C:39-41).
 * All arrays have n_processes entries unless stated.
 * ------------------------------------------------------------------------- */
typedef struct {
    int32_t n_processes;
    int32_t n_gases;
    const int32_t* type;              /* EKMC_CONSERVATIVE / _IONIZATION / _ATTACHMENT */
    const int32_t* is_elastic;        /* 1 for "Elastic" processes */
    const int32_t* is_superelastic;   /* 1 for the reverse direction */
    const int32_t* gas_index;         /* target gas; processes of one gas contiguous */
    const double* threshold;          /* collision->threshold [eV] (positive for both
                                         directions; the superelastic loss is -threshold) */
    const double* target_mass;        /* [kg] */
    const double* rel_density;        /* target-state density / gas density, i.e.
                                         collision->target->density (BoltzmannMC.C:446-451) */
    const double* stat_weight_ratio;  /* g_target/g_product for superelastic entries, else 0 */
    const double* w_parameter;        /* OPB parameter, or the threshold (BoltzmannMC.C:173-183) */
    const int32_t* scattering;        /* EKMC_SCAT_* */
    const double* scattering_params;  /* [2*n_processes] */
    /* raw integral cross sections as read from LXCat (untrimmed): process p
     * owns points [table_offset[p], table_offset[p+1]); superelastic entries
     * own zero points.  table_offset has n_processes+1 entries. */
    const int32_t* table_offset;
    const double* table_energy;       /* [eV], ascending */
    const double* table_sigma;        /* [m^2] */
    const double* gas_fraction;       /* [n_gases] */
} ekmc_mixture;

/* One working condition (WorkingConditions.h:56-142). */
typedef struct {
    double reduced_elec_field;      /* E/N [Td] */
    double elec_field_angle;        /* angle of E to z [deg] */
    double excitation_frequency;    /* f [Hz]; 0 is DC */
    double reduced_mag_field;       /* B/N [Hx]; B is along z */
    double gas_temperature;         /* [K] */
    double gas_density;             /* [m^-3] = p/(k T) */
    int64_t n_electrons;            /* numericsMC.nElectrons for this condition */
} ekmc_condition;

/* numericsMC keys (BoltzmannMC.h:262-376) plus the engine's own knobs.  A
 * field the setup does not define takes the reference default given here by
 * ekmc_numerics_default(). */
typedef struct {
    int32_t precision;                       /* EKMC_PRECISION_* */
    int32_t gas_temperature_effect;          /* EKMC_GASTEMP_* */
    int32_t ionization_operator;             /* EKMC_IONIZ_* */
    int32_t n_interp_points;                 /* 1e4 */
    double initial_elec_temp_over_gas_temp;  /* 0.01 */
    double min_collisions_before_ss;         /* x nElectrons, 0 */
    double max_collisions_before_ss;         /* x nElectrons, 1e100 */
    double max_collisions_after_ss;          /* x nElectrons, 1e100 */
    int32_t n_energy_cells;                  /* 1000 */
    int32_t n_cos_angle_cells;               /* 100 */
    int32_t n_radial_velocity_cells;         /* 200 */
    int32_t n_axial_velocity_cells;          /* 200 */
    int32_t errors_to_be_checked;            /* relError block present */
    int32_t n_integration_phases;            /* 100 */
    double rel_error_mean_energy;            /* 1e100 each */
    double rel_error_flux_drift_velocity;
    double rel_error_bulk_drift_velocity;
    double rel_error_flux_diffusion_coeffs;
    double rel_error_bulk_diffusion_coeffs;
    double rel_error_power_balance;
    double n_integration_points;             /* 200 */
    double n_integrated_ss_times;            /* 0 */
    double integrated_absolute_time;         /* 0 */
    double sync_time_x_max_coll_freq;        /* 1 */
    int32_t sync_over_sampling;              /* 1 */
    int32_t max_sync_points;                 /* safety cap on sampling points, 0 = none */
    uint64_t seed;                           /* Philox key */
    int32_t work_chunk;                      /* electrons per wavefront work unit (at most
                                                65535); 0: 1024 for batches above 3.1 M
                                                electrons, else the batch over the resident
                                                waves, at least 256 */
    int32_t park_lanes;                      /* lanes of a wavefront that wait for a real
                                                collision before it runs the collision step
                                                for all of them (1-64; 0: 32) */
    int32_t env_coarsen;                     /* the null-test envelope of the total collision
                                                frequency, staged in LDS, uses bins of
                                                2^env_coarsen times the fewest rows that fit
                                                256 bins (0-3; 0: finest) */
    int32_t evdf3d_cells;                    /* 0: none; 1-256: also count the full 3D velocity
                                                distribution of every condition from its steady
                                                state on, on evdf3d_cells^3 cells over
                                                [-v_max, v_max]^3 per axis (v_max: the speed
                                                at the top of the EEDF grid, as the 2D EVDF).
                                                An extension: LoKI-MC has the 2D EVDF of
                                                cylindrically symmetric cases only */
} ekmc_numerics;

/* Time-averaged results of one condition (BoltzmannMC.C:1484-1741,1950-2266),
 * in the simulation frame (z along B) unless "rotated" (E along z). */
typedef struct {
    double mean_energy, mean_energy_error;                       /* [eV] */
    double flux_drift_velocity[3], flux_drift_velocity_error[3]; /* [m/s] */
    double flux_diffusion[9], flux_diffusion_error[9];           /* [m^2/s] row-major */
    double bulk_drift_velocity[3], bulk_drift_velocity_error[3];
    double bulk_diffusion[9], bulk_diffusion_error[9];
    double rotated_flux_drift_velocity[3], rotated_flux_drift_velocity_error[3];
    double rotated_bulk_drift_velocity[3], rotated_bulk_drift_velocity_error[3];
    double rotated_flux_diffusion[9], rotated_flux_diffusion_error[9];
    double rotated_bulk_diffusion[9], rotated_bulk_diffusion_error[9];
    double power_field, power_growth;                            /* [eV m^3/s] per electron */
    double power_balance_rel_error;
    double total_ion_rate_coeff, total_att_rate_coeff;           /* [m^3/s] (MC counts) */
    double flux_red_mobility, bulk_red_mobility;                 /* [1/(m V s)] */
    double flux_red_transv_diff, flux_red_long_diff;             /* [1/(m s)] */
    double bulk_red_transv_diff, bulk_red_long_diff;
    double steady_state_time, final_time, total_integrated_time; /* [s] */
    double trial_collision_frequency;                            /* [1/s] */
    double max_elec_energy, max_eedf_energy;                     /* [eV] */
    double elapsed_time;                                         /* wall [s] of solve() */
    uint64_t real_collisions, null_collisions;                   /* whole run */
    uint64_t real_collisions_at_ss, null_collisions_at_ss;
    uint64_t escape_events;   /* flights that outran the trial frequency bound */
    int32_t n_sampling_points, n_integration_points;
    int32_t good_statistical_errors;
    int32_t stopped_by_max_collisions;
} ekmc_results;

/* Per-kernel timing collected with HIP events on the engine's stream. */
typedef struct {
    char name[48];
    uint64_t launches;
    double total_ms;
} ekmc_kernel_stat;

typedef struct ekmc_engine ekmc_engine;

/* ----------------------------------------------------------------------------
 * Engine life cycle (replaces BoltzmannMC's constructor, solve() and members)
 * ------------------------------------------------------------------------- */

/* Reference defaults of numericsMC (BoltzmannMC.h:52-80). */
void ekmc_numerics_default(ekmc_numerics* out);

/* Build an engine for n_cond working conditions sharing one mixture.  With
 * n_cond == 1 and EKMC_PRECISION_FP64 this is the plain one-condition form
 * (BoltzmannMC.h:244-380 + allocateEvaluateVariablesFirstTime, BoltzmannMC.C:29-271);
 * with n_cond > 1 every condition advances in the same batched launches. */
ekmc_status ekmc_create(const ekmc_mixture* mixture, const ekmc_condition* conditions,
                        int32_t n_cond, const ekmc_numerics* numerics, ekmc_engine** out);
void ekmc_destroy(ekmc_engine* engine);

/* Run the whole Monte Carlo method for every condition: initial ensemble,
 * flights and collisions until every stop criterion holds, time averages
 * (BoltzmannMC::solve, BoltzmannMC.C:273-426). */
ekmc_status ekmc_solve(ekmc_engine* engine);

/* Fine-grained driving (what solve() does, one step at a time). */
ekmc_status ekmc_initialize(ekmc_engine* engine);            /* BoltzmannMC.C:428-559 + first sample */
/* Advance every still-running condition by n_sync synchronization intervals,
 * sampling and running the control logic of BoltzmannMC.C:320-384 after each;
 * *n_running receives the number of conditions not yet finished. */
ekmc_status ekmc_step(ekmc_engine* engine, int32_t n_sync, int32_t* n_running);
ekmc_status ekmc_finalize(ekmc_engine* engine);              /* BoltzmannMC.C:386-425 */

/* Work-unit autotuner.  Times short runs of the given batch, one per
 * candidate work unit of the flight kernel (main electrons per wavefront work
 * unit: candidates / 64 electrons per lane), and returns the candidate with the
 * most collision events per second.  Each trial creates its own engine with
 * `numerics` except for work_chunk and the stop criteria (every condition runs
 * through warmup + steps intervals), runs `warmup` untimed and `steps` timed
 * intervals, and destroys it; one trial's device memory is held at a time.
 * candidates NULL or n_candidates 0: {256, 512, 768, 1024, 1536, 2048}.
 * rates, NULL or one double per candidate (6 with the default candidates),
 * receives every candidate's events/s.
 * The work unit changes the keys of the population-control draws, so a run is
 * reproducible for a given work_chunk, not across work units.  No counterpart
 * in the reference: its OpenMP loop over electrons (BoltzmannMC.C:636) has no
 * decomposition to tune. */
ekmc_status ekmc_autotune(const ekmc_mixture* mixture, const ekmc_condition* conditions, int32_t n_cond,
                          const ekmc_numerics* numerics, const int32_t* candidates, int32_t n_candidates,
                          int32_t warmup, int32_t steps, int32_t* best_work_chunk, double* rates);

/* Launch autotuner: the work unit as ekmc_autotune does, then, at that work
 * unit, the collision-step lane threshold (park_lanes: 16, 24, 32, 40, 48),
 * then the envelope coarseness (env_coarsen: 0, 1, 2), each trial a short run
 * of the batch as in ekmc_autotune.  *best receives `numerics` with the
 * fastest work_chunk, park_lanes and env_coarsen; *best_rate (may be NULL)
 * its collision events per second.  No counterpart in the reference. */
ekmc_status ekmc_autotune_launch(const ekmc_mixture* mixture, const ekmc_condition* conditions, int32_t n_cond,
                                 const ekmc_numerics* numerics, int32_t warmup, int32_t steps,
                                 ekmc_numerics* best, double* best_rate);

/* ----------------------------------------------------------------------------
 * Results (the members Output<T> reads)
 * ------------------------------------------------------------------------- */
ekmc_status ekmc_get_results(const ekmc_engine* engine, int32_t cond, ekmc_results* out);
/* Per-process arrays of length n_processes: MC rate coefficients [m^3/s]
 * (BoltzmannMC.C:1625-1636), collision counts after the steady state, and
 * power gains/losses [eV m^3/s] (BoltzmannMC.C:1638-1649). Any pointer may be NULL. */
ekmc_status ekmc_get_process_data(const ekmc_engine* engine, int32_t cond, double* rate_coeffs,
                                  uint64_t* counts, double* power_gain, double* power_loss);
/* EEDF on n_energy_cells cells: cell energies [eV] and eedf [eV^-3/2]
 * (BoltzmannMC.C:1574-1581). */
ekmc_status ekmc_get_eedf(const ekmc_engine* engine, int32_t cond, double* energy_cells, double* eedf);
/* AC only: eedf per phase, [n_integration_phases * n_energy_cells], and the
 * per-phase mean energy / flux velocity (BoltzmannMC.C:1725-1741). */
ekmc_status ekmc_get_eedf_periodic(const ekmc_engine* engine, int32_t cond, double* eedf_periodic,
                                   double* mean_energy_periodic, double* flux_velocity_periodic);
/* AC only: phase-bin centres [n_integration_phases] (BoltzmannMC.C:525-526)
 * and the per-phase averages of the sampled swarm quantities after the steady
 * state (BoltzmannMC.C:1466-1481,1728-1735): mean energy [P], flux and bulk
 * drift velocity [3P], flux and bulk diffusion coefficients [9P, row-major,
 * m^2/s] -- what Output::saveMCTemporalInfoPeriodic writes (Output.h:754-782).
 * Any pointer may be NULL. */
ekmc_status ekmc_get_swarm_periodic(const ekmc_engine* engine, int32_t cond, double* phases, double* mean_energy,
                                    double* flux_velocity, double* bulk_velocity, double* flux_diffusion,
                                    double* bulk_diffusion);
/* With evdf3d_cells > 0: the time-averaged 3D EVDF of a condition,
 * [evdf3d_cells^3] row-major over (v_x, v_y, v_z) in the lab frame, normalised
 * so that its sum times the cell volume is 1 (velocities outside the cube are
 * not counted), and the cell centres of one axis [evdf3d_cells] in m/s.
 * Integer counts per sample, so the sums are exact.  No counterpart in the
 * reference (its EVDF is the 2D one of cylindrically symmetric cases,
 * BoltzmannMC.C:1562-1571).  Either pointer may be NULL. */
ekmc_status ekmc_get_evdf3d(const ekmc_engine* engine, int32_t cond, double* axis_cells, double* evdf3d);
/* Cylindrically symmetric cases only: EVDF [n_radial * n_axial] row-major with
 * radial/axial cell centres, and the first/second anisotropies per energy cell
 * (BoltzmannMC.C:1583-1605). */
ekmc_status ekmc_get_evdf(const ekmc_engine* engine, int32_t cond, double* radial_cells,
                          double* axial_cells, double* evdf, double* efadf, double* esadf);
/* Sampled series (one row per sampling point, BoltzmannMC.C:1410-1482):
 * times[n], mean_energy[n], mean_position[3n], mean_velocity[3n],
 * flux_diffusion[9n], bulk_velocity[3n], bulk_diffusion[9n].  *n_points gets
 * the number of rows; pass capacity in it.  Any array may be NULL. */
ekmc_status ekmc_get_series(const ekmc_engine* engine, int32_t cond, int32_t* n_points,
                            double* times, double* mean_energy, double* mean_position,
                            double* mean_velocity, double* flux_diffusion,
                            double* bulk_velocity, double* bulk_diffusion);
/* Position covariance per sampling point, cov[9n] row-major
 * (positionCovariances, BoltzmannMC.C:1437-1454; written to MCTemporalInfo.txt
 * by Output.h:730-752).  *n_points as for ekmc_get_series. */
ekmc_status ekmc_get_position_covariance(const ekmc_engine* engine, int32_t cond, int32_t* n_points,
                                         double* cov);
/* Copy the current electron ensemble of a condition to host (SoA, FP64). */
ekmc_status ekmc_get_electrons(const ekmc_engine* engine, int32_t cond, double* pos /*3N*/,
                               double* vel /*3N*/, double* energy /*N*/, double* time /*N*/);

/* Running status of one condition, the quantities BoltzmannMC::dispInfo
 * prints (BoltzmannMC.C:1895-1948): simulated time, trial frequency, sampling
 * counters, collision counts and the last sampled mean energy. */
typedef struct {
    double time;                  /* [s] */
    double trial_collision_frequency;
    double steady_state_time;     /* NON_DEF (-123456789) before the steady state */
    double last_mean_energy;      /* [eV] */
    double max_elec_energy;       /* [eV] */
    double table_max_energy;      /* upper end of the current collision tables [eV] */
    uint64_t real_collisions, null_collisions;
    int32_t n_sampling_points, n_integration_points;
    int32_t finished;
    int32_t pad;
    uint64_t escape_events;       /* trial events whose trial frequency was below the total
                                     collision frequency at the electron's energy (0 when the
                                     bound of BoltzmannMC.C:716-802 held) */
} ekmc_cond_status;
ekmc_status ekmc_get_status(const ekmc_engine* engine, int32_t cond, ekmc_cond_status* out);

/* Per-kernel HIP-event timings since creation (or the last reset): the flight
 * kernels by default, every kernel when the environment variable
 * EKMC_KERNEL_TIMING is "all", none when it is "0"; followed
 * by host-side phases (names starting with "host_"): preparing and launching
 * an interval, waiting for it, the per-condition control logic, table
 * rebuilds and histogram downloads. */
ekmc_status ekmc_get_kernel_stats(const ekmc_engine* engine, ekmc_kernel_stat* stats,
                                  int32_t* n_stats);
ekmc_status ekmc_reset_kernel_stats(ekmc_engine* engine);
/* Total flight/collision tests (real + null) over all conditions so far. */
ekmc_status ekmc_get_event_count(const ekmc_engine* engine, uint64_t* events, uint64_t* real);

const char* ekmc_last_error(void);

/* ----------------------------------------------------------------------------
 * Batch operators: the per-electron device functions of the engine applied to
 * caller-supplied electrons and uniform variates.  They exist so a host can
 * check the device arithmetic against any CPU restatement event by event.
 * Each runs the same __device__ code the engine's flight kernel runs.
 * ------------------------------------------------------------------------- */

/* Number of uniform variates per collision event and their slot meaning:
 *  0 free time U(0,1)          1-4 thermal target versor, 4 x U(0,1]
 *  5 channel choice U(0,1]     6 scattered cos(chi) sampler U[0,1]
 *  7 scattered azimuth U(0,1]  8 ionization energy share U[0,1]
 *  9 ejected cos(chi) U[0,1]   10 ejected azimuth U(0,1]
 * (draw order of BoltzmannMC.C:652-674,907-1272 and MathFunctions.C:54-59).
 * In the engine, slot s of event k of an electron with stream id S comes from
 * Philox4x32-10 with counter {S lo, S hi, k, call} and key = seed:
 *   FP64: 53 bits of words (0,1) or (2,3); call 0 {0,5}, 1 {6,7}, 2 {1,2},
 *         3 {3,4}, 4 {8,9}, 5 {10}
 *   FP32: the same calls, top 24 bits of word 0 or 2: the 24 leading bits
 *         of the FP64 variate */
#define EKMC_VARIATES_PER_EVENT 11

/* Collision tables of one condition (interpolateCrossSections,
 * BoltzmannMC.C:561-615) for a given maximum energy, built on the host in
 * FP64: interp and cumul are [n_interp x n_processes] row-major.  Needs no
 * device. */
ekmc_status ekmc_build_tables(const ekmc_mixture* mixture, double gas_density,
                              int32_t n_interp, double max_energy,
                              double* interp, double* cumul, double* total_freq,
                              double* max_freq, double* energy_step);

/* The same tables as the engine builds them on the device for condition
 * `cond` (in the engine's precision, returned as FP64): rebuilds that
 * condition's tables at max_energy and copies them back.  Any output pointer
 * may be NULL. */
ekmc_status ekmc_op_tables(ekmc_engine* engine, int32_t cond, double max_energy, double* interp,
                           double* cumul, double* total_freq, double* max_freq, double* energy_step);

/* Free flights (accelerateElectron, BoltzmannMC.C:804-905) on the device for
 * n electrons of condition `cond`: positions and velocities [3n] in/out,
 * energy [n] out, field gain [n] out. */
ekmc_status ekmc_op_accelerate(ekmc_engine* engine, int32_t cond, int64_t n, const double* time,
                               const double* dt, double* pos, double* vel, double* energy,
                               double* gain);

/* Most events one call of ekmc_op_collide_noalloc takes. */
#define EKMC_NOALLOC_MAX_EVENTS 32768

/* Collisions (performCollision and its callees, BoltzmannMC.C:907-1280) on the
 * device with the engine's current tables of `cond`.  Inputs: vel [3n],
 * energy [n], trial frequency used for the flight [n], variates
 * [n*EKMC_VARIATES_PER_EVENT].  Outputs: process id [n] (-1 null), vel and
 * energy updated in place, ejected velocity [3n] and energy [n] (ionization),
 * energy change [n].  Tables are rebuilt to max_energy first when > 0. */
ekmc_status ekmc_op_collide(ekmc_engine* engine, int32_t cond, double max_energy, int64_t n,
                            double* vel, double* energy, const double* trial_freq,
                            const double* variates, int32_t* process, double* ejected_vel,
                            double* ejected_energy, double* energy_change);

/* One synchronization interval of every running condition, as
 * ekmc_step(engine, 1, ...) runs it, with a copy of condition `cond`'s
 * electron pool taken after the flights and before population control: the
 * passes of BoltzmannMC.C:630-681 without nonParallelCollisionTasks, i.e.
 * every main electron and every electron ejected during the interval at
 * t_sync, attached ones removed.  Every alive slot of the condition (main
 * slots, then the newborn regions), in slot order: positions and velocities
 * [3n], energies [n], clocks [n], Philox streams [n] and event counters [n].
 * *n_pool: capacity in, electrons out (EKMC_ERR_CAPACITY, with the count, when
 * the capacity is smaller).  *t_sync and *trial_freq receive the interval's
 * end and trial frequency.  The interval then completes as in ekmc_step.
 * Any output pointer may be NULL. */
ekmc_status ekmc_op_interval_pool(ekmc_engine* engine, int32_t cond, int64_t* n_pool, double* pos, double* vel,
                                  double* energy, double* time, uint64_t* stream, uint32_t* event,
                                  double* t_sync, double* trial_freq);

/* ekmc_op_collide without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same arguments, kernel and results as ekmc_op_collide, for at most
 * EKMC_NOALLOC_MAX_EVENTS events per call (more: EKMC_ERR_CAPACITY).  The
 * device staging of the events is a static array of the library, one per
 * GPU (5.4 MiB, held until the process ends); every call on a GPU reuses
 * that GPU's array, so calls that overlap on one GPU (from several threads
 * or engines) race.  Every staged value is written by the call itself, so
 * nothing carries over from an earlier call. */
ekmc_status ekmc_op_collide_noalloc(ekmc_engine* engine, int32_t cond, double max_energy, int64_t n,
                                    double* vel, double* energy, const double* trial_freq,
                                    const double* variates, int32_t* process, double* ejected_vel,
                                    double* ejected_energy, double* energy_change);

/* Philox4x32-10 as the engine uses it (counter [4n], key [2n] -> out [4n]). */
ekmc_status ekmc_op_philox(int64_t n, const uint32_t* counter, const uint32_t* key,
                           uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* EKMC_H */
