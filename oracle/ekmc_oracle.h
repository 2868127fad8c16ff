/*
 * ekmc_oracle.h -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
 *
 * C API of the CPU FP64 restatement of LoKI-MC's Monte Carlo loop
 * (Code/LoKI-MC/Sources/BoltzmannMC.C, MathFunctions.C,
 * Headers/AngularScatteringFunctions.h).  Only tests/, the smoke check and
 * the benchmark's CPU-baseline leg may load it; the GPU engine is tested
 * against it and never falls back to it.
 *
 * The reference draws its variates from a thread-local mt19937_64 seeded by
 * std::random_device (MathFunctions.C:31-43), so it cannot be replayed.  The
 * restatement replaces that source with an injectable one: either a tape of
 * uniforms indexed by slot (EKMC_VARIATES_PER_EVENT per event), or the
 * Philox4x32-10 stream keyed by (seed, electron stream, event, slot) that the
 * GPU engine uses.  Everything else follows the reference's arithmetic.
 *
 * How each function is pinned (tests/): Philox by the Random123 known-answer
 * vectors; tables by numpy interpolation, prefix sums and the Klein-Rosseland
 * closed form; flights by ODE integration and field-work identities; channel
 * choice by a brute-force linear scan; scattering by energy/momentum
 * invariants and KS tests against the reference's angular distributions;
 * population control (orc_population_control) by the exact probabilities its
 * rules fix (last-in first-out refills, uniform clones, uniform removal
 * subsets, the growth identity); the trial-frequency energy bound by flights
 * that never exceed it; checkSteadyState by constructed series; the whole
 * loop (orc_solve) by the published Reid-ramp and Lucas-Saelee results and by
 * the sampling counts its control cadence fixes.  The linear interpolation
 * follows GSL's published gsl_interp_linear, whose installed version the
 * reference does not pin (parity unpinned at that version only).
 */
#ifndef EKMC_ORACLE_H
#define EKMC_ORACLE_H

#include <stdint.h>
#include "../include/ekmc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

/* Field/gas description of one working condition, as BoltzmannMC keeps it
 * after evaluateNonConstantVariables (BoltzmannMC.C:428-489). */
typedef struct {
    double elec_field[3];        /* [V/m], already x sqrt(2) for AC */
    double accel[3];             /* -e E / m_e */
    double omega;                /* 2 pi f [rad/s] */
    double cyclotron;            /* e B / m_e [rad/s] */
    double reduced_mag_field;    /* [Hx], only tested against 0 */
    double excitation_frequency; /* [Hz], only tested against 0 */
} orc_field;

/* Build the model (process list, BoltzmannMC.C:29-271 + Collision.C:41-44). */
int orc_model_create(const ekmc_mixture* mix, const ekmc_condition* cond,
                     const ekmc_numerics* num, orc_model** out);
void orc_model_destroy(orc_model* m);
int orc_model_nprocesses(const orc_model* m);
void orc_model_field(const orc_model* m, orc_field* out);
/* interpolateCrossSections (BoltzmannMC.C:561-615); copies the tables out when
 * the pointers are non-NULL. */
int orc_model_tables(orc_model* m, double max_energy, double* interp, double* cumul,
                     double* total_freq, double* max_freq, double* energy_step);
/* maximizationAccelerationEnergy (BoltzmannMC.C:765-802) */
double orc_max_accel_energy(const orc_model* m, double initial_energy, double dt);

/* Batch free flights, accelerateElectron (BoltzmannMC.C:804-905). */
int orc_accelerate(const orc_model* m, int64_t n, const double* time, const double* dt,
                   double* pos, double* vel, double* energy, double* gain);

/* Batch collisions, performCollision & callees (BoltzmannMC.C:907-1280), on
 * the model's current tables with tape variates [n*EKMC_VARIATES_PER_EVENT]. */
int orc_collide(const orc_model* m, int64_t n, double* vel, double* energy,
                const double* trial_freq, const double* variates, int32_t* process,
                double* ejected_vel, double* ejected_energy, double* energy_change);

/* Brute-force channel choice: the commented-out linear scan of
 * BoltzmannMC.C:923-939 written for the cold branch (cumulative sum over
 * processes at the interpolated row); returns -1 for a null collision. */
int orc_choose_linear(const orc_model* m, double energy, double speed, double rand_freq);

/* Philox4x32-10 (counter [4n], key [2n] -> out [4n]). */
void orc_philox(int64_t n, const uint32_t* counter, const uint32_t* key, uint32_t* out);
/* Uniform of slot `slot` of event `event` of electron stream `stream`, in the
 * interval the slot uses (see EKMC_VARIATES_PER_EVENT in ekmc.h). */
double orc_variate(uint64_t seed, uint64_t stream, uint32_t event, int slot);

/* Ensemble advance without population control: n electrons (SoA FP64) with
 * streams and event counters run the reference's passes (BoltzmannMC.C:630-681)
 * until every electron and every ejected electron reaches t_sync at a fixed
 * trial frequency; ejected electrons are appended after the n inputs (up to
 * capacity) and attached electrons get alive = 0.  Returns the pool size. */
int64_t orc_advance_pool(const orc_model* m, uint64_t seed, double trial_freq, double t_sync,
                         int64_t n, int64_t capacity, double* pos, double* vel, double* energy,
                         double* time, double* free_time, double* nu_e, uint64_t* stream,
                         uint32_t* event, int32_t* alive, uint64_t* counters /*[P]*/,
                         uint64_t* n_null);

/* Whole reference loop (BoltzmannMC::solve) on the CPU with n_threads OpenMP
 * threads; fills the same results struct as the engine. */
int orc_solve(orc_model* m, int n_threads, ekmc_results* out, double* rate_coeffs,
              double* eedf_cells, double* eedf);
/* The reference loop for a fixed number of synchronization intervals (each
 * with its passes, serial population control and ensemble sampling) from a
 * fresh ensemble, timed around the interval loop only.  Used as the CPU
 * baseline of the benchmark. */
int orc_run_intervals(orc_model* m, int n_threads, int n_sync, uint64_t* events, uint64_t* real,
                      double* seconds);
/* Sampled series after orc_solve (same layout as ekmc_get_series). */
int orc_series(const orc_model* m, int32_t* n_points, double* times, double* mean_energy,
               double* mean_velocity);

/* nonParallelCollisionTasks (BoltzmannMC.C:1282-1408) on a caller-given
 * ensemble of n electrons after one pass: pid[n] the pass outcome of each
 * electron (process id, EKMC_NULL_COLLISION, EKMC_PARTIAL_FLIGHT or
 * EKMC_NON_DEF), dE[n] its energy change, ej_pos[3n]/ej_energy[n] the electron
 * ejected by electron i when pid[i] is an ionization.  pos[3n], energy[n] and
 * time[n] are updated to the ensemble after attachment refills, clones and
 * removals; the serial draws come from the sequential Philox stream of
 * pop_seed.  *growth receives energyGrowth, counters[P] the collision counts. */
int orc_population_control(orc_model* m, int64_t n, const int32_t* pid, const double* dE, double* pos,
                           double* energy, double* time, const double* ej_pos, const double* ej_energy,
                           uint64_t pop_seed, double* growth, uint64_t* counters);
/* checkSteadyState (BoltzmannMC.C:1787-1893) on a sampled series of n points
 * (sampling times and mean energies) at simulated time `time` with
 * `total_collisions` real collisions so far; returns 1 when the steady state
 * is declared, 0 otherwise, -1 for an empty series or an AC model. */
int orc_check_steady_state(orc_model* m, int32_t n, const double* times, const double* mean_energy,
                           double time, uint64_t total_collisions);

/* Pieces of the reference's sampling/control arithmetic for unit tests. */
double orc_statistical_error(const double* x, int64_t n, double n_bins); /* MathFunctions.C:211-231 */
void orc_histogram(const double* x, int64_t n, const double* nodes, int32_t n_nodes,
                   double* counts);                                    /* MathFunctions.C:61-81 */

#ifdef __cplusplus
}
#endif
#endif
