// outputs.h -- the output files of a LoKI-MC run (Output.h:180-823) from
// the engine's results, with the host FP64 post-processing they need
// (evaluateSwarmParameters / evaluateRateCoeff / evaluatePower,
// BoltzmannMC.C:1950-2384; Collision.C:291-397).
#pragma once
#include <string>
#include <vector>

#include "ekmc.h"
#include "gas_model.h"

namespace lokimc {

// Everything one job's output files need, pulled from the engine.
struct JobData {
    ekmc_condition cond{};
    ekmc_results res{};
    std::vector<double> rateMC, powerGain, powerLoss;  // per process
    std::vector<uint64_t> counts;
    std::vector<double> energyCells, eedf;             // EEDF cells
    bool cylindrical = false;
    std::vector<double> radialCells, axialCells, evdf, efadf, esadf;
    int nSeries = 0;
    std::vector<double> times, meanEnergy, meanPosition, meanVelocity, fluxDiffusion, bulkVelocity, bulkDiffusion,
        positionCovariance;
    double nElectrons = 0;
    // AC conditions: phase-bin centres, EEDF per phase [phase][cell], and the
    // per-phase averages of the sampled swarm quantities
    int nPhases = 0;
    std::vector<double> phases, eedfPeriodic, meanEnergyPeriodic, fluxVelocityPeriodic, bulkVelocityPeriodic,
        fluxDiffusionPeriodic, bulkDiffusionPeriodic;
};

// Pull a condition's results out of a finished engine.  `allAt180`: every
// job of the scan has the field angle 180 degrees (the reference writes the
// velocity and anisotropy distributions only then, WorkingConditions.h:164-172).
JobData collectJob(const ekmc_engine* eng, int cond, const ekmc_condition& c, const ekmc_numerics& num, int nProcesses,
                   bool allAt180);

// Swarm parameters keyed as the reference's swarmParam map.
struct SwarmParams {
    std::vector<std::pair<std::string, double>> items;
    double get(const std::string& k) const;
    void set(const std::string& k, double v);
};
SwarmParams swarmParameters(const JobData& j, const GasMixture& m);

// EEDF-convolution rate coefficients per collision (inelastic, superelastic
// or NaN when the collision is not reversible), for the job's EEDF or, with
// `eedf`, for that EEDF on the job's cells.
void eedfRateCoefficients(const JobData& j, const GasMixture& m, std::vector<double>& ine, std::vector<double>& sup,
                          const double* eedf = nullptr);

// Output files requested by `dataFiles` (output.dataFiles) into `folder`
// (the job's sub-folder when there are several jobs).
struct OutputOptions {
    std::vector<std::string> dataFiles;
    std::string variableCondition;  // reducedElecField, reducedMagField, elecFieldAngle, excitationFrequency
};
void writeJobFiles(const std::string& folder, const JobData& j, const GasMixture& m, const OutputOptions& o);
// Look-up tables over all jobs (lookUpTableSwarm/Power/RateCoeff.txt).
void writeLookUpTables(const std::string& folder, const std::vector<JobData>& jobs, const GasMixture& m,
                       const OutputOptions& o);

}  // namespace lokimc
