// gas_model.h -- LXCat files (Parse.C:43-295) and the electron-kinetics gas
// mixture of a setup (Setup.h:229-459: gases, states, collisions, their
// properties) flattened into the engine's process list (BoltzmannMC.C:29-271).
#pragma once
#include <string>
#include <vector>

#include "ekmc.h"
#include "setup_file.h"

namespace lokimc {

// A state as written in LXCat: gas(ionCharg,eleLevel,v=vib,J=rot)
// (Parse::getRawState, Parse.C:412-473).
struct RawState {
    std::string gas, ionCharg, eleLevel, vibLevel, rotLevel;
    bool operator==(const RawState& o) const {
        return gas == o.gas && ionCharg == o.ionCharg && eleLevel == o.eleLevel && vibLevel == o.vibLevel &&
               rotLevel == o.rotLevel;
    }
    // State name as the reference prints it (State.h:122-135).
    std::string name() const;
};
RawState parseState(const std::string& text);

struct LXCatEntry {
    std::string type;      // Elastic, Effective, Attachment, Ionization, Excitation, Vibrational, Rotational
    std::string subtype;   // momentum-transfer or integral
    bool isReverse = false;
    double threshold = 0;
    RawState target;
    std::vector<RawState> products;
    std::vector<double> productStoi;
    std::vector<double> energy, sigma;
};
// Entries of the files, in order; paths are relative to `inputDir`.
std::vector<LXCatEntry> readLXCat(const std::vector<std::string>& files, const std::string& inputDir);

// The mixture of a setup: process arrays for ekmc_mixture plus what the
// output files need (descriptions and categories).
struct GasMixture {
    std::vector<std::string> gasNames;     // gases with collisions, in process order
    std::vector<std::string> allGasNames;  // every gas of the LXCat files (targets and products)
    std::vector<double> gasMass;           // [kg]
    std::vector<double> gasFraction;
    // per process (inelastic and superelastic directions separate)
    std::vector<int32_t> type, isElastic, isSuperelastic, gasIndex, scattering, tableOffset;
    std::vector<double> threshold, targetMass, relDensity, statWeightRatio, wParameter, scatteringParams,
        tableEnergy, tableSigma;
    std::vector<std::string> lxcatType;    // collision type of the process (Excitation, ...)
    std::vector<int> collisionId;          // index of the collision among non-Effective collisions
    std::vector<std::string> description;  // Collision::description (Collision.C:88-122)
    // per collision (the rate-coefficient files list collisions, not directions)
    std::vector<int> collisionFirstProcess;
    std::vector<int> collisionIsReverse;
    std::vector<std::string> collisionDescription;
    std::vector<double> collisionTargetDensity, collisionProductDensity;
    // EEDF-derived quantities need each collision's cross sections
    std::vector<std::vector<double>> collisionEnergy, collisionSigmaIntegral;  // integral cross section
    std::vector<std::vector<double>> collisionEnergyMomTransf, collisionSigmaMomTransf;  // momentum-transfer
    std::vector<double> collisionThreshold, collisionStatWeightRatio;
    std::vector<std::string> collisionType;
    std::vector<int> collisionGas;         // slot in gasNames, -1 for a gas without flights
    // 0: a collision of the flights (first the process order's collisions);
    // 1: an Effective collision, 2: an extra one (LXCatFilesExtra), both
    // listed in the rate-coefficient files only, after the kind-0 ones
    std::vector<int> collisionKind;
    std::vector<int> collisionRefId;       // the reference's Collision::ID (creation order), printed + 1

    int nProcesses() const { return (int)type.size(); }
    // View for the C ABI (pointers into this object).
    ekmc_mixture view() const;
};

// Build the mixture of `setup` (electronKinetics: LXCatFiles(Extra),
// gasProperties, stateProperties with wildcard states and property functions,
// effectiveCrossSectionPopulations, anisotropicScattering) at the gas
// temperature `gasTemperature` [K] (the argument of population functions).
// Throws InputError for input the reference rejects.
GasMixture buildMixture(const SetupFile& setup, const std::string& inputDir, double gasTemperature);

}  // namespace lokimc
