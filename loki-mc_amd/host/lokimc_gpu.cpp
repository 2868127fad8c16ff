// lokimc_gpu.cpp -- command-line driver with the reference's inputs and
// outputs (lokimc.C:18-93): reads a LoKI-MC setup file and its LXCat files,
// runs every job of the working-condition scan as one batched engine on the
// GPU, and writes the output files of Output.h into Output/<folder>.
//
//   lokimc_gpu <setup file> [--input DIR] [--output DIR] [--precision fp64|mixed]
//              [--seed N] [--autotune] [--dry-run]
//
// --input is the folder the setup's relative paths start from (the
// reference's Input/), --output the folder holding the output folder (the
// reference's Output/).  --autotune times short runs of the scan per
// candidate work unit (ekmc_autotune) and runs the scan with the fastest;
// without it the engine's default work unit is used.  --dry-run parses
// everything and prints the jobs and the process list as JSON without
// touching the GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "ekmc.h"
#include "gas_model.h"
#include "outputs.h"
#include "setup_file.h"

using namespace lokimc;

namespace {

constexpr double kBoltzmann = 1.38064852e-23;  // Constant.h:6-23

struct Jobs {
    std::vector<ekmc_condition> conds;
    std::string variable;  // the scanned condition
};

// WorkingConditions.h:56-142 for eedfType boltzmannMC.
Jobs workingConditions(const SetupFile& s, int64_t nElectrons) {
    const double p = evalNumber(s.value("workingConditions.gasPressure"));
    const double T = evalNumber(s.value("workingConditions.gasTemperature"));
    auto vec = [&](const char* key) {
        const std::string v = s.value(std::string("workingConditions.") + key);
        if (v.empty())
            throw InputError("Error in the configuration of the working conditions. When choosing 'boltzmannMC' eedfType, the 'reducedElecField', 'reducedMagField', 'elecFieldAngle' and 'excitationFrequency' must be defined.");
        return evalVector(v);
    };
    const std::vector<double> en = vec("reducedElecField"), bn = vec("reducedMagField"), an = vec("elecFieldAngle"),
                              fr = vec("excitationFrequency");
    int multi = (en.size() > 1) + (bn.size() > 1) + (an.size() > 1) + (fr.size() > 1);
    if (multi > 1)
        throw InputError("Error in the configuration of the working conditions. Only one of 'reducedElecField', 'reducedMagField', 'elecFieldAngle' and 'excitationFrequency' may have multiple values.");
    if (p < 0 || T < 0) throw InputError("Error in the configuration of the working conditions. The working conditions must be all non-negative.");
    Jobs j;
    j.variable = bn.size() > 1 ? "reducedMagField" : an.size() > 1 ? "elecFieldAngle" : fr.size() > 1 ? "excitationFrequency" : "reducedElecField";
    const size_t n = std::max(std::max(en.size(), bn.size()), std::max(an.size(), fr.size()));
    for (size_t k = 0; k < n; ++k) {
        ekmc_condition c{};
        c.reduced_elec_field = en.size() > 1 ? en[k] : en[0];
        c.reduced_mag_field = bn.size() > 1 ? bn[k] : bn[0];
        c.elec_field_angle = an.size() > 1 ? an[k] : an[0];
        c.excitation_frequency = fr.size() > 1 ? fr[k] : fr[0];
        c.gas_temperature = T;
        c.gas_density = p / (kBoltzmann * T);
        c.n_electrons = nElectrons;
        j.conds.push_back(c);
    }
    return j;
}

// numericsMC keys (BoltzmannMC.h:262-376).
ekmc_numerics numerics(const SetupFile& s) {
    ekmc_numerics n;
    ekmc_numerics_default(&n);
    const std::string io = s.value("electronKinetics.ionizationOperatorType");
    if (io == "equalSharing") n.ionization_operator = EKMC_IONIZ_EQUAL_SHARING;
    else if (io == "oneTakesAll") n.ionization_operator = EKMC_IONIZ_ONE_TAKES_ALL;
    else if (io == "usingSDCS") n.ionization_operator = EKMC_IONIZ_USING_SDCS;
    else if (io == "randomUniform") n.ionization_operator = EKMC_IONIZ_RANDOM_UNIFORM;
    else throw InputError("Unknown ionizationOperatorType '" + io + "'.");
    const std::string k = "electronKinetics.numericsMC.";
    const std::string gt = s.value(k + "gasTemperatureEffect");
    if (gt == "false") n.gas_temperature_effect = EKMC_GASTEMP_FALSE;
    else if (gt == "true") n.gas_temperature_effect = EKMC_GASTEMP_TRUE;
    else if (gt == "smartActivation") n.gas_temperature_effect = EKMC_GASTEMP_SMART;
    else throw InputError("Error: 'gasTemperatureEffect' must be 'false', 'true' or 'smartActivation'.");
    n.initial_elec_temp_over_gas_temp = s.number(k + "initialElecTempOverGasTemp", n.initial_elec_temp_over_gas_temp);
    n.min_collisions_before_ss = s.number(k + "minCollisionsBeforeSteadyState", n.min_collisions_before_ss);
    n.max_collisions_before_ss = s.number(k + "maxCollisionsBeforeSteadyState", n.max_collisions_before_ss);
    n.max_collisions_after_ss = s.number(k + "maxCollisionsAfterSteadyState", n.max_collisions_after_ss);
    n.n_energy_cells = (int32_t)s.number(k + "nEnergyCells", n.n_energy_cells);
    n.n_cos_angle_cells = (int32_t)s.number(k + "nCosAngleCells", n.n_cos_angle_cells);
    n.n_radial_velocity_cells = (int32_t)s.number(k + "nRadialVelocityCells", n.n_radial_velocity_cells);
    n.n_axial_velocity_cells = (int32_t)s.number(k + "nAxialVelocityCells", n.n_axial_velocity_cells);
    n.errors_to_be_checked = s.has(k + "relError") ? 1 : 0;
    n.rel_error_mean_energy = s.number(k + "relError.meanEnergy", n.rel_error_mean_energy);
    n.rel_error_flux_drift_velocity = s.number(k + "relError.fluxDriftVelocity", n.rel_error_flux_drift_velocity);
    n.rel_error_bulk_drift_velocity = s.number(k + "relError.bulkDriftVelocity", n.rel_error_bulk_drift_velocity);
    n.rel_error_flux_diffusion_coeffs = s.number(k + "relError.fluxDiffusionCoeffs", n.rel_error_flux_diffusion_coeffs);
    n.rel_error_bulk_diffusion_coeffs = s.number(k + "relError.bulkDiffusionCoeffs", n.rel_error_bulk_diffusion_coeffs);
    n.rel_error_power_balance = s.number(k + "relError.powerBalance", n.rel_error_power_balance);
    n.n_integration_points = s.number(k + "nIntegrationPoints", n.n_integration_points);
    n.n_integrated_ss_times = s.number(k + "nIntegratedSSTimes", n.n_integrated_ss_times);
    n.integrated_absolute_time = s.number(k + "integratedAbsoluteTime", n.integrated_absolute_time);
    n.n_integration_phases = (int32_t)s.number(k + "nIntegrationPhases", n.n_integration_phases);
    n.n_interp_points = (int32_t)s.number(k + "nInterpPoints", n.n_interp_points);
    n.sync_time_x_max_coll_freq = s.number(k + "synchronizationTimeXMaxCollisionFrequency", n.sync_time_x_max_coll_freq);
    n.sync_over_sampling = (int32_t)s.number(k + "synchronizationOverSampling", n.sync_over_sampling);
    return n;
}

std::string jsonString(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}

void dryRun(const Jobs& jobs, const ekmc_numerics& n, const GasMixture& m) {
    std::printf("{\"variable\": %s, \"jobs\": [", jsonString(jobs.variable).c_str());
    for (size_t k = 0; k < jobs.conds.size(); ++k) {
        const ekmc_condition& c = jobs.conds[k];
        std::printf("%s{\"reduced_elec_field\": %.17g, \"reduced_mag_field\": %.17g, \"elec_field_angle\": %.17g, "
                    "\"excitation_frequency\": %.17g, \"gas_temperature\": %.17g, \"gas_density\": %.17g, \"n_electrons\": %lld}",
                    k ? ", " : "", c.reduced_elec_field, c.reduced_mag_field, c.elec_field_angle, c.excitation_frequency,
                    c.gas_temperature, c.gas_density, (long long)c.n_electrons);
    }
    std::printf("], \"numerics\": {\"ionization_operator\": %d, \"gas_temperature_effect\": %d, \"n_integration_points\": %.17g, "
                "\"n_integrated_ss_times\": %.17g, \"errors_to_be_checked\": %d, \"n_energy_cells\": %d, "
                "\"initial_elec_temp_over_gas_temp\": %.17g},",
                n.ionization_operator, n.gas_temperature_effect, n.n_integration_points, n.n_integrated_ss_times,
                n.errors_to_be_checked, n.n_energy_cells, n.initial_elec_temp_over_gas_temp);
    std::printf(" \"gases\": [");
    for (size_t g = 0; g < m.gasNames.size(); ++g)
        std::printf("%s{\"name\": %s, \"mass\": %.17g, \"fraction\": %.17g}", g ? ", " : "", jsonString(m.gasNames[g]).c_str(),
                    m.gasMass[g], m.gasFraction[g]);
    std::printf("], \"processes\": [");
    for (int p = 0; p < m.nProcesses(); ++p)
        std::printf("%s{\"description\": %s, \"type\": %d, \"elastic\": %d, \"superelastic\": %d, \"gas\": %d, \"threshold\": %.17g, "
                    "\"rel_density\": %.17g, \"stat_weight_ratio\": %.17g, \"w\": %.17g, \"scattering\": %d, \"points\": %d}",
                    p ? ", " : "", jsonString(m.description[p]).c_str(), m.type[p], m.isElastic[p], m.isSuperelastic[p],
                    m.gasIndex[p], m.threshold[p], m.relDensity[p], m.statWeightRatio[p], m.wParameter[p], m.scattering[p],
                    m.tableOffset[p + 1] - m.tableOffset[p]);
    // per collision: densities and both raw cross sections
    auto array = [](const std::vector<double>& v) {
        std::string o = "[";
        char buf[32];
        for (size_t i = 0; i < v.size(); ++i) {
            std::snprintf(buf, sizeof buf, "%s%.17g", i ? ", " : "", v[i]);
            o += buf;
        }
        return o + "]";
    };
    std::printf("], \"collisions\": [");
    for (size_t c = 0; c < m.collisionDescription.size(); ++c)
        std::printf("%s{\"description\": %s, \"kind\": %d, \"ref_id\": %d, \"type\": %s, \"threshold\": %.17g, \"target_density\": %.17g, "
                    "\"product_density\": %.17g, \"stat_weight_ratio\": %.17g, \"integral_energy\": %s, \"integral\": %s, "
                    "\"mom_transf_energy\": %s, \"mom_transf\": %s}",
                    c ? ", " : "", jsonString(m.collisionDescription[c]).c_str(), m.collisionKind[c], m.collisionRefId[c],
                    jsonString(m.collisionType[c]).c_str(),
                    m.collisionThreshold[c], m.collisionTargetDensity[c], m.collisionProductDensity[c],
                    m.collisionStatWeightRatio[c], array(m.collisionEnergy[c]).c_str(), array(m.collisionSigmaIntegral[c]).c_str(),
                    array(m.collisionEnergyMomTransf[c]).c_str(), array(m.collisionSigmaMomTransf[c]).c_str());
    std::printf("]}\n");
}

// Message::error (Message.C:6-16): the message goes to errorLog.txt and the run ends.
[[noreturn]] void fatal(const std::string& msg) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    {
        std::ofstream log("errorLog.txt", std::ios::app);
        log << msg << "\n";
    }
    std::exit(1);
}

}  // namespace

int main(int argc, char** argv) {
    std::string setupPath, inputDir = "Input", outputDir = "Output", precision = "fp64";
    uint64_t seed = 0;
    bool dry = false, autotune = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) fatal("missing value after " + a);
            return argv[++i];
        };
        if (a == "--input") inputDir = next();
        else if (a == "--output") outputDir = next();
        else if (a == "--precision") precision = next();
        else if (a == "--seed") seed = std::stoull(next());
        else if (a == "--dry-run") dry = true;
        else if (a == "--autotune") autotune = true;
        else if (!a.empty() && a[0] == '-') fatal("unknown option " + a);
        else setupPath = a;
    }
    if (setupPath.empty()) {
        std::fprintf(stderr, "usage: lokimc_gpu <setup file> [--input DIR] [--output DIR] [--precision fp64|mixed] [--seed N] "
                     "[--autotune] [--dry-run]\n");
        return 2;
    }
    try {
        SetupFile setup = SetupFile::fromFile(setupPath);
        if (setup.value("electronKinetics.eedfType") != "boltzmannMC")
            throw InputError("eedfType must be 'boltzmannMC' (this program runs the Monte Carlo electron kinetics only).");
        const double nEl = setup.number("electronKinetics.numericsMC.nElectrons", -1);
        if (nEl < 1) throw InputError("electronKinetics.numericsMC.nElectrons must be defined.");
        const Jobs jobs = workingConditions(setup, (int64_t)std::llround(nEl));
        ekmc_numerics num = numerics(setup);
        num.seed = seed;
        if (precision == "mixed") num.precision = EKMC_PRECISION_MIXED;
        else if (precision != "fp64") throw InputError("--precision must be fp64 or mixed");
        const GasMixture mix = buildMixture(setup, inputDir, evalNumber(setup.value("workingConditions.gasTemperature")));
        if (dry) {
            dryRun(jobs, num, mix);
            return 0;
        }
        const ekmc_mixture view = mix.view();
        if (autotune) {
            ekmc_numerics best{};
            if (ekmc_autotune_launch(&view, jobs.conds.data(), (int32_t)jobs.conds.size(), &num, 100, 100, &best,
                                     nullptr) != EKMC_OK)
                fatal(ekmc_last_error());
            num = best;
            std::printf("autotuned launch: work unit %d electrons, collision step at %d lanes, envelope coarsening %d\n",
                        best.work_chunk, best.park_lanes, best.env_coarsen);
        }
        ekmc_engine* eng = nullptr;
        if (ekmc_create(&view, jobs.conds.data(), (int32_t)jobs.conds.size(), &num, &eng) != EKMC_OK) fatal(ekmc_last_error());
        if (ekmc_solve(eng) != EKMC_OK) fatal(ekmc_last_error());
        std::vector<JobData> data;
        bool allAt180 = true;
        for (const ekmc_condition& c : jobs.conds) allAt180 &= c.elec_field_angle == 180.0;
        for (size_t k = 0; k < jobs.conds.size(); ++k)
            data.push_back(collectJob(eng, (int)k, jobs.conds[k], num, mix.nProcesses(), allAt180));
        ekmc_destroy(eng);
        const std::string folderName = setup.value("output.folder").empty() ? "lokimc_gpu" : setup.value("output.folder");
        const std::string folder = outputDir + "/" + folderName;
        if (setup.value("output.isOn") != "false") {
            std::filesystem::create_directories(folder);
            {
                std::ofstream st(folder + "/setup.txt");
                st << setup.normalized();
            }
            OutputOptions opt;
            opt.dataFiles = setup.childNames("output.dataFiles");
            opt.variableCondition = jobs.variable;
            for (size_t k = 0; k < data.size(); ++k) {
                std::string dir = folder;
                if (data.size() > 1) {
                    const ekmc_condition& c = jobs.conds[k];
                    const double v = jobs.variable == "reducedMagField" ? c.reduced_mag_field
                                     : jobs.variable == "elecFieldAngle" ? c.elec_field_angle
                                     : jobs.variable == "excitationFrequency" ? c.excitation_frequency
                                                                              : c.reduced_elec_field;
                    char buf[100];
                    std::snprintf(buf, sizeof buf, "%g", v);
                    dir += "/" + jobs.variable + "_" + buf;
                    std::filesystem::create_directories(dir);
                }
                writeJobFiles(dir, data[k], mix, opt);
            }
            writeLookUpTables(folder, data, mix, opt);
        }
        for (size_t k = 0; k < data.size(); ++k)
            std::printf("job %zu: %s = %g, mean energy %.6g eV, flux drift %.6g m/s, %llu real + %llu null collisions\n", k,
                        jobs.variable.c_str(),
                        jobs.variable == "reducedMagField" ? jobs.conds[k].reduced_mag_field
                        : jobs.variable == "elecFieldAngle" ? jobs.conds[k].elec_field_angle
                        : jobs.variable == "excitationFrequency" ? jobs.conds[k].excitation_frequency
                                                                 : jobs.conds[k].reduced_elec_field,
                        data[k].res.mean_energy, std::fabs(data[k].res.rotated_flux_drift_velocity[2]),
                        (unsigned long long)data[k].res.real_collisions, (unsigned long long)data[k].res.null_collisions);
    } catch (const std::exception& e) {
        fatal(e.what());
    }
    return 0;
}
