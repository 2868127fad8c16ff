// outputs.cpp -- see outputs.h.
#include "outputs.h"

#include <cmath>
#include <cstdio>
#include <map>
#include <stdexcept>

namespace lokimc {

namespace {

constexpr double kCharge = 1.6021766208e-19;  // Constant.h:6-23
constexpr double kMass = 9.10938356e-31;
constexpr double kPi = 3.14159265358979323846;  // M_PI

void check(ekmc_status s, const char* what) {
    if (s != EKMC_OK) throw std::runtime_error(std::string(what) + ": " + ekmc_last_error());
}

struct File {
    FILE* f;
    explicit File(const std::string& path) : f(std::fopen(path.c_str(), "w")) {
        if (!f) throw std::runtime_error("The file '" + path + "' could not be created.");
    }
    ~File() { if (f) std::fclose(f); }
};

bool wants(const OutputOptions& o, const char* name) {
    for (const std::string& d : o.dataFiles)
        if (d == name) return true;
    return false;
}

// Cross section of collision c at `energy` (Collision::interpolatedCrossSection,
// Collision.C:231-289): linear between the table points, zero outside, and
// zero at or below the threshold of an inelastic collision.
double sigmaAt(const GasMixture& m, int c, bool momTransf, double energy) {
    const std::vector<double>& e = momTransf ? m.collisionEnergyMomTransf[c] : m.collisionEnergy[c];
    const std::vector<double>& sig = momTransf ? m.collisionSigmaMomTransf[c] : m.collisionSigmaIntegral[c];
    const bool elasticLike = m.collisionType[c] == "Elastic" || m.collisionType[c] == "Effective";
    if (!elasticLike && !(energy > m.collisionThreshold[c])) return 0.0;
    if (e.empty() || energy < e.front() || energy > e.back()) return 0.0;
    size_t lo = 0, hi = e.size() - 1;
    while (hi - lo > 1) {
        const size_t mid = (lo + hi) / 2;
        if (e[mid] <= energy) lo = mid; else hi = mid;
    }
    const double dx = e[hi] - e[lo];
    return dx > 0 ? sig[lo] + (energy - e[lo]) / dx * (sig[hi] - sig[lo]) : sig[lo];
}

// Energy grid of the EEDF: nodes 0..maxE and cell centres (Grid.C:32-48).
struct EnergyGrid {
    double step = 0;
    std::vector<double> node, cell;
};
EnergyGrid gridOf(const JobData& j) {
    EnergyGrid g;
    const size_t n = j.energyCells.size();
    if (n == 0) return g;
    g.step = 2.0 * j.energyCells[0];
    g.cell = j.energyCells;
    g.node.resize(n + 1);
    for (size_t i = 0; i <= n; ++i) g.node[i] = g.step * (double)i;
    return g;
}

// Cross section on the nodes; a superelastic one by Klein-Rosseland
// (Collision.C:291-353).
std::vector<double> onNodes(const GasMixture& m, int c, const EnergyGrid& g, bool momTransf, bool super) {
    std::vector<double> out(g.node.size(), 0.0);
    for (size_t i = 0; i < g.node.size(); ++i) {
        const double u = g.node[i];
        if (!super) {
            out[i] = sigmaAt(m, c, momTransf, u);
        } else if (u != 0) {
            const double thr = m.collisionThreshold[c];
            out[i] = (1.0 + thr / u) * sigmaAt(m, c, momTransf, u + thr) * m.collisionStatWeightRatio[c];
        }
    }
    return out;
}

}  // namespace

double SwarmParams::get(const std::string& k) const {
    for (const auto& kv : items)
        if (kv.first == k) return kv.second;
    return 0.0;
}
void SwarmParams::set(const std::string& k, double v) {
    for (auto& kv : items)
        if (kv.first == k) { kv.second = v; return; }
    items.emplace_back(k, v);
}

JobData collectJob(const ekmc_engine* eng, int cond, const ekmc_condition& c, const ekmc_numerics& num, int nProcesses,
                   bool allAt180) {
    JobData j;
    j.cond = c;
    j.nElectrons = (double)c.n_electrons;
    check(ekmc_get_results(eng, cond, &j.res), "ekmc_get_results");
    j.rateMC.resize(nProcesses);
    j.powerGain.resize(nProcesses);
    j.powerLoss.resize(nProcesses);
    j.counts.resize(nProcesses);
    check(ekmc_get_process_data(eng, cond, j.rateMC.data(), j.counts.data(), j.powerGain.data(), j.powerLoss.data()),
          "ekmc_get_process_data");
    j.energyCells.resize(num.n_energy_cells);
    j.eedf.resize(num.n_energy_cells);
    check(ekmc_get_eedf(eng, cond, j.energyCells.data(), j.eedf.data()), "ekmc_get_eedf");
    j.cylindrical = allAt180 && c.elec_field_angle == 180.0;
    if (j.cylindrical) {
        j.radialCells.resize(num.n_radial_velocity_cells);
        j.axialCells.resize(num.n_axial_velocity_cells);
        j.evdf.resize((size_t)num.n_radial_velocity_cells * num.n_axial_velocity_cells);
        j.efadf.resize(num.n_energy_cells);
        j.esadf.resize(num.n_energy_cells);
        check(ekmc_get_evdf(eng, cond, j.radialCells.data(), j.axialCells.data(), j.evdf.data(), j.efadf.data(),
                            j.esadf.data()),
              "ekmc_get_evdf");
    }
    int32_t n = j.res.n_sampling_points;
    j.times.resize(n); j.meanEnergy.resize(n); j.meanPosition.resize(3 * (size_t)n); j.meanVelocity.resize(3 * (size_t)n);
    j.fluxDiffusion.resize(9 * (size_t)n); j.bulkVelocity.resize(3 * (size_t)n); j.bulkDiffusion.resize(9 * (size_t)n);
    check(ekmc_get_series(eng, cond, &n, j.times.data(), j.meanEnergy.data(), j.meanPosition.data(), j.meanVelocity.data(),
                          j.fluxDiffusion.data(), j.bulkVelocity.data(), j.bulkDiffusion.data()),
          "ekmc_get_series");
    j.nSeries = n;
    j.positionCovariance.resize(9 * (size_t)n);
    int32_t nc = n;
    check(ekmc_get_position_covariance(eng, cond, &nc, j.positionCovariance.data()), "ekmc_get_position_covariance");
    if (c.excitation_frequency != 0) {
        const int np = num.n_integration_phases, ncell = num.n_energy_cells;
        j.nPhases = np;
        j.phases.resize(np);
        j.eedfPeriodic.resize((size_t)np * ncell);
        j.meanEnergyPeriodic.resize(np);
        j.fluxVelocityPeriodic.resize(3 * (size_t)np);
        j.bulkVelocityPeriodic.resize(3 * (size_t)np);
        j.fluxDiffusionPeriodic.resize(9 * (size_t)np);
        j.bulkDiffusionPeriodic.resize(9 * (size_t)np);
        check(ekmc_get_eedf_periodic(eng, cond, j.eedfPeriodic.data(), nullptr, nullptr), "ekmc_get_eedf_periodic");
        check(ekmc_get_swarm_periodic(eng, cond, j.phases.data(), j.meanEnergyPeriodic.data(), j.fluxVelocityPeriodic.data(),
                                      j.bulkVelocityPeriodic.data(), j.fluxDiffusionPeriodic.data(),
                                      j.bulkDiffusionPeriodic.data()),
              "ekmc_get_swarm_periodic");
    }
    return j;
}

void eedfRateCoefficients(const JobData& j, const GasMixture& m, std::vector<double>& ine, std::vector<double>& sup,
                          const double* eedf) {
    // Collision::evaluateRateCoeff (Collision.C:355-397)
    const EnergyGrid g = gridOf(j);
    if (!eedf) eedf = j.eedf.data();
    const double factor = std::sqrt(2.0 * kCharge / kMass);
    const int nc = (int)m.collisionDescription.size();
    ine.assign(nc, 0.0);
    sup.assign(nc, std::nan(""));
    for (int c = 0; c < nc; ++c) {
        const std::vector<double> sig = onNodes(m, c, g, false, false);
        const int lmin = (int)std::floor(m.collisionThreshold[c] / g.step);
        const int part = (int)sig.size() - 1 - lmin;
        if (part <= 0) {
            ine[c] = 0;
            if (m.collisionIsReverse[c]) sup[c] = 0;
            continue;
        }
        double si = 0, ss = 0;
        for (int k = 0; k < part; ++k) {
            const double aux = 0.5 * (sig[lmin + k] + sig[lmin + k + 1]) * g.cell[lmin + k];
            si += aux * eedf[lmin + k];
            ss += aux * eedf[k];
        }
        ine[c] = factor * si * g.step;
        if (m.collisionIsReverse[c]) sup[c] = factor * m.collisionStatWeightRatio[c] * ss * g.step;
    }
}

SwarmParams swarmParameters(const JobData& j, const GasMixture& m) {
    // evaluateSwarmParameters (BoltzmannMC.C:2064-2266)
    const ekmc_results& r = j.res;
    const double N = j.cond.gas_density, ENsi = j.cond.reduced_elec_field * 1e-21, f = j.cond.excitation_frequency;
    SwarmParams s;
    s.set("fluxRedTransvDiffCoeff", N * (r.rotated_flux_diffusion[0] + r.rotated_flux_diffusion[4]) / 2.0);
    s.set("fluxRedTransvDiffCoeffError", N * (r.rotated_flux_diffusion_error[0] + r.rotated_flux_diffusion_error[4]) / 2.0);
    s.set("fluxRedLongDiffCoeff", N * r.rotated_flux_diffusion[8]);
    s.set("fluxRedLongDiffCoeffError", N * r.rotated_flux_diffusion_error[8]);
    std::vector<double> ine, sup;
    eedfRateCoefficients(j, m, ine, sup);
    double ion = 0, att = 0;
    for (size_t c = 0; c < ine.size(); ++c) {
        if (m.collisionKind[c] != 0) continue;
        if (m.collisionType[c] == "Ionization") ion += m.collisionTargetDensity[c] * ine[c];
        else if (m.collisionType[c] == "Attachment") att += m.collisionTargetDensity[c] * ine[c];
    }
    s.set("totalIonRateCoeff", ion);
    s.set("totalAttRateCoeff", att);
    auto drift = [&](const char* pre, const double* vrot, const double* verr, const double* drot, const double* derr) {
        const std::string p = pre;
        if (f != 0) return;
        const double v = std::fabs(vrot[2]);
        s.set(p + "RedMobCoeff", v / ENsi);
        s.set(p + "RedMobCoeffError", std::fabs(verr[2]) / ENsi);
        s.set(p + "RedTownsendCoeff", ion / v);
        s.set(p + "RedAttCoeff", att / v);
        const double mob = s.get(p + "RedMobCoeff"), dt = s.get(p + "RedTransvDiffCoeff");
        s.set(p + "CharacEnergy", dt / mob);
        s.set(p + "CharacEnergyError", s.get(p + "RedTransvDiffCoeffError") / mob + s.get(p + "RedMobCoeffError") * dt / (mob * mob));
        (void)drot; (void)derr;
    };
    drift("flux", r.rotated_flux_drift_velocity, r.rotated_flux_drift_velocity_error, r.rotated_flux_diffusion,
          r.rotated_flux_diffusion_error);
    s.set("bulkRedTransvDiffCoeff", N * (r.rotated_bulk_diffusion[0] + r.rotated_bulk_diffusion[4]) / 2.0);
    s.set("bulkRedTransvDiffCoeffError", N * (r.rotated_bulk_diffusion_error[0] + r.rotated_bulk_diffusion_error[4]) / 2.0);
    s.set("bulkRedLongDiffCoeff", N * r.rotated_bulk_diffusion[8]);
    s.set("bulkRedLongDiffCoeffError", N * r.rotated_bulk_diffusion_error[8]);
    drift("bulk", r.rotated_bulk_drift_velocity, r.rotated_bulk_drift_velocity_error, r.rotated_bulk_diffusion,
          r.rotated_bulk_diffusion_error);
    if (f == 0) {
        const double v = std::fabs(r.rotated_bulk_drift_velocity[2]);
        const double sst = 0.5 * v + std::sqrt(0.25 * v * v - N * r.rotated_bulk_diffusion[8] * (ion - att));
        s.set("effSSTAverageVelocity", sst);
        s.set("effSSTRedTownsendCoeff", ion / sst);
        s.set("effSSTRedAttCoeff", att / sst);
    }
    s.set("meanEnergy", r.mean_energy);
    s.set("meanEnergyError", r.mean_energy_error);
    s.set("Te", 2.0 / 3.0 * r.mean_energy);
    s.set("TeError", 2.0 / 3.0 * r.mean_energy_error);
    // two-term expressions from the EEDF
    const EnergyGrid g = gridOf(j);
    const int Nn = (int)g.node.size(), Nc = (int)g.cell.size();
    if (Nc < 2) return s;
    const double factor = std::sqrt(2.0 * kCharge / kMass) / 3.0;
    std::vector<double> mt(Nn, 0.0), er(Nn, 0.0), erc(Nn, 0.0);
    for (size_t c = 0; c < m.collisionDescription.size(); ++c) {
        if (m.collisionKind[c] != 0) continue;  // Effective and extra collisions are not part of the gas's set
        const double dens = m.collisionTargetDensity[c];
        const std::vector<double> smt = onNodes(m, (int)c, g, true, false), sint = onNodes(m, (int)c, g, false, false);
        const double ratio = kMass / m.gasMass[m.collisionGas[c]];
        for (int i = 0; i < Nn; ++i) {
            mt[i] += smt[i] * dens;
            if (m.collisionType[c] == "Elastic") {
                er[i] += 2.0 * ratio * sint[i] * dens;
            } else {
                const double w = sint[i] * dens;
                if (i > 0) er[i] += w * m.collisionThreshold[c] / g.node[i];
                erc[i] += w;
            }
        }
        if (m.collisionIsReverse[c]) {
            const double pd = m.collisionProductDensity[c];
            const std::vector<double> supI = onNodes(m, (int)c, g, false, true), supM = onNodes(m, (int)c, g, true, true);
            for (int i = 0; i < Nn; ++i) {
                const double w = supI[i] * pd;
                if (i > 0) er[i] += w * m.collisionThreshold[c] / g.node[i];
                mt[i] += supM[i] * pd;
                erc[i] += w;
            }
        }
    }
    double smt = 0, ser = 0, serc = 0;
    for (int i = 0; i < Nc; ++i) {
        const double w = j.eedf[i] * g.cell[i];
        smt += w * 0.5 * (mt[i + 1] + mt[i]);
        ser += w * 0.5 * (er[i + 1] + er[i]);
        serc += w * 0.5 * (erc[i + 1] + erc[i]);
    }
    s.set("momTransfFreq", N * 3.0 * factor * g.step * smt);
    s.set("energyRelaxFreq", N * 3.0 * factor * g.step * ser);
    s.set("energyRelaxFreqClassic", N * 3.0 * factor * g.step * serc);
    std::vector<double> aux = mt;
    for (int i = 1; i < Nn; ++i) aux[i] += 1.0 / (3.0 * factor * std::sqrt(g.node[i])) * (ion - att);
    double rde = 0, rme = 0, rd = 0, rm = 0;
    for (int i = 0; i < Nn - 1; ++i) {
        rde += g.cell[i] * g.cell[i] * j.eedf[i] / (aux[i] + aux[i + 1]);
        rd += g.cell[i] * j.eedf[i] / (aux[i] + aux[i + 1]);
    }
    for (int i = 1; i < Nn - 1; ++i) {
        const double de = j.eedf[i] - j.eedf[i - 1];
        rme += g.node[i] * g.node[i] * de / aux[i];
        rm += g.node[i] * de / aux[i];
    }
    s.set("redDiffCoeffEnergy_eedf", 2.0 * factor * g.step * rde);
    s.set("redMobCoeffEnergy_eedf", -factor * rme);
    s.set("redDiffCoeff_eedf", 2.0 * factor * g.step * rd);
    s.set("redMobCoeff_DC_eedf", -factor * rm);
    s.set("characEnergy_eedf", s.get("redDiffCoeff_eedf") / s.get("redMobCoeff_DC_eedf"));
    return s;
}

namespace {

// evaluatePower (BoltzmannMC.C:1950-2062)
struct Power {
    std::map<std::string, double> total;
    std::map<std::string, std::map<std::string, double>> gas;  // sorted by gas name, as the reference's map
};
Power powerOf(const JobData& j, const GasMixture& m) {
    Power p;
    const char* keys[] = {"field", "elasticNet", "elasticGain", "elasticLoss", "carNet", "carGain", "carLoss",
                          "excitationIne", "excitationSup", "excitationNet", "vibrationalIne", "vibrationalSup",
                          "vibrationalNet", "rotationalIne", "rotationalSup", "rotationalNet", "ionizationIne",
                          "attachmentIne", "inelastic", "superelastic", "eDensGrowth"};
    for (const char* k : keys) p.total[k] = 0;
    p.total["field"] = j.res.power_field;
    for (int i = 0; i < m.nProcesses(); ++i)
        if (m.lxcatType[i] == "Elastic") {
            p.total["elasticGain"] += j.powerGain[i];
            p.total["elasticLoss"] += j.powerLoss[i];
        }
    p.total["elasticNet"] = p.total["elasticGain"] + p.total["elasticLoss"];
    for (size_t gi = 0; gi < m.allGasNames.size(); ++gi) {
        std::map<std::string, double> g;
        for (const char* k : {"excitationIne", "excitationSup", "excitationNet", "vibrationalIne", "vibrationalSup",
                              "vibrationalNet", "rotationalIne", "rotationalSup", "rotationalNet", "ionizationIne",
                              "attachmentIne"})
            g[k] = 0;
        for (int i = 0; i < m.nProcesses(); ++i) {
            if (m.gasNames[m.gasIndex[i]] != m.allGasNames[gi]) continue;
            const double v = j.powerGain[i] + j.powerLoss[i];
            const std::string& t = m.lxcatType[i];
            const bool sup = m.isSuperelastic[i] != 0;
            if (t == "Excitation") g[sup ? "excitationSup" : "excitationIne"] += v;
            else if (t == "Vibrational") g[sup ? "vibrationalSup" : "vibrationalIne"] += v;
            else if (t == "Rotational") g[sup ? "rotationalSup" : "rotationalIne"] += v;
            else if (t == "Ionization") g["ionizationIne"] += v;
            else if (t == "Attachment") g["attachmentIne"] += v;
        }
        g["excitationNet"] = g["excitationIne"] + g["excitationSup"];
        g["vibrationalNet"] = g["vibrationalIne"] + g["vibrationalSup"];
        g["rotationalNet"] = g["rotationalIne"] + g["rotationalSup"];
        for (const char* k : {"excitationIne", "excitationSup", "vibrationalIne", "vibrationalSup", "rotationalIne",
                              "rotationalSup", "ionizationIne", "attachmentIne"})
            p.total[k] += g[k];
        p.gas[m.allGasNames[gi]] = g;
    }
    auto& t = p.total;
    t["excitationNet"] = t["excitationIne"] + t["excitationSup"];
    t["vibrationalNet"] = t["vibrationalIne"] + t["vibrationalSup"];
    t["rotationalNet"] = t["rotationalIne"] + t["rotationalSup"];
    t["inelastic"] = t["excitationIne"] + t["vibrationalIne"] + t["rotationalIne"] + t["ionizationIne"] + t["attachmentIne"];
    t["superelastic"] = t["excitationSup"] + t["vibrationalSup"] + t["rotationalSup"];
    t["eDensGrowth"] = j.res.power_growth;
    t["balance"] = t["field"] + t["elasticNet"] + t["inelastic"] + t["superelastic"] + t["eDensGrowth"];
    double gain = 0;
    for (const char* k : {"field", "elasticGain", "elasticLoss", "excitationSup", "excitationIne", "vibrationalSup",
                          "vibrationalIne", "rotationalSup", "rotationalIne", "eDensGrowth"})
        if (t[k] > 0) gain += t[k];
    t["relativeBalance"] = std::fabs(t["balance"]) / gain;
    return p;
}

void writeEedf(const std::string& dir, const JobData& j) {
    File out(dir + "/eedf.txt");
    if (!j.cylindrical) {
        std::fprintf(out.f, "Energy(eV)           EEDF(eV^-(3/2))\n");
        for (size_t i = 0; i < j.eedf.size(); ++i) std::fprintf(out.f, "%-20.14e %-20.14e \n", j.energyCells[i], j.eedf[i]);
    } else {
        std::fprintf(out.f, "Energy(eV)           EEDF(eV^-(3/2))      First Anisotropy     Second Anisotropy    \n");
        for (size_t i = 0; i < j.eedf.size(); ++i)
            std::fprintf(out.f, "%-20.14e %-20.14e %-20.14e %-20.14e\n", j.energyCells[i], j.eedf[i], j.efadf[i], j.esadf[i]);
    }
    if (j.cond.excitation_frequency == 0) return;
    // one EEDF per phase of the period (Output.h:207-220)
    File per(dir + "/eedf_periodic.txt");
    const size_t nc = j.eedf.size();
    for (int ph = 0; ph < j.nPhases; ++ph) {
        std::fprintf(per.f, "%-20.14e ", j.phases[ph]);
        for (size_t i = 0; i < nc; ++i) std::fprintf(per.f, "%-20.14e ", j.eedfPeriodic[(size_t)ph * nc + i]);
        std::fprintf(per.f, "\n");
    }
}

void writeEvdf(const std::string& dir, const JobData& j) {
    File out(dir + "/evdf.txt");
    std::fprintf(out.f, "v_r(m/s)           v_z(m/s)           EVDF(m-3s-3)       \n");
    const size_t nr = j.radialCells.size(), na = j.axialCells.size();
    for (size_t i = 0; i < nr; ++i)
        for (size_t k = 0; k < na; ++k)
            std::fprintf(out.f, "%-18.10e %-18.10e %-18.10e\n", j.radialCells[i], j.axialCells[k], j.evdf[i * na + k]);
}

void writeVector(FILE* f, const char* lead, const char* label, const double* v, const double* e, bool primed) {
    double rel[3];
    for (int k = 0; k < 3; ++k) rel[k] = e[k] / std::fabs(v[k]) * 100.0;
    if (!primed) {
        std::fprintf(f, "%s| v_x |   | %-15.8e |                       | %-9.3e%% |\n", lead, v[0], rel[0]);
        std::fprintf(f, "%s| v_y | = | %-15.8e |  (m/s)    ; Rel. std: | %-9.3e%% |\n", lead, v[1], rel[1]);
        std::fprintf(f, "%s| v_z |   | %-15.8e |                       | %-9.3e%% |\n\n", lead, v[2], rel[2]);
    } else {
        std::fprintf(f, "%s| v_x' |   | %-15.8e |                       | %-9.3e%% |\n", lead, v[0], rel[0]);
        std::fprintf(f, "%s| v_y' | = | %-15.8e |  (m/s)    ; Rel. std: | %-9.3e%% |\n", lead, v[1], rel[1]);
        std::fprintf(f, "%s| v_z' |   | %-15.8e |                       | %-9.3e%% |\n\n", lead, v[2], rel[2]);
    }
    (void)label;
}

void writeMatrix(FILE* f, double N, const double* d, const double* e, bool primed) {
    double r[9], x[9];
    for (int k = 0; k < 9; ++k) { x[k] = d[k] * N; r[k] = e[k] / std::fabs(d[k]) * 100.0; }
    if (!primed) {
        std::fprintf(f, "                      | ND_xx ND_xy ND_xz|   | %-15.8e %-15.8e %-15.8e |                      | %-9.3e%% %-9.3e%% %-9.3e%% |\n", x[0], x[1], x[2], r[0], r[1], r[2]);
        std::fprintf(f, "                      | ND_yx ND_yy ND_yz| = | %-15.8e %-15.8e %-15.8e | (1/(ms)) ; Rel. std: | %-9.3e%% %-9.3e%% %-9.3e%% |\n", x[3], x[4], x[5], r[3], r[4], r[5]);
        std::fprintf(f, "                      | ND_zx ND_zy ND_zz|   | %-15.8e %-15.8e %-15.8e |                      | %-9.3e%% %-9.3e%% %-9.3e%% |\n\n", x[6], x[7], x[8], r[6], r[7], r[8]);
    } else {
        std::fprintf(f, "                | ND_x'x' ND_x'y' ND_x'z'|   | %-15.8e %-15.8e %-15.8e |                      | %-9.3e%% %-9.3e%% %-9.3e%% |\n", x[0], x[1], x[2], r[0], r[1], r[2]);
        std::fprintf(f, "                | ND_y'x' ND_y'y' ND_y'z'| = | %-15.8e %-15.8e %-15.8e | (1/(ms)) ; Rel. std: | %-9.3e%% %-9.3e%% %-9.3e%% |\n", x[3], x[4], x[5], r[3], r[4], r[5]);
        std::fprintf(f, "                | ND_z'x' ND_z'y' ND_z'z'|   | %-15.8e %-15.8e %-15.8e |                      | %-9.3e%% %-9.3e%% %-9.3e%% |\n\n", x[6], x[7], x[8], r[6], r[7], r[8]);
    }
}

std::string banner(int left, const std::string& text, int right) {
    return std::string(left, '*') + " " + text + " " + std::string(right, '*');
}

void writeSwarm(const std::string& dir, const JobData& j, const SwarmParams& s) {
    File out(dir + "/swarmParameters.txt");
    FILE* f = out.f;
    const ekmc_results& r = j.res;
    const ekmc_condition& c = j.cond;
    const double N = c.gas_density;
    const char* lead = "                                   ";
    const char* leadP = "                                  ";
    std::fprintf(f, "                    Reduced electric field = %#.14e (Td)\n", c.reduced_elec_field);
    std::fprintf(f, "                      Electric field angle = %#.14e (Degrees)\n", c.elec_field_angle);
    std::fprintf(f, "                      Excitation frequency = %#.14e (Hz)\n", c.excitation_frequency);
    std::fprintf(f, "                    Reduced magnetic field = %#.14e (Hx)\n", c.reduced_mag_field);
    for (int pass = 0; pass < 2; ++pass) {
        const bool flux = pass == 0;
        const std::string p = flux ? "flux" : "bulk";
        std::fprintf(f, "\n%s\n\n", banner(35, flux ? "Flux parameters" : "Bulk parameters", 35).c_str());
        writeVector(f, lead, "", flux ? r.flux_drift_velocity : r.bulk_drift_velocity,
                    flux ? r.flux_drift_velocity_error : r.bulk_drift_velocity_error, false);
        writeMatrix(f, N, flux ? r.flux_diffusion : r.bulk_diffusion, flux ? r.flux_diffusion_error : r.bulk_diffusion_error, false);
        std::fprintf(f, "Parameters after rotation to a ref. frame (x'y'z'), where E-field is along z'\n\n");
        writeVector(f, leadP, "", flux ? r.rotated_flux_drift_velocity : r.rotated_bulk_drift_velocity,
                    flux ? r.rotated_flux_drift_velocity_error : r.rotated_bulk_drift_velocity_error, true);
        writeMatrix(f, N, flux ? r.rotated_flux_diffusion : r.rotated_bulk_diffusion,
                    flux ? r.rotated_flux_diffusion_error : r.rotated_bulk_diffusion_error, true);
        std::fprintf(f, "  Reduced transverse diffusion coefficient = %#.14e (1/(ms)) ; Rel. std: %-9.3e%%\n",
                     s.get(p + "RedTransvDiffCoeff"), s.get(p + "RedTransvDiffCoeffError") / s.get(p + "RedTransvDiffCoeff") * 100.0);
        std::fprintf(f, "Reduced longitudinal diffusion coefficient = %#.14e (1/(ms)) ; Rel. std: %-9.3e%%\n",
                     s.get(p + "RedLongDiffCoeff"), s.get(p + "RedLongDiffCoeffError") / s.get(p + "RedLongDiffCoeff") * 100.0);
        if (c.reduced_mag_field == 0 && c.excitation_frequency == 0) {
            std::fprintf(f, "              Reduced mobility coefficient = %#.14e (1/(msV)); Rel. std: %-9.3e%%\n",
                         s.get(p + "RedMobCoeff"), s.get(p + "RedMobCoeffError") / s.get(p + "RedMobCoeff") * 100.0);
            std::fprintf(f, "                     Characteristic energy = %#.14e (eV)     ; Rel. std: %-9.3e%%\n",
                         s.get(p + "CharacEnergy"), s.get(p + "CharacEnergyError") / s.get(p + "CharacEnergy") * 100.0);
        }
        if (c.excitation_frequency == 0) {
            std::fprintf(f, "              Reduced Townsend coefficient = %#.14e (m2)\n", s.get(p + "RedTownsendCoeff"));
            std::fprintf(f, "            Reduced attachment coefficient = %#.14e (m2)\n", s.get(p + "RedAttCoeff"));
        }
    }
    std::fprintf(f, "\n%s\n\n", banner(15, "Effective SST parameters deduced from the TOF simulation", 15).c_str());
    std::fprintf(f, "                     SST averaged velocity = %#.14e (m/s)\n", s.get("effSSTAverageVelocity"));
    std::fprintf(f, "          SST reduced Townsend coefficient = %#.14e (m/s)\n", s.get("effSSTRedTownsendCoeff"));
    std::fprintf(f, "        SST reduced attachment coefficient = %#.14e (m/s)\n", s.get("effSSTRedAttCoeff"));
    std::fprintf(f, "\n%s\n\n", banner(35, "Energy parameters", 35).c_str());
    std::fprintf(f, "                               Mean energy = %#.14e (eV) ; Rel. std: %-9.3e%%\n", s.get("meanEnergy"),
                 s.get("meanEnergyError") / s.get("meanEnergy") * 100.0);
    std::fprintf(f, "                      Electron temperature = %#.14e (eV) ; Rel. std: %-9.3e%%\n", s.get("Te"),
                 s.get("TeError") / s.get("Te") * 100.0);
    std::fprintf(f, "\n%s\n\n", banner(27, "Parameters obtained from the EEDF", 27).c_str());
    std::fprintf(f, "                    Ionization coefficient = %#.14e (m-3)\n", s.get("totalIonRateCoeff"));
    std::fprintf(f, "                    Attachment coefficient = %#.14e (m-3)\n", s.get("totalAttRateCoeff"));
    std::fprintf(f, "               Momentum-transfer frequency = %#.14e (s-1)\n", s.get("momTransfFreq"));
    std::fprintf(f, "               Energy-relaxation frequency = %#.14e (s-1)\n", s.get("energyRelaxFreq"));
    std::fprintf(f, "      Reduced energy diffusion coefficient = %#.14e (eV/(ms))\n", s.get("redDiffCoeffEnergy_eedf"));
    std::fprintf(f, "                   Reduced energy mobility = %#.14e (eV/(msV))\n", s.get("redMobCoeffEnergy_eedf"));
    std::fprintf(f, "             Reduced diffusion coefficient = %#.14e (1/(ms))\n", s.get("redDiffCoeff_eedf"));
    std::fprintf(f, "        DC non-magnetized reduced mobility = %#.14e (1/(msV))\n", s.get("redMobCoeff_DC_eedf"));
    std::fprintf(f, "                     Characteristic energy = %#.14e (eV)\n\n", s.get("characEnergy_eedf"));
}

void writePower(const std::string& dir, const Power& pw) {
    File out(dir + "/powerBalance.txt");
    FILE* f = out.f;
    auto t = pw.total;
    const std::string dash(73, '-');
    std::fprintf(f, "                               Field = %#+.14e (eVm3/s)\n", t["field"]);
    std::fprintf(f, "           Elastic collisions (gain) = %#+.14e (eVm3/s)\n", t["elasticGain"]);
    std::fprintf(f, "           Elastic collisions (loss) = %#+.14e (eVm3/s)\n", t["elasticLoss"]);
    std::fprintf(f, "                          CAR (gain) = %#+.14e (eVm3/s)\n", t["carGain"]);
    std::fprintf(f, "                          CAR (loss) = %#+.14e (eVm3/s)\n", t["carLoss"]);
    std::fprintf(f, "     Excitation inelastic collisions = %#+.14e (eVm3/s)\n", t["excitationIne"]);
    std::fprintf(f, "  Excitation superelastic collisions = %#+.14e (eVm3/s)\n", t["excitationSup"]);
    std::fprintf(f, "    Vibrational inelastic collisions = %#+.14e (eVm3/s)\n", t["vibrationalIne"]);
    std::fprintf(f, " Vibrational superelastic collisions = %#+.14e (eVm3/s)\n", t["vibrationalSup"]);
    std::fprintf(f, "     Rotational inelastic collisions = %#+.14e (eVm3/s)\n", t["rotationalIne"]);
    std::fprintf(f, "  Rotational superelastic collisions = %#+.14e (eVm3/s)\n", t["rotationalSup"]);
    std::fprintf(f, "               Ionization collisions = %#+.14e (eVm3/s)\n", t["ionizationIne"]);
    std::fprintf(f, "               Attachment collisions = %#+.14e (eVm3/s)\n", t["attachmentIne"]);
    std::fprintf(f, "             Electron density growth = %#+.14e (eVm3/s) +\n", t["eDensGrowth"]);
    std::fprintf(f, " %s\n", dash.c_str());
    std::fprintf(f, "                       Power Balance = %#+.14e (eVm3/s)\n", t["balance"]);
    std::fprintf(f, "              Relative Power Balance = %#.14e%%\n\n", t["relativeBalance"] * 100.0);
    std::fprintf(f, "           Elastic collisions (gain) = %#+.14e (eVm3/s)\n", t["elasticGain"]);
    std::fprintf(f, "           Elastic collisions (loss) = %#+.14e (eVm3/s) +\n", t["elasticLoss"]);
    std::fprintf(f, " %s\n", dash.c_str());
    std::fprintf(f, "            Elastic collisions (net) = %#+.14e (eVm3/s)\n\n", t["elasticNet"]);
    std::fprintf(f, "                          CAR (gain) = %#+.14e (eVm3/s)\n", t["carGain"]);
    std::fprintf(f, "                          CAR (gain) = %#+.14e (eVm3/s) +\n", t["carLoss"]);
    std::fprintf(f, " %s\n", dash.c_str());
    std::fprintf(f, "                           CAR (net) = %#+.14e (eVm3/s)\n\n", t["carNet"]);
    std::fprintf(f, "     Excitation inelastic collisions = %#+.14e (eVm3/s)\n", t["excitationIne"]);
    std::fprintf(f, "  Excitation superelastic collisions = %#+.14e (eVm3/s) +\n", t["excitationSup"]);
    std::fprintf(f, " %s\n", dash.c_str());
    std::fprintf(f, "         Excitation collisions (net) = %#+.14e (eVm3/s)\n\n", t["excitationNet"]);
    std::fprintf(f, "    Vibrational inelastic collisions = %#+.14e (eVm3/s)\n", t["vibrationalIne"]);
    std::fprintf(f, " Vibrational superelastic collisions = %#+.14e (eVm3/s) +\n", t["vibrationalSup"]);
    std::fprintf(f, " %s\n", dash.c_str());
    std::fprintf(f, "        Vibrational collisions (net) = %#+.14e (eVm3/s)\n\n", t["vibrationalNet"]);
    std::fprintf(f, "     Rotational inelastic collisions = %#+.14e (eVm3/s)\n", t["rotationalIne"]);
    std::fprintf(f, "  Rotational superelastic collisions = %#+.14e (eVm3/s) +\n", t["rotationalSup"]);
    std::fprintf(f, " %s\n", dash.c_str());
    std::fprintf(f, "         Rotational collisions (net) = %#+.14e (eVm3/s)\n", t["rotationalNet"]);
    for (auto kv : pw.gas) {
        auto& g = kv.second;
        const int right = 39 - (int)kv.first.size();
        std::fprintf(f, "\n%s\n\n", banner(37, kv.first, right > 0 ? right : 0).c_str());
        std::fprintf(f, "     Excitation inelastic collisions = %#+.14e (eVm3/s)\n", g["excitationIne"]);
        std::fprintf(f, "  Excitation superelastic collisions = %#+.14e (eVm3/s) +\n", g["excitationSup"]);
        std::fprintf(f, " %s\n", dash.c_str());
        std::fprintf(f, "         Excitation collisions (net) = %#+.14e (eVm3/s)\n\n", g["excitationNet"]);
        std::fprintf(f, "    Vibrational inelastic collisions = %#+.14e (eVm3/s)\n", g["vibrationalIne"]);
        std::fprintf(f, " Vibrational superelastic collisions = %#+.14e (eVm3/s) +\n", g["vibrationalSup"]);
        std::fprintf(f, " %s\n", dash.c_str());
        std::fprintf(f, "        Vibrational collisions (net) = %#+.14e (eVm3/s)\n\n", g["vibrationalNet"]);
        std::fprintf(f, "     Rotational inelastic collisions = %#+.14e (eVm3/s)\n", g["rotationalIne"]);
        std::fprintf(f, "  Rotational superelastic collisions = %#+.14e (eVm3/s) +\n", g["rotationalSup"]);
        std::fprintf(f, " %s\n", dash.c_str());
        std::fprintf(f, "         Rotational collisions (net) = %#+.14e (eVm3/s)\n\n", g["rotationalNet"]);
        std::fprintf(f, "               Ionization collisions = %#+.14e (eVm3/s)\n", g["ionizationIne"]);
        std::fprintf(f, "               Attachment collisions = %#+.14e (eVm3/s)\n", g["attachmentIne"]);
    }
}

// Rows of the rate-coefficient files (Output.h:419-538): the collisions of
// the flights, then the Effective ones, then (in their own section) the extra
// ones, each with the reference's collision ID.
void rateRow(FILE* f, const GasMixture& m, int c, double ine, double sup) {
    if (!m.collisionIsReverse[c])
        std::fprintf(f, "%4d %20.14e (N/A)                %s\n", m.collisionRefId[c] + 1, ine, m.collisionDescription[c].c_str());
    else
        std::fprintf(f, "%4d %20.14e %20.14e %s\n", m.collisionRefId[c] + 1, ine, sup, m.collisionDescription[c].c_str());
}

bool hasExtra(const GasMixture& m) {
    for (int k : m.collisionKind)
        if (k == 2) return true;
    return false;
}

// The "# ID Description" block and the column header of a periodic or
// look-up rate-coefficient file (Output.h:478-507, 620-640).
std::string rateColumns(FILE* f, const GasMixture& m, std::string header) {
    std::fprintf(f, "%s\n# %-76s #\n", std::string(80, '#').c_str(), "ID   Description");
    auto column = [&](int c) {
        const int id = m.collisionRefId[c] + 1;
        std::fprintf(f, "# %-4d %-71s #\n", id, m.collisionDescription[c].c_str());
        std::string a = "R" + std::to_string(id) + "_ine(m3/s)";
        header += a + std::string(a.size() < 22 ? 22 - a.size() : 1, ' ');
        if (m.collisionIsReverse[c]) {
            a = "R" + std::to_string(id) + "_sup(m3/s)";
            header += a + std::string(a.size() < 22 ? 22 - a.size() : 1, ' ');
        }
    };
    const int nc = (int)m.collisionDescription.size();
    for (int c = 0; c < nc; ++c)
        if (m.collisionKind[c] != 2) column(c);
    std::fprintf(f, "#%s#\n# %-76s #\n#%s#\n# %-76s #\n", std::string(78, ' ').c_str(), "*** Extra rate coefficients ***",
                 std::string(78, '#').c_str(), "ID   Description");
    for (int c = 0; c < nc; ++c)
        if (m.collisionKind[c] == 2) column(c);
    std::fprintf(f, "%s\n\n%s\n", std::string(80, '#').c_str(), header.c_str());
    return header;
}

void rateValues(FILE* f, const GasMixture& m, const std::vector<double>& ine, const std::vector<double>& sup) {
    const int nc = (int)m.collisionDescription.size();
    for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < nc; ++c) {
            if ((m.collisionKind[c] == 2) != (pass == 1)) continue;
            std::fprintf(f, "%-21.14e ", ine[c]);
            if (m.collisionIsReverse[c]) std::fprintf(f, "%-21.14e ", sup[c]);
        }
}

void writeRates(const std::string& dir, const JobData& j, const GasMixture& m) {
    std::vector<double> ine, sup;
    eedfRateCoefficients(j, m, ine, sup);
    const int nc = (int)m.collisionDescription.size();
    {
        File out(dir + "/rateCoefficients.txt");
        std::fprintf(out.f, " ID  Ine.R.Coeff.(m3/s)   Sup.R.Coeff.(m3/s)   Description\n");
        for (int c = 0; c < nc; ++c)
            if (m.collisionKind[c] != 2) rateRow(out.f, m, c, ine[c], sup[c]);
        if (hasExtra(m)) {
            std::fprintf(out.f, "\n%s\n* Extra Rate Coefficients *\n%s\n\n", std::string(27, '*').c_str(), std::string(27, '*').c_str());
            for (int c = 0; c < nc; ++c)
                if (m.collisionKind[c] == 2) rateRow(out.f, m, c, ine[c], sup[c]);
        }
    }
    {
        File out(dir + "/rateCoefficientsMC.txt");
        std::fprintf(out.f, " ID  Ine.R.Coeff.(m3/s)   Sup.R.Coeff.(m3/s)   Description\n");
        for (int c = 0; c < nc; ++c) {
            if (m.collisionKind[c] != 0) continue;
            const int p = m.collisionFirstProcess[c];
            rateRow(out.f, m, c, j.rateMC[p], m.collisionIsReverse[c] ? j.rateMC[p + 1] : 0.0);
        }
    }
    const double omega = j.cond.excitation_frequency * 2.0 * kPi;
    if (omega == 0) return;
    // rate coefficients of the EEDF of each phase (Output.h:468-535)
    File out(dir + "/rateCoefficients_periodic.txt");
    std::string header;
    for (const char* v : {"Phase(rad)", "Phase(s)", "E/N(Td)"}) header += std::string(v) + std::string(22 - std::string(v).size(), ' ');
    rateColumns(out.f, m, header);
    const double EN = j.cond.reduced_elec_field;
    const size_t ncell = j.eedf.size();
    for (int ph = 0; ph < j.nPhases; ++ph) {
        std::vector<double> pi, ps;
        eedfRateCoefficients(j, m, pi, ps, j.eedfPeriodic.data() + (size_t)ph * ncell);
        std::fprintf(out.f, "%-21.14e %-21.14e %-21.14e ", j.phases[ph], j.phases[ph] / omega, std::sqrt(2.0) * EN * std::cos(j.phases[ph]));
        rateValues(out.f, m, pi, ps);
        std::fprintf(out.f, "\n");
    }
}

void writeTemporal(const std::string& dir, const JobData& j) {
    File out(dir + "/MCTemporalInfo.txt");
    const char* vars[] = {"Time(s)", "MeanEnergy(eV)", "xPos(m)", "yPos(m)", "zPos(m)", "xSqWidth(m2)", "ySqWidth(m2)",
                          "zSqWidth(m2)", "xVel(m/s)", "yVel(m/s)", "zVel(m/s)"};
    std::string header;
    for (const char* v : vars) header += std::string(v) + std::string(20 - std::string(v).size(), ' ');
    std::fprintf(out.f, "%s\n", header.c_str());
    for (int i = 0; i < j.nSeries; ++i)
        std::fprintf(out.f, "%-19.10e %-19.10e %-19.10e %-19.10e %-19.10e %-19.10e %-19.10e %-19.10e %-19.10e %-19.10e %-19.10e \n",
                     j.times[i], j.meanEnergy[i], j.meanPosition[3 * i], j.meanPosition[3 * i + 1], j.meanPosition[3 * i + 2],
                     j.positionCovariance[9 * i], j.positionCovariance[9 * i + 4], j.positionCovariance[9 * i + 8],
                     j.meanVelocity[3 * i], j.meanVelocity[3 * i + 1], j.meanVelocity[3 * i + 2]);
}

void writeDetails(const std::string& dir, const JobData& j) {
    File out(dir + "/MCSimDetails.txt");
    FILE* f = out.f;
    const ekmc_results& r = j.res;
    const double real = (double)r.real_collisions, nul = (double)r.null_collisions;
    std::fprintf(f, "                          number of electrons: %e\n", j.nElectrons);
    std::fprintf(f, "                      trialCollisionFrequency: %e s-1\n\n", r.trial_collision_frequency);
    std::fprintf(f, "                        final simulation time: %e s\n", r.final_time);
    std::fprintf(f, "                            steady-state time: %e s\n", r.steady_state_time);
    std::fprintf(f, "                 number of integration points: %d\n\n", r.n_integration_points);
    std::fprintf(f, "********************************** Collisions ************************************\n\n");
    std::fprintf(f, "number of real collisions before steady-state: %e\n", (double)r.real_collisions_at_ss);
    std::fprintf(f, " number of real collisions after steady-state: %e\n", real - (double)r.real_collisions_at_ss);
    std::fprintf(f, "----------------------------------------------------------------------------------\n");
    std::fprintf(f, "              total number of real collisions: %e\n\n", real);
    std::fprintf(f, "number of null collisions before steady-state: %e\n", (double)r.null_collisions_at_ss);
    std::fprintf(f, " number of null collisions after steady-state: %e\n", nul - (double)r.null_collisions_at_ss);
    std::fprintf(f, "----------------------------------------------------------------------------------\n");
    std::fprintf(f, "              total number of null collisions: %e\n\n", nul);
    std::fprintf(f, "                  fraction of real collisions: %e%%\n", real / (real + nul) * 100.0);
    std::fprintf(f, "                  fraction of null collisions: %e%%\n", nul / (real + nul) * 100.0);
    std::fprintf(f, "----------------------------------------------------------------------------------\n");
    std::fprintf(f, "                                        total: %e%%\n\n", 100.0);
    std::fprintf(f, "************************** Simulation relative errors ****************************\n\n");
    std::fprintf(f, "                                  Mean energy: %.4e\n", r.mean_energy_error / r.mean_energy);
    double rel[3], rd[3];
    for (int k = 0; k < 3; ++k) {
        rel[k] = r.flux_drift_velocity_error[k] / std::fabs(r.flux_drift_velocity[k]);
        rd[k] = r.flux_diffusion_error[4 * k] / std::fabs(r.flux_diffusion[4 * k]);
    }
    std::fprintf(f, "                          Flux drift velocity: %.4e %.4e %.4e\n", rel[0], rel[1], rel[2]);
    std::fprintf(f, "                  Flux diffusion coefficients: %.4e %.4e %.4e\n", rd[0], rd[1], rd[2]);
    std::fprintf(f, "                                Power balance: %.4e\n", r.power_balance_rel_error);
    std::fprintf(f, "**********************************************************************************\n\n");
    std::fprintf(f, "                                 Elapsed time: %e s\n", r.elapsed_time);
}

}  // namespace

// Output.h:754-782
void writeTemporalPeriodic(const std::string& dir, const JobData& j) {
    const double omega = j.cond.excitation_frequency * 2.0 * kPi;
    if (omega == 0) return;
    File out(dir + "/MCTemporalInfo_periodic.txt");
    const char* vars[] = {"Phase(rad)", "Phase(s)", "E/N(Td)", "MeanEnergy(eV)", "FluxV_x(m/s)", "FluxV_y(m/s)", "FluxV_z(m/s)",
                          "BulkV_x(m/s)", "BulkV_y(m/s)", "BulkV_z(m/s)", "FluxND_xx(1/(ms))", "FluxND_xy(1/(ms))",
                          "FluxND_xz(1/(ms))", "FluxND_yx(1/(ms))", "FluxND_yy(1/(ms))", "FluxND_yz(1/(ms))",
                          "FluxND_zx(1/(ms))", "FluxND_zy(1/(ms))", "FluxND_zz(1/(ms))", "BulkND_xx(1/(ms))",
                          "BulkND_xy(1/(ms))", "BulkND_xz(1/(ms))", "BulkND_yx(1/(ms))", "BulkND_yy(1/(ms))",
                          "BulkND_yz(1/(ms))", "BulkND_zx(1/(ms))", "BulkND_zy(1/(ms))", "BulkND_zz(1/(ms))"};
    std::string header;
    for (const char* v : vars) header += std::string(v) + std::string(20 - std::string(v).size(), ' ');
    std::fprintf(out.f, "%s\n", header.c_str());
    const double EN = j.cond.reduced_elec_field, N = j.cond.gas_density;
    for (int ph = 0; ph < j.nPhases; ++ph) {
        std::vector<double> row = {j.phases[ph], j.phases[ph] / omega, std::sqrt(2.0) * EN * std::cos(j.phases[ph]),
                                   j.meanEnergyPeriodic[ph]};
        for (int a = 0; a < 3; ++a) row.push_back(j.fluxVelocityPeriodic[3 * ph + a]);
        for (int a = 0; a < 3; ++a) row.push_back(j.bulkVelocityPeriodic[3 * ph + a]);
        for (int a = 0; a < 9; ++a) row.push_back(N * j.fluxDiffusionPeriodic[9 * ph + a]);
        for (int a = 0; a < 9; ++a) row.push_back(N * j.bulkDiffusionPeriodic[9 * ph + a]);
        for (size_t k = 0; k < row.size(); ++k) std::fprintf(out.f, k + 1 < row.size() ? "%-19.10e " : "%-19.10e\n", row[k]);
    }
}

void writeJobFiles(const std::string& dir, const JobData& j, const GasMixture& m, const OutputOptions& o) {
    if (wants(o, "eedf")) writeEedf(dir, j);
    if (wants(o, "evdf") && j.cylindrical) writeEvdf(dir, j);
    if (wants(o, "powerBalance")) writePower(dir, powerOf(j, m));
    if (wants(o, "swarmParameters")) writeSwarm(dir, j, swarmParameters(j, m));
    if (wants(o, "rateCoefficients")) writeRates(dir, j, m);
    if (wants(o, "MCTemporalInfo")) writeTemporal(dir, j);
    if (wants(o, "MCTemporalInfo_periodic")) writeTemporalPeriodic(dir, j);
    if (wants(o, "MCSimDetails")) writeDetails(dir, j);
}

void writeLookUpTables(const std::string& dir, const std::vector<JobData>& jobs, const GasMixture& m,
                       const OutputOptions& o) {
    // Output.h:540-728 (boltzmannMC branch), written once all jobs are done
    if (!wants(o, "lookUpTable") || jobs.size() < 2) return;
    std::string varCond = "RedElecField(Td)      ";
    if (o.variableCondition == "reducedMagField") varCond = "RedMagField(Hx)       ";
    else if (o.variableCondition == "elecFieldAngle") varCond = "elecFieldAngle(degr)  ";
    else if (o.variableCondition == "excitationFrequency") varCond = "ExcitationFreq(Hz)    ";
    auto var = [&](const JobData& j) {
        if (o.variableCondition == "reducedMagField") return j.cond.reduced_mag_field;
        if (o.variableCondition == "elecFieldAngle") return j.cond.elec_field_angle;
        if (o.variableCondition == "excitationFrequency") return j.cond.excitation_frequency;
        return j.cond.reduced_elec_field;
    };
    auto pad = [](const std::string& s, size_t w) { return s + std::string(s.size() < w ? w - s.size() : 1, ' '); };
    File swarm(dir + "/lookUpTableSwarm.txt"), power(dir + "/lookUpTablePower.txt"), rate(dir + "/lookUpTableRateCoeff.txt"),
        eedf(dir + "/lookUpTableEedf.txt"), grid(dir + "/lookUpTableEGrid.txt");
    {
        const char* vars[] = {"FluxND_xx(1/(ms))", "FluxND_xy(1/(ms))", "FluxND_xz(1/(ms))", "FluxND_yx(1/(ms))", "FluxND_yy(1/(ms))",
                              "FluxND_yz(1/(ms))", "FluxND_zx(1/(ms))", "FluxND_zy(1/(ms))", "FluxND_zz(1/(ms))", "FluxV_x(m/s)",
                              "FluxV_y(m/s)", "FluxV_z(m/s)", "BulkND_xx(1/(ms))", "BulkND_xy(1/(ms))", "BulkND_xz(1/(ms))",
                              "BulkND_yx(1/(ms))", "BulkND_yy(1/(ms))", "BulkND_yz(1/(ms))", "BulkND_zx(1/(ms))", "BulkND_zy(1/(ms))",
                              "BulkND_zz(1/(ms))", "BulkV_x(m/s)", "BulkV_y(m/s)", "BulkV_z(m/s)", "SSTVelocity(m/s)",
                              "SSTRedTow(m2)", "SSTRedAtt(m2)", "MeanE(eV)", "EleTemp(eV)", "IonCoeff(m3/s)", "AttCoeff(m3/s)",
                              "MomTransfFreq(1/s)", "EnergyRelFreq(1/s)", "RedDiffE_eedf(eV/(ms))", "RedMobE_eedf(eV/(msV))",
                              "RedDiff_eedf(1/(ms))", "RedMob_DC_eedf(1/(msV))", "CharacEnergy_eedf(eV)", "trialCollFreq(1/s)"};
        std::string h = varCond;
        for (const char* v : vars) h += pad(v, 24);
        std::fprintf(swarm.f, "%s\n", h.c_str());
        const char* pvars[] = {"PowerField(eVm3/s)", "PwrElaGain(eVm3/s)", "PwrElaLoss(eVm3/s)", "PwrElaNet(eVm3/s)",
                               "PwrEleGain(eVm3/s)", "PwrEleLoss(eVm3/s)", "PwrEleNet(eVm3/s)", "PwrVibGain(eVm3/s)",
                               "PwrVibLoss(eVm3/s)", "PwrVibNet(eVm3/s)", "PwrRotGain(eVm3/s)", "PwrRotLoss(eVm3/s)",
                               "PwrRotNet(eVm3/s)", "PwrIon(eVm3/s)", "PwrAtt(eVm3/s)", "PwrGrowth(eVm3/s)",
                               "PwrBalance(eVm3/s)", "RelPwrBalance"};
        h = varCond;
        for (const char* v : pvars) h += pad(v, 22);
        std::fprintf(power.f, "%s\n", h.c_str());
        rateColumns(rate.f, m, varCond);
    }
    for (const JobData& j : jobs) {
        const double cv = var(j);
        const SwarmParams s = swarmParameters(j, m);
        const ekmc_results& r = j.res;
        const double N = j.cond.gas_density;
        std::vector<double> row;
        for (int k = 0; k < 9; ++k) row.push_back(r.flux_diffusion[k] * N);
        for (int k = 0; k < 3; ++k) row.push_back(r.flux_drift_velocity[k]);
        for (int k = 0; k < 9; ++k) row.push_back(r.bulk_diffusion[k] * N);
        for (int k = 0; k < 3; ++k) row.push_back(r.bulk_drift_velocity[k]);
        for (const char* k : {"effSSTAverageVelocity", "effSSTRedTownsendCoeff", "effSSTRedAttCoeff", "meanEnergy", "Te",
                              "totalIonRateCoeff", "totalAttRateCoeff", "momTransfFreq", "energyRelaxFreq",
                              "redDiffCoeffEnergy_eedf", "redMobCoeffEnergy_eedf", "redDiffCoeff_eedf",
                              "redMobCoeff_DC_eedf", "characEnergy_eedf"})
            row.push_back(s.get(k));
        row.push_back(r.trial_collision_frequency);
        std::fprintf(swarm.f, "%-21.14e ", cv);
        for (double v : row) std::fprintf(swarm.f, "%-23.14e ", v);
        std::fprintf(swarm.f, "\n");

        Power p = powerOf(j, m);
        std::fprintf(power.f, "%-21.14e ", cv);
        for (const char* k : {"field", "elasticGain", "elasticLoss", "elasticNet", "excitationSup", "excitationIne",
                              "excitationNet", "vibrationalSup", "vibrationalIne", "vibrationalNet", "rotationalSup",
                              "rotationalIne", "rotationalNet", "ionizationIne", "attachmentIne", "eDensGrowth", "balance"})
            std::fprintf(power.f, "%-21.14e ", p.total[k]);
        std::fprintf(power.f, "%19.14e%%\n", p.total["relativeBalance"] * 100.0);

        std::vector<double> ine, sup;
        eedfRateCoefficients(j, m, ine, sup);
        std::fprintf(rate.f, "%-21.14e ", cv);
        rateValues(rate.f, m, ine, sup);
        std::fprintf(rate.f, "\n");

        std::fprintf(eedf.f, "%-20.14e ", cv);
        std::fprintf(grid.f, "%-20.14e ", cv);
        for (size_t i = 0; i < j.energyCells.size(); ++i) {
            std::fprintf(eedf.f, "%-20.14e ", j.eedf[i]);
            std::fprintf(grid.f, "%-20.14e ", j.energyCells[i]);
        }
        std::fprintf(eedf.f, "\n");
        std::fprintf(grid.f, "\n");
    }
}

}  // namespace lokimc
