// tables.h -- host-side process list and collision tables (FP64).
#pragma once
#include <string>
#include <vector>
#include "../../include/ekmc.h"

namespace ekmc {

// The flattened process list of BoltzmannMC::allocateEvaluateVariablesFirstTime
// (BoltzmannMC.C:29-271) with each forward process's integral cross section
// prepared as the reference holds it: Elastic tables start at 0 eV
// (Collision.C:41-44) and every table starts at its threshold with a zero
// (BoltzmannMC.C:139-150).
struct ProcessList {
    int P = 0, G = 0;
    std::vector<int> type, isElastic, isSuper, gas, scat;
    std::vector<double> threshold, energyLoss, tMass, rMass, relDens, statW, wParam, params;
    std::vector<double> minE, maxE;
    std::vector<std::vector<double>> energy, sigma;  // forward processes only
    std::vector<int> gasFirst, gasLast;
    std::vector<double> gasFrac;
    double energyMaxElastic = 1e100;
};

// Validates and flattens an ekmc_mixture; returns an error text or "".
std::string build_process_list(const ekmc_mixture& mix, ProcessList& out);

// interpolateCrossSections (BoltzmannMC.C:561-615) on an I-point grid up to
// maxEnergy for gas density N.  interp/cumul are [I x P] row-major.
struct Tables {
    int I = 0, P = 0;
    double maxEnergy = -1, step = 0;
    std::vector<double> interp, cumul, nutot, numax;
};
void build_tables(const ProcessList& pl, double N, int I, double maxEnergy, Tables& t);

}  // namespace ekmc
