// eedf_rates_probe.cpp -- test harness (not part of the product): the
// EEDF-convolution rate coefficients of the host post-processing
// (lokimc::eedfRateCoefficients, after Collision::evaluateRateCoeff,
// Collision.C:355-397) for the mixture of a setup file and a Maxwellian EEDF
// at temperature kT [eV] on n cells of width `step` [eV].
//   eedf_rates_probe <setup> <inputDir> <kT> <step> <n>
// prints one line per collision: description, inelastic and superelastic
// rate coefficients [m^3/s] (%.17g; nan when the collision is not reversible).
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "gas_model.h"
#include "outputs.h"
#include "setup_file.h"

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s setup inputDir kT step n\n", argv[0]);
        return 2;
    }
    using namespace lokimc;
    const SetupFile setup = SetupFile::fromFile(argv[1]);
    const GasMixture mix = buildMixture(setup, argv[2], evalNumber(setup.value("workingConditions.gasTemperature")));
    const double kT = std::atof(argv[3]), step = std::atof(argv[4]);
    const int n = std::atoi(argv[5]);
    JobData j;
    j.energyCells.resize(n);
    j.eedf.resize(n);
    const double norm = 2.0 / std::sqrt(3.14159265358979323846) / std::pow(kT, 1.5);
    for (int i = 0; i < n; ++i) {
        j.energyCells[i] = (i + 0.5) * step;
        j.eedf[i] = norm * std::exp(-j.energyCells[i] / kT);
    }
    std::vector<double> ine, sup;
    eedfRateCoefficients(j, mix, ine, sup);
    for (size_t c = 0; c < ine.size(); ++c)
        std::printf("%s %.17g %.17g\n", mix.collisionDescription[c].c_str(), ine[c], sup[c]);
    return 0;
}
