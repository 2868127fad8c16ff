// tables.cpp -- host FP64 process list and null-collision tables.
#include "tables.h"

#include <cmath>
#include "common.h"

namespace ekmc {

std::string build_process_list(const ekmc_mixture& mix, ProcessList& pl) {
    if (mix.n_processes <= 0 || mix.n_gases <= 0) return "mixture has no processes or no gases";
    const int P = mix.n_processes;
    pl = ProcessList();
    pl.P = P;
    pl.G = mix.n_gases;
    pl.type.resize(P); pl.isElastic.resize(P); pl.isSuper.resize(P); pl.gas.resize(P); pl.scat.resize(P);
    pl.threshold.resize(P); pl.energyLoss.resize(P); pl.tMass.resize(P); pl.rMass.resize(P);
    pl.relDens.resize(P); pl.statW.resize(P); pl.wParam.resize(P); pl.params.resize(2 * P);
    pl.minE.resize(P); pl.maxE.resize(P); pl.energy.resize(P); pl.sigma.resize(P);
    for (int p = 0; p < P; ++p) {
        pl.type[p] = mix.type[p];
        if (pl.type[p] < 0 || pl.type[p] > 2) return "invalid process type";
        pl.isElastic[p] = mix.is_elastic[p];
        pl.isSuper[p] = mix.is_superelastic[p];
        pl.gas[p] = mix.gas_index[p];
        if (pl.gas[p] < 0 || pl.gas[p] >= pl.G) return "gas index out of range";
        if (p > 0 && pl.gas[p] < pl.gas[p - 1]) return "processes of a gas must be contiguous";
        pl.scat[p] = mix.scattering[p];
        if (pl.scat[p] < 0 || pl.scat[p] > 5) return "invalid angular scattering model";
        if (pl.scat[p] == EKMC_SCAT_MOMENTUM_CONSERVATION_IONIZATION && pl.type[p] != EKMC_IONIZATION)
            return "momentumConservationIonization assigned to a process that is not Ionization";
        pl.params[2 * p] = mix.scattering_params[2 * p];
        pl.params[2 * p + 1] = mix.scattering_params[2 * p + 1];
        pl.threshold[p] = mix.threshold[p];
        pl.tMass[p] = mix.target_mass[p];
        if (!(pl.tMass[p] > 0)) return "target masses must be positive";
        pl.rMass[p] = kMass * pl.tMass[p] / (kMass + pl.tMass[p]);
        pl.relDens[p] = mix.rel_density[p];
        pl.statW[p] = mix.stat_weight_ratio[p];
        pl.wParam[p] = mix.w_parameter[p];
        if (pl.isSuper[p]) {
            // superelastic partner of process p-1 (BoltzmannMC.C:209-266)
            if (p == 0 || pl.isSuper[p - 1] || pl.gas[p] != pl.gas[p - 1])
                return "a superelastic process must follow its inelastic partner";
            pl.energyLoss[p] = -pl.threshold[p];
            pl.minE[p] = 0;
            pl.maxE[p] = pl.maxE[p - 1] - pl.threshold[p];
            continue;
        }
        pl.energyLoss[p] = pl.threshold[p];
        const int b = mix.table_offset[p], e = mix.table_offset[p + 1];
        if (e - b < 1) return "a forward process needs a cross-section table";
        std::vector<double> en(mix.table_energy + b, mix.table_energy + e);
        std::vector<double> sg(mix.table_sigma + b, mix.table_sigma + e);
        for (size_t i = 1; i < en.size(); ++i)
            if (!(en[i - 1] < en[i])) return "cross-section energies must be strictly increasing";
        if (pl.isElastic[p] && en[0] != 0) { en.insert(en.begin(), 0.0); sg.insert(sg.begin(), sg[0]); }
        const double thr = pl.threshold[p];
        size_t drop = 0;
        while (drop < en.size() && en[drop] < thr) ++drop;
        en.erase(en.begin(), en.begin() + drop);
        sg.erase(sg.begin(), sg.begin() + drop);
        if (en.empty() || en[0] != thr) { en.insert(en.begin(), thr); sg.insert(sg.begin(), 0.0); }
        if (en.size() < 2) return "a cross-section table needs two points above threshold";
        pl.minE[p] = thr;
        pl.maxE[p] = en.back();
        if (pl.isElastic[p]) pl.energyMaxElastic = std::fmin(pl.energyMaxElastic, pl.maxE[p]);
        pl.energy[p] = std::move(en);
        pl.sigma[p] = std::move(sg);
    }
    pl.gasFirst.assign(pl.G, -1);
    pl.gasLast.assign(pl.G, -1);
    pl.gasFrac.assign(mix.gas_fraction, mix.gas_fraction + pl.G);
    for (int p = 0; p < P; ++p) {
        int g = pl.gas[p];
        if (pl.gasFirst[g] < 0) pl.gasFirst[g] = p;
        pl.gasLast[g] = p;
    }
    for (int g = 0; g < pl.G; ++g)
        if (pl.gasFirst[g] < 0) return "every gas needs at least one process";
    return "";
}

namespace {

// Linear interpolation of an ascending table at ascending query points, as
// gsl_interp_linear with a gsl_interp_accel cursor evaluates it: the
// interval is the one with x[i] <= v < x[i+1] (the last point belongs to the
// last interval), and y = y_i + (v - x_i)/dx*(y_{i+1} - y_i).
struct Cursor {
    const std::vector<double>& x;
    const std::vector<double>& y;
    size_t i = 0;
    Cursor(const std::vector<double>& x_, const std::vector<double>& y_) : x(x_), y(y_) {}
    double at(double v) {
        const size_t last = x.size() - 2;
        if (v < x[i]) i = 0;
        while (i < last && x[i + 1] <= v) ++i;
        double dx = x[i + 1] - x[i];
        return dx > 0.0 ? y[i] + (v - x[i]) / dx * (y[i + 1] - y[i]) : 0.0;
    }
};

}  // namespace

void build_tables(const ProcessList& pl, double N, int I, double maxEnergy, Tables& t) {
    const int P = pl.P;
    t.I = I;
    t.P = P;
    t.maxEnergy = maxEnergy;
    t.step = maxEnergy / (double)(I - 1);
    t.interp.assign((size_t)I * P, 0.0);
    t.cumul.assign((size_t)I * P, 0.0);
    t.nutot.assign(I, 0.0);
    t.numax.assign(I, 0.0);
    std::vector<Cursor> cur;
    cur.reserve(P);
    for (int p = 0; p < P; ++p) cur.emplace_back(pl.isSuper[p] ? pl.energy[p - 1] : pl.energy[p],
                                                 pl.isSuper[p] ? pl.sigma[p - 1] : pl.sigma[p]);
    double running = 0;
    for (int i = 0; i < I; ++i) {
        const double energy = i * t.step;
        double acc = 0;
        double* irow = &t.interp[(size_t)i * P];
        double* crow = &t.cumul[(size_t)i * P];
        for (int p = 0; p < P; ++p) {
            double value = 0;
            if (pl.isSuper[p]) {
                // Klein-Rosseland: (g_i/g_f)(1 + E_th/u) sigma_if(u + E_th)
                if (energy > pl.minE[p] && energy <= pl.maxE[p]) {
                    const double thr = pl.minE[p - 1];
                    value = pl.statW[p] * (1.0 + thr / energy) * cur[p].at(energy + thr) * pl.relDens[p];
                }
            } else if (energy >= pl.minE[p] && energy <= pl.maxE[p]) {
                value = cur[p].at(energy) * pl.relDens[p];
            }
            irow[p] = value;
            acc += value;
            crow[p] = acc;
        }
        acc *= N * std::sqrt(energy * 2.0 * kCharge / kMass);
        t.nutot[i] = acc;
        running = std::fmax(acc, running);
        t.numax[i] = running;
    }
}

}  // namespace ekmc
