// ekmc_oracle.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// CPU FP64 restatement of LoKI-MC v1.1.0's Monte Carlo loop, written from the
// reference sources (paths relative to Code/LoKI-MC/):
//   Sources/BoltzmannMC.C      process list, tables, flights, collisions,
//                              population control, sampling, control, averages
//   Sources/MathFunctions.C    Box-Muller, histograms, Euler rotation,
//                              cart2sph, batch-means error
//   Headers/AngularScatteringFunctions.h   cos(chi) samplers
//   Headers/Constant.h, Headers/GeneralDefinitions.h
//   Sources/Collision.C:41-44  zero-energy point of Elastic tables
// GSL's linear interpolation (gsl_interp_linear + gsl_interp_accel, called at
// BoltzmannMC.C:162-164,589,598) is restated from its published algorithm.
//
// The only intended difference from the reference is the variate source:
// slot-indexed uniforms (a tape, or Philox4x32-10 keyed by seed / electron
// stream / event / slot) instead of the unseeded thread-local mt19937_64 of
// MathFunctions.C:31-43.  Used by tests/ and bench.py's CPU-baseline leg only.

#include "ekmc_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Constant.h:6-23
constexpr double kBoltzmann = 1.38064852e-23;
constexpr double kCharge = 1.6021766208e-19;
constexpr double kMass = 9.10938356e-31;
constexpr double kNonDef = -123456789;
constexpr double kPi = M_PI;

// ---------------------------------------------------------------- Philox
// Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11: ten rounds of two 32x32->64 multiplies with
// Weyl-sequence key bumps between rounds.
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Variate slots (ekmc.h) and their intervals:
//   0: (0,1)   1-4: (0,1]   5: (0,1]   6: [0,1]   7: (0,1]
//   8: [0,1]   9: [0,1]     10: (0,1]
// A slot's 53-bit integer x comes from two 32-bit words (2h, 2h+1) of one
// Philox call of the event: x = ((w0 << 32) | w1) >> 11.  Slot -> (call, h):
// call 0 {0,5}, call 1 {6,7}, call 2 {1,2}, call 3 {3,4}, call 4 {8,9},
// call 5 {10}: the free time and the channel-choice variate share call 0.
enum Interval { OPEN_OPEN, OPEN_CLOSED, CLOSED_CLOSED, CLOSED_OPEN };
constexpr Interval kSlotInterval[EKMC_VARIATES_PER_EVENT] = {
    OPEN_OPEN, OPEN_CLOSED, OPEN_CLOSED, OPEN_CLOSED, OPEN_CLOSED, OPEN_CLOSED,
    CLOSED_CLOSED, OPEN_CLOSED, CLOSED_CLOSED, CLOSED_CLOSED, OPEN_CLOSED};

inline double to_unit(uint64_t x53, Interval iv) {
    const double s = 1.0 / 9007199254740992.0;  // 2^-53
    switch (iv) {
        case OPEN_OPEN: return ((double)x53 + 0.5) * s;
        case OPEN_CLOSED: return ((double)x53 + 1.0) * s;
        default: return (double)x53 * s;  // [0,1): the value 1 has probability 0 here
    }
}

enum Purpose : uint32_t { P_EVENT = 0, P_INIT = 1, P_POPCTL = 2, P_CHILD = 3, P_CLONE = 4 };

inline void philox_words(uint64_t seed, uint64_t stream, uint32_t event, uint32_t purpose,
                         uint32_t call, uint32_t out[4]) {
    uint32_t ctr[4] = {(uint32_t)stream, (uint32_t)(stream >> 32), event, (purpose << 16) | call};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    philox4x32_10(ctr, key, out);
}

constexpr int kSlotCall[EKMC_VARIATES_PER_EVENT] = {0, 2, 2, 3, 3, 0, 1, 1, 4, 4, 5};
constexpr int kSlotHalf[EKMC_VARIATES_PER_EVENT] = {0, 0, 1, 0, 1, 1, 0, 1, 0, 1, 0};

inline double slot_variate(uint64_t seed, uint64_t stream, uint32_t event, int slot) {
    uint32_t w[4];
    philox_words(seed, stream, event, P_EVENT, (uint32_t)kSlotCall[slot], w);
    int h = kSlotHalf[slot];
    uint64_t x = ((((uint64_t)w[2 * h]) << 32) | w[2 * h + 1]) >> 11;
    return to_unit(x, kSlotInterval[slot]);
}

inline uint64_t child_stream(uint64_t seed, uint64_t parent, uint32_t parent_event) {
    uint32_t w[4];
    philox_words(seed, parent, parent_event, P_CHILD, 0, w);
    return ((uint64_t)w[1] << 32) | w[0];
}

// A variate source: either Philox slots of (stream, event) or a tape row.
struct Variates {
    const double* tape = nullptr;   // [EKMC_VARIATES_PER_EVENT]
    uint64_t seed = 0, stream = 0;
    uint32_t event = 0;
    double u(int slot) const {
        if (tape) return tape[slot];
        return slot_variate(seed, stream, event, slot);
    }
};

// -------------------------------------------------------- GSL linear eval
// gsl_interp_linear evaluated through gsl_interp_accel_find/gsl_interp_bsearch:
// index = largest i with x[i] <= v (x[n-1] maps to n-2), then
// y = y_lo + (v - x_lo)/dx*(y_hi - y_lo), or 0 when dx is not positive.
double gsl_linear(const std::vector<double>& x, const std::vector<double>& y, double v) {
    size_t lo = 0, hi = x.size() - 1;
    while (hi > lo + 1) {
        size_t i = (hi + lo) / 2;
        if (x[i] > v) hi = i; else lo = i;
    }
    double xl = x[lo], xh = x[lo + 1], yl = y[lo], yh = y[lo + 1];
    double dx = xh - xl;
    return dx > 0.0 ? yl + (v - xl) / dx * (yh - yl) : 0.0;
}

// ------------------------------------------------------------ math helpers
struct Sph { double norm, sinT, cosT, sinP, cosP; };

// MathFunctions::cart2sph (MathFunctions.C:143-163)
bool cart2sph(const double v[3], Sph& s) {
    double x = v[0], y = v[1], z = v[2];
    double x2My2 = x * x + y * y;
    double nxy = std::sqrt(x2My2);
    s.norm = std::sqrt(x2My2 + z * z);
    if (s.norm == 0) return false;
    s.sinT = nxy / s.norm;
    s.cosT = z / s.norm;
    if (nxy != 0) { s.sinP = y / nxy; s.cosP = x / nxy; }
    else { s.sinP = 1; s.cosP = 0; }
    return true;
}

// MathFunctions::eulerTransformation (MathFunctions.C:129-141), Yousfi 1994
inline void euler(double sinChi, double cosChi, double sinEta, double cosEta, const Sph& s,
                  double out[3]) {
    double sinChiSinEta = sinChi * sinEta;
    double aux = sinChi * cosEta * s.cosT + cosChi * s.sinT;
    out[0] = -sinChiSinEta * s.sinP + aux * s.cosP;
    out[1] = sinChiSinEta * s.cosP + aux * s.sinP;
    out[2] = -sinChi * cosEta * s.sinT + cosChi * s.cosT;
}

inline double sq_norm(const double v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
inline double energy_of(const double v[3]) { return 0.5 * kMass * sq_norm(v) / kCharge; }

// ------------------------------------------------------------------ model
struct Electron {
    double r[3], v[3], eps, t, tfree, nue;
    uint64_t stream;
    uint32_t event;
};

}  // namespace

struct orc_model {
    // processes (BoltzmannMC.C:29-271)
    int P = 0, G = 0, I = 10000;
    std::vector<int> type, isElastic, isSuper, gasIdx, scat;
    std::vector<double> energyLoss, threshold, tMass, rMass, relDens, statW, wParam, minE, maxE;
    std::vector<double> params, thermalStd;
    std::vector<std::vector<double>> E, S;
    std::vector<int> firstP, lastP;
    std::vector<double> gasFrac;
    double energyMaxElastic = 1e100;
    // condition
    ekmc_condition cond;
    ekmc_numerics num;
    double N = 0, Tg = 0, gasEnergy = 0;
    orc_field f;
    double elecFieldAngle = 180, reducedElecField = 0;
    bool cylSym = true;
    double energySharingFactor = 0.5;
    // tables (BoltzmannMC.C:561-615)
    double maxInterp = -1, step = 0;
    std::vector<double> interp, cumul, nutot, numax;

    // ---- solver state (orc_solve)
    int64_t nE = 0;
    std::vector<Electron> el;
    std::vector<int> pid;
    std::vector<Electron> ej;
    std::vector<double> dE, gainF;
    double time = 0, ssTime = kNonDef, trial = 0, maxElecEnergy = 0;
    uint64_t popDraws = 0;
    std::vector<double> sT, sMeanE, sMeanR, sMeanV, sFlux, sCov, sBulkV, sBulkD;
    int nSamp = 0, nSync = 0, idx = 0, nIntPts = 0, firstIntIdx = 0;
    double totalIntTime = 0, timeIntervalInteg = 0;
    double energyGainField = 0, energyGrowth = 0;
    std::vector<double> gainP, lossP;
    std::vector<uint64_t> counters;
    uint64_t total = 0, nullc = 0, atSS = 0, nullAtSS = 0, afterSS = 0;
    bool good = false;
    // distributions
    int nCells = 1000;
    double maxEedfE = 0, eedfStep = 0;
    std::vector<double> eedfNodes, eedfCells, eehSum;
    std::vector<double> eehSumPer;  // [phases x cells]
    int nPhases = 0;
    double phaseStep = 0;
    std::vector<double> perN, perE, perV;
    // averages
    double avgE = kNonDef, avgEErr = 0;
    double fluxV[3], fluxVErr[3], fluxD[9], fluxDErr[9], bulkV[3], bulkVErr[3], bulkD[9], bulkDErr[9];
    double powerBalanceRelError = 0;
};

namespace {

// Klein-Rosseland superelastic and forward interpolation of one process at a
// grid energy (BoltzmannMC.C:577-606)
double table_value(const orc_model& m, int p, double energy) {
    if (m.isSuper[p]) {
        if (energy > m.minE[p] && energy <= m.maxE[p]) {
            return m.statW[p] * (1.0 + m.minE[p - 1] / energy) *
                   gsl_linear(m.E[p - 1], m.S[p - 1], energy + m.minE[p - 1]) * m.relDens[p];
        }
        return 0;
    }
    if (energy >= m.minE[p] && energy <= m.maxE[p])
        return gsl_linear(m.E[p], m.S[p], energy) * m.relDens[p];
    return 0;
}

void interpolate_cross_sections(orc_model& m, double maxEnergy) {
    if (maxEnergy == m.maxInterp) return;
    m.maxInterp = maxEnergy;
    m.step = maxEnergy / (double)(m.I - 1);
    double maxF = 0;
    for (int i = 0; i < m.I; ++i) {
        double energy = i * m.step;
        double cur = 0;
        for (int p = 0; p < m.P; ++p) {
            double value = table_value(m, p, energy);
            m.interp[(size_t)i * m.P + p] = value;
            cur += value;
            m.cumul[(size_t)i * m.P + p] = cur;
        }
        cur *= m.N * std::sqrt(energy * 2.0 * kCharge / kMass);
        m.nutot[i] = cur;
        maxF = std::fmax(cur, maxF);
        m.numax[i] = maxF;
    }
}

// accelerateElectron (BoltzmannMC.C:804-905)
void accelerate(const orc_model& m, double tE, double dt, double r[3], double v[3], double& eps,
                double& gain) {
    const orc_field& f = m.f;
    double prev = eps;
    const double e = kCharge, me = kMass;
    if (f.reduced_mag_field == 0) {
        if (f.excitation_frequency == 0) {
            for (int k = 0; k < 3; ++k) {
                r[k] += v[k] * dt + f.accel[k] * (0.5 * dt * dt);
                v[k] += f.accel[k] * dt;
            }
        } else {
            double w = f.omega, phi = w * tE;
            double sinPhi = std::sin(phi), cosPhi = std::cos(phi);
            double wDt = w * dt, phase = wDt + phi;
            double sinPhase = std::sin(phase), cosPhase = std::cos(phase);
            double e_me_w = e / (me * w);
            double aux1 = e_me_w / w * (cosPhase + wDt * sinPhi - cosPhi);
            double aux2 = e_me_w * (sinPhi - sinPhase);
            r[0] += v[0] * dt + f.elec_field[0] * aux1;
            r[1] += v[1] * dt;
            r[2] += v[2] * dt + f.elec_field[2] * aux1;
            v[0] += f.elec_field[0] * aux2;
            v[2] += f.elec_field[2] * aux2;
        }
    } else {
        double vx0 = v[0], vy0 = v[1], vz0 = v[2];
        if (f.excitation_frequency == 0) {
            double W = f.cyclotron, WDt = W * dt;
            double sW = std::sin(WDt), cW = std::cos(WDt);
            double sW_W = sW / W, aux2 = (cW - 1.0) / W;
            double vEx = e * f.elec_field[0] / (me * W);
            double az = e * f.elec_field[2] / me;
            r[0] += vx0 * sW_W + (vy0 + vEx) * aux2;
            r[1] += -vx0 * aux2 + vy0 * sW_W + vEx * (sW_W - dt);
            r[2] += vz0 * dt - 0.5 * az * dt * dt;
            v[0] = vx0 * cW - vy0 * sW - vEx * sW;
            v[1] = vx0 * sW + vy0 * cW + vEx * (cW - 1.0);
            v[2] = vz0 - az * dt;
        } else if (std::abs(f.omega - f.cyclotron) / f.omega < 1E-6) {
            double W = f.cyclotron, phi = W * tE;
            double sinPhi = std::sin(phi), cosPhi = std::cos(phi);
            double WDt = W * dt;
            double sW = std::sin(WDt), cW = std::cos(WDt);
            double phase = WDt + phi;
            double sinPhase = std::sin(phase), cosPhase = std::cos(phase);
            double cosOpp = std::cos(phi - WDt);
            double e_me_W = e / (me * W);
            double vEx = e_me_W * f.elec_field[0], vEz = e_me_W * f.elec_field[2];
            double sW_W = sW / W, cWm1_W = (cW - 1.0) / W, cosPhase_W = cosPhase / W;
            r[0] += vx0 * sW_W + vy0 * cWm1_W -
                    0.25 * vEx * (cosPhase_W - cosOpp / W + 2.0 * dt * sinPhase);
            r[1] += -vx0 * cWm1_W + vy0 * sW_W +
                    0.5 * vEx * (dt * cosPhase_W - 2.0 * cWm1_W * sinPhi - cosPhi * sW_W);
            r[2] += vz0 * dt + vEz * (cosPhase_W - cosPhi / W + dt * sinPhi);
            v[0] = vx0 * cW - vy0 * sW - 0.5 * vEx * (WDt * cosPhase + cosPhi * sW);
            v[1] = vx0 * sW + vy0 * cW -
                   0.5 * vEx * (WDt * cW * sinPhi + (WDt * cosPhi - sinPhi) * sW);
            v[2] = vz0 + vEz * (sinPhi - sinPhase);
        } else {
            double w = f.omega, W = f.cyclotron, phi = w * tE;
            double sinPhi = std::sin(phi), cosPhi = std::cos(phi);
            double wDt = w * dt;
            double swD = std::sin(wDt), cwD = std::cos(wDt);
            double phase = wDt + phi;
            double sinPhase = std::sin(phase), cosPhase = std::cos(phase);
            double WDt = W * dt;
            double sW = std::sin(WDt), cW = std::cos(WDt);
            double w2 = w * w, W2 = W * W, w2mW2 = w2 - W2;
            double vEx = e / (me * W) * f.elec_field[0];
            double vEz = e / (me * w) * f.elec_field[2];
            double vEx_d = vEx / w2mW2, WvEx_d = W * vEx_d;
            double sW_W = sW / W, cWm1_W = (cW - 1.0) / W, wSinPhi = w * sinPhi;
            r[0] += vx0 * sW_W + vy0 * cWm1_W +
                    WvEx_d * (cosPhi * (cwD - cW) + (wSinPhi * sW_W - sinPhi * swD));
            r[1] += -vx0 * cWm1_W + vy0 * sW_W -
                    vEx_d / w * ((W2 + w2 * cW - w2) * sinPhi + W2 * (w * cosPhi * sW_W - sinPhase));
            r[2] += vz0 * dt + vEz * ((cosPhase - cosPhi) / w + dt * sinPhi);
            v[0] = vx0 * cW - vy0 * sW + WvEx_d * w * (cW * sinPhi - sinPhase + cosPhi / w * W * sW);
            v[1] = vx0 * sW + vy0 * cW + vEx_d * W2 * (cosPhase - cosPhi * cW + wSinPhi * sW_W);
            v[2] = vz0 + vEz * (sinPhi - sinPhase);
        }
    }
    eps = energy_of(v);
    gain = eps - prev;
}

// AngularScatteringFunctions.h:27-74; slot is the variate slot of the draw
double sample_cos_chi(const orc_model& m, int p, double energy, double energyAfter,
                      const Variates& var, int slot) {
    switch (m.scat[p]) {
        case EKMC_SCAT_FORWARD: return 1;
        case EKMC_SCAT_BORN_DIPOLE: {
            double loss = m.energyLoss[p];
            double ratio = loss / std::pow(std::sqrt(energyAfter) + std::sqrt(energy), 2);
            double r2 = ratio * ratio;
            return 1.0 + 2.0 * r2 / (1.0 - r2) * (1.0 - std::pow(r2, -var.u(slot)));
        }
        case EKMC_SCAT_SURENDRA:
            return (2.0 + energy - 2.0 * std::pow(1.0 + energy, var.u(slot))) / energy;
        case EKMC_SCAT_COULOMB_SCREEN: {
            double eps = m.params[2 * p] == 0 ? energy : energyAfter;
            double sp = m.params[2 * p + 1] / eps;
            double R = var.u(slot);
            return (sp + 1.0 - (2.0 * sp + 1.0) * R) / (sp + 1.0 - R);
        }
        default: return 1.0 - 2.0 * var.u(slot);
    }
}

struct CollisionOut {
    int pid = EKMC_NULL_COLLISION;
    double dE = 0, dEoverE = 0;
    double ejV[3] = {0, 0, 0}, ejEps = 0;
    bool zeroVelocity = false;
};

// conservativeCollision (BoltzmannMC.C:1115-1193)
void conservative(const orc_model& m, int p, double v[3], const double V[3], double& eps,
                  bool thermal, const Variates& var, CollisionOut& o) {
    double M = m.tMass[p], mu = m.rMass[p], loss = m.energyLoss[p], inc = eps;
    if (thermal) {
        double rel[3] = {v[0] - V[0], v[1] - V[1], v[2] - V[2]};
        Sph s;
        if (!cart2sph(rel, s)) { o.zeroVelocity = true; return; }
        double relE = 0.5 * mu * s.norm * s.norm / kCharge;
        double relAfter = relE - loss;
        if (relAfter <= 0) { o.pid = EKMC_NULL_COLLISION; return; }
        double cc = sample_cos_chi(m, p, relE, relAfter, var, 6);
        double sc = std::sqrt(1.0 - cc * cc);
        double eta = 2.0 * kPi * var.u(7);
        double u[3];
        euler(sc, cc, std::sin(eta), std::cos(eta), s, u);
        double sp = std::sqrt(s.norm * s.norm - 2.0 / mu * loss * kCharge);
        for (int k = 0; k < 3; ++k)
            v[k] = M / (kMass + M) * (sp * u[k]) + (kMass * v[k] + M * V[k]) / (kMass + M);
    } else {
        Sph s;
        if (!cart2sph(v, s)) { o.zeroVelocity = true; return; }
        double after = inc - loss;
        if (after <= 0) { o.pid = EKMC_NULL_COLLISION; return; }
        double cc = sample_cos_chi(m, p, inc, after, var, 6);
        double sc = std::sqrt(1.0 - cc * cc);
        double eta = 2.0 * kPi * var.u(7);
        double u[3];
        euler(sc, cc, std::sin(eta), std::cos(eta), s, u);
        double sp = std::sqrt((s.norm * s.norm - 2.0 / kMass * loss * kCharge) *
                              (1.0 - 2.0 * mu / (kMass + M) * (1.0 - cc)));
        for (int k = 0; k < 3; ++k) v[k] = sp * u[k];
    }
    eps = energy_of(v);
    o.dE = eps - inc;
    o.dEoverE = o.dE / inc;
}

// ionizationCollision (BoltzmannMC.C:1195-1272)
void ionization(const orc_model& m, int p, double v[3], double& eps, const Variates& var,
                CollisionOut& o) {
    double inc = eps, thr = m.energyLoss[p];
    if (inc < thr) { o.pid = EKMC_NULL_COLLISION; return; }
    Sph s;
    if (!cart2sph(v, s)) { o.zeroVelocity = true; return; }
    double net = inc - thr, ejE;
    int op = m.num.ionization_operator;
    if (op == EKMC_IONIZ_USING_SDCS)
        ejE = m.wParam[p] * std::tan(var.u(8) * std::atan(net / (2.0 * m.wParam[p])));
    else if (op == EKMC_IONIZ_RANDOM_UNIFORM)
        ejE = var.u(8) * net;
    else
        ejE = m.energySharingFactor * net;
    eps = net - ejE;
    double sCs, cCs, sCe, cCe, sEs, cEs, sEe, cEe;
    if (m.scat[p] == EKMC_SCAT_MOMENTUM_CONSERVATION_IONIZATION) {
        cCs = std::sqrt(eps / net);
        sCs = std::sqrt(1.0 - cCs * cCs);
        double eta = 2.0 * kPi * var.u(7);
        sEs = std::sin(eta); cEs = std::cos(eta);
        cCe = std::sqrt(ejE / net);
        sCe = std::sqrt(1.0 - cCe * cCe);
        sEe = -sEs; cEe = -cEs;
    } else {
        cCs = sample_cos_chi(m, p, inc, eps, var, 6);
        sCs = std::sqrt(1.0 - cCs * cCs);
        double eta = 2.0 * kPi * var.u(7);
        sEs = std::sin(eta); cEs = std::cos(eta);
        cCe = sample_cos_chi(m, p, inc, ejE, var, 9);
        sCe = std::sqrt(1.0 - cCe * cCe);
        double etaE = 2.0 * kPi * var.u(10);
        sEe = std::sin(etaE); cEe = std::cos(etaE);
    }
    double u[3];
    euler(sCs, cCs, sEs, cEs, s, u);
    double sp = std::sqrt(2.0 * eps * kCharge / kMass);
    for (int k = 0; k < 3; ++k) v[k] = sp * u[k];
    euler(sCe, cCe, sEe, cEe, s, u);
    double spE = std::sqrt(2.0 * ejE * kCharge / kMass);
    for (int k = 0; k < 3; ++k) o.ejV[k] = spE * u[k];
    o.ejEps = ejE;
    o.dE = -thr;
    o.dEoverE = -thr / inc;
}

// performCollision (BoltzmannMC.C:907-1113)
CollisionOut perform_collision(const orc_model& m, double v[3], double& eps, double nue,
                               const Variates& var) {
    CollisionOut o;
    double V[3] = {0, 0, 0};
    int chosen = EKMC_NULL_COLLISION;
    int gte = m.num.gas_temperature_effect;
    bool thermal = gte == EKMC_GASTEMP_TRUE || (gte == EKMC_GASTEMP_SMART && eps < 20.0 * m.gasEnergy);
    const int I = m.I, P = m.P;
    if (thermal) {
        // MathFunctions::unitNormalRand3 (MathFunctions.C:54-59)
        double r1 = var.u(1), r2 = var.u(2), r3 = var.u(3), r4 = var.u(4);
        double aux1 = std::sqrt(-2.0 * std::log(r1)), aux2 = 2.0 * kPi * r2;
        double vers[3] = {aux1 * std::cos(aux2), aux1 * std::sin(aux2),
                          std::sqrt(-2.0 * std::log(r3)) * std::cos(2.0 * kPi * r4)};
        double randCS = nue * var.u(5) / m.N;
        double prevSum = 0;
        for (int g = 0; g < m.G && chosen == EKMC_NULL_COLLISION; ++g) {
            if (m.gasFrac[g] == 0) continue;
            int L = m.firstP[g], R = m.lastP[g];
            for (int k = 0; k < 3; ++k) V[k] = vers[k] * m.thermalStd[L];
            double d[3] = {v[0] - V[0], v[1] - V[1], v[2] - V[2]};
            double relSpeed = std::sqrt(sq_norm(d));
            double over = 0.5 * m.rMass[L] * relSpeed * relSpeed / kCharge / m.step;
            int k1 = (int)std::fmin(over, I - 1);
            int k2 = (int)std::fmin(k1 + 1, I - 1);
            double w1 = (double)k2 - over;
            w1 = w1 < 0 ? 0.0 : 1.0;
            double w2 = 1.0 - w1;
            const double* c1 = &m.cumul[(size_t)k1 * P];
            const double* c2 = &m.cumul[(size_t)k2 * P];
            const double* i1 = &m.interp[(size_t)k1 * P];
            const double* i2 = &m.interp[(size_t)k2 * P];
            double ref = 0;
            if (L > 0) ref = w1 * c1[L - 1] + w2 * c2[L - 1];
            double lim = prevSum + (w1 * c1[R] + w2 * c2[R] - ref) * relSpeed;
            if (randCS > lim) { prevSum = lim; continue; }
            while (L != R) {
                int t = (L + R) / 2;
                double tv = prevSum + (w1 * c1[t] + w2 * c2[t] - ref) * relSpeed;
                if (randCS < tv) R = t;
                else if (randCS > tv) L = t + 1;
                else { chosen = t; break; }
            }
            if (L == R) chosen = L;
            while (w1 * i1[chosen] == 0 && w2 * i2[chosen] == 0) --chosen;
        }
        if (chosen == EKMC_NULL_COLLISION) { o.pid = chosen; return o; }
    } else {
        double randF = nue * var.u(5);
        double over = eps / m.step;
        int k1 = (int)std::fmin(over, I - 1);
        int k2 = (int)std::fmin(k1 + 1, I - 1);
        double w1 = (double)k2 - over;
        if (w1 < 0) w1 = 0.0;
        double w2 = 1.0 - w1;
        if (randF > w1 * m.nutot[k1] + w2 * m.nutot[k2]) { o.pid = EKMC_NULL_COLLISION; return o; }
        const double* c1 = &m.cumul[(size_t)k1 * P];
        const double* c2 = &m.cumul[(size_t)k2 * P];
        const double* i1 = &m.interp[(size_t)k1 * P];
        const double* i2 = &m.interp[(size_t)k2 * P];
        double randCS = randF / m.N / std::sqrt(sq_norm(v));
        int L = 0, R = P - 1;
        while (L != R) {
            int t = (L + R) / 2;
            double tv = w1 * c1[t] + w2 * c2[t];
            if (randCS < tv) R = t;
            else if (randCS > tv) L = t + 1;
            else { chosen = t; break; }
        }
        if (L == R) chosen = L;
        while (w1 * i1[chosen] == 0 && w2 * i2[chosen] == 0) --chosen;
    }
    o.pid = chosen;
    if (m.type[chosen] == EKMC_CONSERVATIVE) conservative(m, chosen, v, V, eps, thermal, var, o);
    else if (m.type[chosen] == EKMC_IONIZATION) ionization(m, chosen, v, eps, var, o);
    else { o.dE = -eps; o.dEoverE = -1; }  // attachmentCollision (BoltzmannMC.C:1274-1280)
    return o;
}

// One pass step of electronDynamicsUntilSynchronization for one electron
// (BoltzmannMC.C:637-680).  Returns NON_DEF (already at t_sync), PARTIAL, or
// the collision id; fills `co` for collisions.
int pass_step(const orc_model& m, uint64_t seed, double trial, double tsync, Electron& e,
              CollisionOut& co, double& gain) {
    if (e.t == tsync) return EKMC_NON_DEF;
    if (e.tfree == kNonDef) {
        e.tfree = -std::log(slot_variate(seed, e.stream, e.event, 0)) / trial;
        e.nue = trial;
    }
    if (e.t + e.tfree > tsync) {
        double dt = tsync - e.t;
        accelerate(m, e.t, dt, e.r, e.v, e.eps, gain);
        e.t = tsync;
        e.tfree -= dt;
        return EKMC_PARTIAL_FLIGHT;
    }
    accelerate(m, e.t, e.tfree, e.r, e.v, e.eps, gain);
    e.t += e.tfree;
    Variates var;
    var.seed = seed; var.stream = e.stream; var.event = e.event;
    co = perform_collision(m, e.v, e.eps, e.nue, var);
    e.event += 1;
    e.tfree = -std::log(slot_variate(seed, e.stream, e.event, 0)) / trial;
    e.nue = trial;
    return co.pid;
}

double max_accel_energy(const orc_model& m, double initialEnergy, double deltaT) {
    // maximizationAccelerationEnergy (BoltzmannMC.C:765-802)
    double e = kCharge, me = kMass, e_me = e / me;
    double Ex0 = std::abs(m.f.elec_field[0]), Ez0 = std::abs(m.f.elec_field[2]);
    double Ex02 = Ex0 * Ex0, Ez02 = Ez0 * Ez0, E02 = Ex02 + Ez02, E0 = std::sqrt(E02);
    double v0 = std::sqrt(initialEnergy * e * 2.0 / me);
    double w = m.f.omega, W = m.f.cyclotron;
    double g = (E0 * v0 + 0.5 * e_me * E02 * deltaT) * deltaT;
    if (W == 0) {
        if (w != 0) g = std::fmin(g, 2.0 / w * (e_me * E02 / w + v0 * (Ex0 + Ez0)));
    } else {
        if (w == 0) {
            g = std::fmin(g, 0.5 * e_me * Ez02 * deltaT * deltaT + (2.0 * e_me * Ex02 / W + 3.0 * v0 * Ex0) / W +
                                 v0 * deltaT * Ez0);
        } else if (std::abs(w - W) / w < 1E-6) {
            double W2 = W * W;
            g = std::fmin(g, 2.0 * e_me * Ez02 / W2 + e_me * Ex02 / (8.0 * W2) * (4.0 + W * deltaT * (2.0 + W * deltaT)) +
                                 (v0 * deltaT + v0 / W) * Ex0 + 2.0 * v0 * Ez0 / W);
        } else {
            double w2 = w * w, W2 = W * W, d = w2 - W2;
            g = std::fmin(g, 2.0 * e_me * Ez02 / w2 + 0.5 * e_me * Ex02 / (d * d) * (5.0 * w2 + 8.0 * w * W + 5.0 * W * W) +
                                 3.0 * v0 * Ex0 / std::abs(w - W) + 2.0 * v0 * Ez0 / w);
        }
    }
    return initialEnergy + g;
}

// checkMaxCollisionFrequency (BoltzmannMC.C:716-763)
void check_max_collision_frequency(orc_model& m) {
    double maxBefore = 0;
    for (auto& e : m.el) maxBefore = std::max(maxBefore, e.eps);
    double thermalC = 0;
    int gte = m.num.gas_temperature_effect;
    if (gte == EKMC_GASTEMP_TRUE || gte == EKMC_GASTEMP_SMART) thermalC = 10.0 * m.gasEnergy;
    bool updated = true;
    while (updated) {
        updated = false;
        double maxEnergy = max_accel_energy(m, maxBefore, 10.0 / m.trial) + thermalC;
        if (maxEnergy > m.maxInterp || 2.5 * maxEnergy < m.maxInterp) {
            if (2.0 * maxEnergy < m.energyMaxElastic) interpolate_cross_sections(m, 2.0 * maxEnergy);
            else interpolate_cross_sections(m, m.energyMaxElastic);
        }
        int k = (int)std::fmin(std::ceil(maxEnergy / m.step), m.I - 1);
        if (m.trial < m.numax[k]) { updated = true; m.trial *= 1.1; }
    }
}

double pop_uniform(orc_model& m) {
    // sequential U[0,1) stream for the serial population control
    uint32_t w[4];
    philox_words(m.num.seed, m.popDraws >> 1, 0, P_POPCTL, 0, w);
    int h = (int)(m.popDraws & 1);
    ++m.popDraws;
    uint64_t x = ((((uint64_t)w[2 * h]) << 32) | w[2 * h + 1]) >> 11;
    return (double)x / 9007199254740992.0;
}

// nonParallelCollisionTasks (BoltzmannMC.C:1282-1408)
void non_parallel_tasks(orc_model& m) {
    const int64_t N = m.nE;
    std::vector<char> isAtt(N, 0);
    std::vector<int64_t> ejIds, attIds;
    for (int64_t i = 0; i < N; ++i) {
        int pid = m.pid[i];
        if (pid == EKMC_NON_DEF) continue;
        m.energyGainField += m.gainF[i];
        if (pid == EKMC_NULL_COLLISION) { ++m.nullc; continue; }
        if (pid == EKMC_PARTIAL_FLIGHT) continue;
        ++m.counters[pid];
        ++m.total;
        if (m.dE[i] >= 0) m.gainP[pid] += m.dE[i];
        else m.lossP[pid] += m.dE[i];
        if (m.type[pid] == EKMC_IONIZATION) ejIds.push_back(i);
        else if (m.type[pid] == EKMC_ATTACHMENT) { isAtt[i] = 1; attIds.push_back(i); }
    }
    int64_t nEj = (int64_t)ejIds.size();
    for (int64_t a : attIds) {
        if (nEj > 0) {
            int64_t j = ejIds.back();
            Electron& d = m.el[a];
            const Electron& s = m.ej[j];
            std::memcpy(d.r, s.r, sizeof d.r);
            std::memcpy(d.v, s.v, sizeof d.v);
            d.eps = s.eps;
            d.t = m.el[j].t;
            d.tfree = kNonDef;
            d.nue = m.trial;
            d.stream = s.stream;
            d.event = 0;
            isAtt[a] = 0;
            ejIds.pop_back();
            --nEj;
        } else {
            int64_t r = (int64_t)std::fmin(pop_uniform(m) * N, N - 1);
            while (isAtt[r]) r = (int64_t)std::fmin(pop_uniform(m) * N, N - 1);
            m.energyGrowth += m.el[r].eps;
            Electron c = m.el[r];
            c.stream = child_stream(m.num.seed ^ (P_CLONE * 0x9E3779B97F4A7C15ull), c.stream, (uint32_t)a);
            c.event = 0;
            m.el[a] = c;
        }
    }
    while (nEj > 0) {
        int64_t r = (int64_t)std::fmin(pop_uniform(m) * (N + nEj), N + nEj - 1);
        if (r < N) {
            int64_t j = ejIds.back();
            m.energyGrowth -= m.el[r].eps;
            Electron& d = m.el[r];
            const Electron& s = m.ej[j];
            std::memcpy(d.r, s.r, sizeof d.r);
            std::memcpy(d.v, s.v, sizeof d.v);
            d.eps = s.eps;
            d.t = m.el[j].t;
            d.tfree = kNonDef;
            d.nue = m.trial;
            d.stream = s.stream;
            d.event = 0;
        } else {
            r -= N;
            m.energyGrowth -= m.ej[ejIds[r]].eps;
            std::swap(ejIds[r], ejIds[nEj - 1]);
        }
        ejIds.pop_back();
        --nEj;
    }
}

void electron_dynamics_until_sync(orc_model& m) {
    check_max_collision_frequency(m);
    double dts = m.num.sync_time_x_max_coll_freq / m.trial;
    double tsync = m.time + dts;
    ++m.nSync;
    const uint64_t seed = m.num.seed;
    bool any = true;
    while (any) {
        check_max_collision_frequency(m);
        const double trial = m.trial;
        int anyFlag = 0;
#pragma omp parallel for schedule(static) reduction(| : anyFlag)
        for (int64_t i = 0; i < m.nE; ++i) {
            Electron& e = m.el[i];
            CollisionOut co;
            double gain = 0;
            uint64_t parentStream = e.stream;
            uint32_t parentEvent = e.event;
            int r = pass_step(m, seed, trial, tsync, e, co, gain);
            m.pid[i] = r;
            if (r == EKMC_NON_DEF) continue;
            m.gainF[i] = gain;
            if (r == EKMC_PARTIAL_FLIGHT) continue;
            anyFlag = 1;
            if (r >= 0) {
                m.dE[i] = co.dE;
                if (m.type[r] == EKMC_IONIZATION) {
                    Electron& j = m.ej[i];
                    std::memcpy(j.r, e.r, sizeof j.r);
                    std::memcpy(j.v, co.ejV, sizeof j.v);
                    j.eps = co.ejEps;
                    j.stream = child_stream(seed, parentStream, parentEvent);
                }
            }
        }
        non_parallel_tasks(m);
        any = anyFlag != 0;
    }
    m.time = tsync;
}

void histogram_count(const double* x, int64_t n, const std::vector<double>& nodes, double* counts) {
    // MathFunctions::histogramCount (MathFunctions.C:61-81)
    int nc = (int)nodes.size() - 1;
    double first = nodes[0], step = nodes[1] - first;
    for (int64_t i = 0; i < n; ++i) {
        int c = (int)((x[i] - first) / step);
        if (c < nc) counts[c] += 1;
    }
}

std::vector<double> linspaced(int n, double a, double b) {
    // Eigen::ArrayXd::LinSpaced(n, a, b): a + i*(b-a)/(n-1), last value exactly b
    std::vector<double> out(n);
    if (n == 1) { out[0] = b; return out; }
    double step = (b - a) / (n - 1);
    for (int i = 0; i < n; ++i) out[i] = (i == n - 1) ? b : a + i * step;
    return out;
}

void calculate_mean(orc_model& m) {
    // calculateMeanDataForSwarmParams (BoltzmannMC.C:1410-1482)
    int i = m.idx;
    auto grow = [&](std::vector<double>& v, int w) { if ((int)v.size() < (i + 1) * w) v.resize((size_t)(2 * (i + 1)) * w, 0.0); };
    grow(m.sT, 1); grow(m.sMeanE, 1); grow(m.sMeanR, 3); grow(m.sMeanV, 3); grow(m.sFlux, 9);
    grow(m.sCov, 9); grow(m.sBulkV, 3); grow(m.sBulkD, 9);
    double mx = 0;
    for (auto& e : m.el) mx = std::max(mx, e.eps);
    m.maxElecEnergy = std::fmax(m.maxElecEnergy, mx);
    const double N = (double)m.nE;
    double sE = 0, sR[3] = {0, 0, 0}, sV[3] = {0, 0, 0}, cov[9] = {0}, fd[9] = {0};
    for (auto& e : m.el) {
        sE += e.eps;
        for (int a = 0; a < 3; ++a) {
            sR[a] += e.r[a]; sV[a] += e.v[a];
            for (int b = 0; b < 3; ++b) { cov[3 * a + b] += e.r[a] * e.r[b]; fd[3 * a + b] += e.r[a] * e.v[b]; }
        }
    }
    m.sMeanE[i] = sE / N;
    for (int a = 0; a < 3; ++a) { m.sMeanR[3 * i + a] = sR[a] / N; m.sMeanV[3 * i + a] = sV[a] / N; }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            m.sCov[9 * i + 3 * a + b] = cov[3 * a + b] / N - m.sMeanR[3 * i + a] * m.sMeanR[3 * i + b];
            m.sFlux[9 * i + 3 * a + b] = fd[3 * a + b] / N - m.sMeanR[3 * i + a] * m.sMeanV[3 * i + b];
        }
    if (i != 0) {
        double dt = m.sT[i] - m.sT[i - 1];
        for (int a = 0; a < 3; ++a) m.sBulkV[3 * i + a] = (m.sMeanR[3 * i + a] - m.sMeanR[3 * (i - 1) + a]) / dt;
        for (int a = 0; a < 9; ++a) m.sBulkD[9 * i + a] = 0.5 * (m.sCov[9 * i + a] - m.sCov[9 * (i - 1) + a]) / dt;
    } else {
        for (int a = 0; a < 3; ++a) m.sBulkV[a] = m.sMeanV[a];
        for (int a = 0; a < 9; ++a) m.sBulkD[a] = m.sFlux[a];
    }
    if (m.f.omega != 0 && m.ssTime != kNonDef) {
        double phase = std::fmod(m.f.omega * m.time, 2.0 * kPi);
        int ph = (int)std::fmin(std::floor(phase / m.phaseStep), m.nPhases - 1);
        m.perN[ph] += 1.0;
        m.perE[ph] += m.sMeanE[i];
        for (int a = 0; a < 3; ++a) m.perV[3 * ph + a] += m.sMeanV[3 * i + a];
    }
}

void time_dependent_distributions(orc_model& m) {
    // getTimeDependDistributions (BoltzmannMC.C:1492-1572); the periodic
    // re-binning is written as a plain linear split.
    if (m.maxElecEnergy > m.maxEedfE) {
        double oldStep = m.eedfStep;
        std::vector<double> oldNodes = m.eedfNodes, oldSum = m.eehSum, oldPer = m.eehSumPer;
        m.maxEedfE = 1.2 * m.maxElecEnergy;
        m.eedfNodes = linspaced(m.nCells + 1, 0.0, m.maxEedfE);
        m.eedfStep = m.eedfNodes[1];
        m.eedfCells = linspaced(m.nCells, m.eedfStep / 2.0, m.maxEedfE - m.eedfStep / 2.0);
        std::fill(m.eehSum.begin(), m.eehSum.end(), 0.0);
        std::fill(m.eehSumPer.begin(), m.eehSumPer.end(), 0.0);
        for (int o = 0; o < m.nCells; ++o) {
            double l = oldNodes[o], r = oldNodes[o + 1];
            int nl = (int)(l / m.eedfStep), nr = (int)(r / m.eedfStep);
            if (nl == nr) {
                m.eehSum[nl] += oldSum[o];
                for (int ph = 0; ph < m.nPhases; ++ph) m.eehSumPer[(size_t)ph * m.nCells + nl] += oldPer[(size_t)ph * m.nCells + o];
            } else {
                double lf = (m.eedfNodes[nr] - l) / oldStep;
                m.eehSum[nl] += lf * oldSum[o];
                m.eehSum[nr] += (1.0 - lf) * oldSum[o];
                for (int ph = 0; ph < m.nPhases; ++ph) {
                    m.eehSumPer[(size_t)ph * m.nCells + nl] += lf * oldPer[(size_t)ph * m.nCells + o];
                    m.eehSumPer[(size_t)ph * m.nCells + nr] += (1.0 - lf) * oldPer[(size_t)ph * m.nCells + o];
                }
            }
        }
    }
    std::vector<double> eps(m.nE);
    for (int64_t i = 0; i < m.nE; ++i) eps[i] = m.el[i].eps;
    std::vector<double> eeh(m.nCells, 0.0);
    histogram_count(eps.data(), m.nE, m.eedfNodes, eeh.data());
    for (int c = 0; c < m.nCells; ++c) m.eehSum[c] += eeh[c];
    if (m.f.omega != 0) {
        double phase = std::fmod(m.f.omega * m.time, 2.0 * kPi);
        int ph = (int)std::fmin(std::floor(phase / m.phaseStep), m.nPhases - 1);
        for (int c = 0; c < m.nCells; ++c) m.eehSumPer[(size_t)ph * m.nCells + c] += eeh[c];
    }
}

double mean_of(const std::vector<double>& a, int w, int col, int first, int n) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[(size_t)(first + i) * w + col];
    return s / n;
}

double stat_error_strided(const std::vector<double>& a, int w, int col, int first, int n, double nBins) {
    // MathFunctions::statisticalError (MathFunctions.C:211-231)
    if (n == 0) return 0;
    int nb = (int)nBins;
    int np = (int)(n / nBins);
    double mean = mean_of(a, w, col, first, n);
    double sum = 0;
    for (int b = 0; b < nb; ++b) sum += std::pow(mean_of(a, w, col, first + b * np, np) - mean, 2);
    return std::sqrt(sum) / nBins;
}

void time_averages(orc_model& m) {
    int f = m.firstIntIdx, n = m.nIntPts;
    m.avgE = mean_of(m.sMeanE, 1, 0, f, n);
    m.avgEErr = stat_error_strided(m.sMeanE, 1, 0, f, n, 50);
    for (int a = 0; a < 3; ++a) {
        m.fluxV[a] = mean_of(m.sMeanV, 3, a, f, n);
        m.fluxVErr[a] = stat_error_strided(m.sMeanV, 3, a, f, n, 50);
        m.bulkV[a] = mean_of(m.sBulkV, 3, a, f, n);
        m.bulkVErr[a] = stat_error_strided(m.sBulkV, 3, a, f, n, 50);
    }
    for (int a = 0; a < 9; ++a) {
        m.fluxD[a] = mean_of(m.sFlux, 9, a, f, n);
        m.fluxDErr[a] = stat_error_strided(m.sFlux, 9, a, f, n, 50);
        m.bulkD[a] = mean_of(m.sBulkD, 9, a, f, n);
        m.bulkDErr[a] = stat_error_strided(m.sBulkD, 9, a, f, n, 50);
    }
}

void check_power_balance(orc_model& m) {
    double g = m.energyGainField, l = 0;
    if (m.energyGrowth > 0) g += m.energyGrowth; else l += m.energyGrowth;
    for (int p = 0; p < m.P; ++p) { g += m.gainP[p]; l += m.lossP[p]; }
    m.powerBalanceRelError = std::abs(g + l) / g;
}

void check_statistical_errors(orc_model& m) {
    // BoltzmannMC.C:1743-1767
    time_averages(m);
    check_power_balance(m);
    const ekmc_numerics& q = m.num;
    double nf = m.fluxV[0] * m.fluxV[0] + m.fluxV[1] * m.fluxV[1] + m.fluxV[2] * m.fluxV[2];
    double nb = m.bulkV[0] * m.bulkV[0] + m.bulkV[1] * m.bulkV[1] + m.bulkV[2] * m.bulkV[2];
    double ef = (std::abs(m.fluxV[0]) * m.fluxVErr[0] + std::abs(m.fluxV[1]) * m.fluxVErr[1] + std::abs(m.fluxV[2]) * m.fluxVErr[2]) / nf;
    double eb = (std::abs(m.bulkV[0]) * m.bulkVErr[0] + std::abs(m.bulkV[1]) * m.bulkVErr[1] + std::abs(m.bulkV[2]) * m.bulkVErr[2]) / nb;
    if (m.avgEErr / m.avgE <= q.rel_error_mean_energy && ef <= q.rel_error_flux_drift_velocity &&
        m.fluxDErr[0] / m.fluxD[0] <= q.rel_error_flux_diffusion_coeffs &&
        m.fluxDErr[4] / m.fluxD[4] <= q.rel_error_flux_diffusion_coeffs &&
        m.fluxDErr[8] / m.fluxD[8] <= q.rel_error_flux_diffusion_coeffs &&
        eb <= q.rel_error_bulk_drift_velocity &&
        m.bulkDErr[0] / m.bulkD[0] <= q.rel_error_bulk_diffusion_coeffs &&
        m.bulkDErr[4] / m.bulkD[4] <= q.rel_error_bulk_diffusion_coeffs &&
        m.bulkDErr[8] / m.bulkD[8] <= q.rel_error_bulk_diffusion_coeffs &&
        m.powerBalanceRelError <= q.rel_error_power_balance)
        m.good = true;
}

void check_steady_state(orc_model& m) {
    // BoltzmannMC.C:1787-1893
    double t1 = 0.5 * m.time, t2 = 0.75 * m.time;
    double a1 = 0, a2 = 0, rs = 0;
    int n1 = 0, n2 = 0;
    for (int i = 0; i < m.nSamp; ++i) {
        double st = m.sT[i];
        if (st >= t1 && st <= t2) { a1 += m.sMeanE[i]; ++n1; }
        else if (st > t2) { double v = m.sMeanE[i]; a2 += v; rs += v * v; ++n2; }
    }
    a1 /= n1; a2 /= n2;
    rs = std::sqrt((rs / n2 - a2 * a2) / n2);
    double maxBefore = m.num.max_collisions_before_ss * (double)m.nE;
    if ((a1 >= a2 && rs < 0.01) || (double)m.total >= maxBefore) {
        m.firstIntIdx = m.idx;
        m.nIntPts = 1;
        m.ssTime = m.time;
        m.atSS = m.total;
        m.nullAtSS = m.nullc;
        m.energyGainField = 0;
        std::fill(m.counters.begin(), m.counters.end(), 0);
        std::fill(m.gainP.begin(), m.gainP.end(), 0.0);
        std::fill(m.lossP.begin(), m.lossP.end(), 0.0);
        m.energyGrowth = 0;
        m.totalIntTime = 0;
        m.maxEedfE = 1.2 * m.maxElecEnergy;
        m.eedfNodes = linspaced(m.nCells + 1, 0.0, m.maxEedfE);
        m.eedfStep = m.eedfNodes[1];
        m.eedfCells = linspaced(m.nCells, m.eedfStep / 2.0, m.maxEedfE - m.eedfStep / 2.0);
        m.eehSum.assign(m.nCells, 0.0);
        if (m.f.omega != 0) m.eehSumPer.assign((size_t)m.nPhases * m.nCells, 0.0);
        time_dependent_distributions(m);
    }
}

void evaluate_non_constant(orc_model& m) {
    // BoltzmannMC.C:428-559
    m.time = 0;
    m.ssTime = kNonDef;
    m.el.assign(m.nE, Electron{});
    double dev = std::sqrt(kBoltzmann * m.num.initial_elec_temp_over_gas_temp * m.Tg / kMass);
    for (int64_t i = 0; i < m.nE; ++i) {
        Electron& e = m.el[i];
        uint32_t w[4], w2[4];
        philox_words(m.num.seed, (uint64_t)i, 0, P_INIT, 0, w);
        philox_words(m.num.seed, (uint64_t)i, 0, P_INIT, 1, w2);
        auto u = [](uint32_t a, uint32_t b) {
            return to_unit(((((uint64_t)a) << 32) | b) >> 11, OPEN_CLOSED);
        };
        double r1 = u(w[0], w[1]), r2 = u(w[2], w[3]), r3 = u(w2[0], w2[1]), r4 = u(w2[2], w2[3]);
        double aux1 = std::sqrt(-2.0 * std::log(r1)), aux2 = 2.0 * kPi * r2;
        double n3[3] = {aux1 * std::cos(aux2), aux1 * std::sin(aux2),
                        std::sqrt(-2.0 * std::log(r3)) * std::cos(2.0 * kPi * r4)};
        for (int k = 0; k < 3; ++k) { e.r[k] = 0; e.v[k] = n3[k] * dev; }
        e.eps = 0.5 * kMass * sq_norm(e.v) / kCharge;
        e.t = 0;
        e.tfree = kNonDef;
        e.stream = (uint64_t)i;
        e.event = 0;
    }
    double mx = 0;
    for (auto& e : m.el) mx = std::max(mx, e.eps);
    m.maxInterp = -1;
    interpolate_cross_sections(m, 2.0 * mx);
    m.trial = m.numax[m.I - 1];
    for (auto& e : m.el) e.nue = m.trial;
    m.maxElecEnergy = mx;
    m.avgE = kNonDef;
    m.energyGainField = m.energyGrowth = 0;
    m.total = m.nullc = 0;
    m.counters.assign(m.P, 0);
    m.gainP.assign(m.P, 0.0);
    m.lossP.assign(m.P, 0.0);
    m.good = false;
    m.pid.assign(m.nE, 0);
    m.ej.assign(m.nE, Electron{});
    m.dE.assign(m.nE, 0.0);
    m.gainF.assign(m.nE, 0.0);
    m.popDraws = 0;
    m.sT.assign(100, 0.0); m.sMeanE.assign(100, 0.0); m.sMeanR.assign(300, 0.0);
    m.sMeanV.assign(300, 0.0); m.sFlux.assign(900, 0.0); m.sCov.assign(900, 0.0);
    m.sBulkV.assign(300, 0.0); m.sBulkD.assign(900, 0.0);
    m.nPhases = (m.f.omega != 0) ? m.num.n_integration_phases : 0;
    if (m.nPhases > 0) {
        m.phaseStep = 2.0 * kPi / m.nPhases;
        m.perN.assign(m.nPhases, 0.0); m.perE.assign(m.nPhases, 0.0); m.perV.assign(3 * m.nPhases, 0.0);
    }
    m.nCells = m.num.n_energy_cells;
    m.eehSum.assign(m.nCells, 0.0);
    m.eehSumPer.assign((size_t)m.nPhases * m.nCells, 0.0);
    m.maxEedfE = 0;
    m.totalIntTime = 0;
    m.nIntPts = 0;
    m.firstIntIdx = 0;
}

}  // namespace

// ======================================================================= API

extern "C" {

int orc_model_create(const ekmc_mixture* mix, const ekmc_condition* cond, const ekmc_numerics* num,
                     orc_model** out) {
    if (!mix || !cond || !num || !out || mix->n_processes <= 0) return EKMC_ERR_INVALID_ARGUMENT;
    orc_model* m = new orc_model();
    m->P = mix->n_processes;
    m->G = mix->n_gases;
    m->I = num->n_interp_points;
    m->cond = *cond;
    m->num = *num;
    const int P = m->P;
    m->type.resize(P); m->isElastic.resize(P); m->isSuper.resize(P); m->gasIdx.resize(P); m->scat.resize(P);
    m->energyLoss.resize(P); m->threshold.resize(P); m->tMass.resize(P); m->rMass.resize(P);
    m->relDens.resize(P); m->statW.resize(P); m->wParam.resize(P); m->minE.resize(P); m->maxE.resize(P);
    m->params.resize(2 * P); m->thermalStd.resize(P); m->E.resize(P); m->S.resize(P);
    m->Tg = cond->gas_temperature;
    m->N = cond->gas_density;
    m->gasEnergy = 1.5 * kBoltzmann * m->Tg / kCharge;
    for (int p = 0; p < P; ++p) {
        m->type[p] = mix->type[p];
        m->isElastic[p] = mix->is_elastic[p];
        m->isSuper[p] = mix->is_superelastic[p];
        m->gasIdx[p] = mix->gas_index[p];
        m->scat[p] = mix->scattering[p];
        m->threshold[p] = mix->threshold[p];
        m->tMass[p] = mix->target_mass[p];
        m->rMass[p] = kMass * m->tMass[p] / (kMass + m->tMass[p]);
        m->relDens[p] = mix->rel_density[p];
        m->statW[p] = mix->stat_weight_ratio[p];
        m->wParam[p] = mix->w_parameter[p];
        m->params[2 * p] = mix->scattering_params[2 * p];
        m->params[2 * p + 1] = mix->scattering_params[2 * p + 1];
        m->thermalStd[p] = std::sqrt(kBoltzmann * m->Tg / m->tMass[p]);
        if (m->isSuper[p]) {
            if (p == 0) { delete m; return EKMC_ERR_INVALID_ARGUMENT; }
            m->energyLoss[p] = -m->threshold[p];
            m->minE[p] = 0;
            m->maxE[p] = m->maxE[p - 1] - m->threshold[p];
            continue;
        }
        m->energyLoss[p] = m->threshold[p];
        std::vector<double> e(mix->table_energy + mix->table_offset[p], mix->table_energy + mix->table_offset[p + 1]);
        std::vector<double> s(mix->table_sigma + mix->table_offset[p], mix->table_sigma + mix->table_offset[p + 1]);
        if (e.empty()) { delete m; return EKMC_ERR_INVALID_ARGUMENT; }
        // Collision.C:41-44: Elastic tables start at 0 eV with the first value
        if (m->isElastic[p] && e[0] != 0) { e.insert(e.begin(), 0.0); s.insert(s.begin(), s[0]); }
        // BoltzmannMC.C:139-150: drop points below threshold, add (threshold, 0)
        double thr = m->threshold[p];
        while (!e.empty() && e[0] < thr) { e.erase(e.begin()); s.erase(s.begin()); }
        if (e.empty() || e[0] != thr) { e.insert(e.begin(), thr); s.insert(s.begin(), 0.0); }
        if (e.size() < 2) { delete m; return EKMC_ERR_INVALID_ARGUMENT; }
        m->minE[p] = thr;
        m->maxE[p] = e.back();
        if (m->isElastic[p]) m->energyMaxElastic = std::fmin(m->energyMaxElastic, m->maxE[p]);
        m->E[p] = e;
        m->S[p] = s;
    }
    m->firstP.assign(m->G, 0);
    m->lastP.assign(m->G, -1);
    m->gasFrac.assign(mix->gas_fraction, mix->gas_fraction + m->G);
    for (int g = 0; g < m->G; ++g) {
        int first = -1, last = -1;
        for (int p = 0; p < P; ++p)
            if (m->gasIdx[p] == g) { if (first < 0) first = p; last = p; }
        m->firstP[g] = first;
        m->lastP[g] = last;
    }
    m->interp.assign((size_t)m->I * P, 0.0);
    m->cumul.assign((size_t)m->I * P, 0.0);
    m->nutot.assign(m->I, 0.0);
    m->numax.assign(m->I, 0.0);
    if (num->ionization_operator == EKMC_IONIZ_ONE_TAKES_ALL) m->energySharingFactor = 0;
    // fields (BoltzmannMC.C:453-489)
    double Ev = cond->reduced_elec_field * 1e-21 * m->N;
    double ang = cond->elec_field_angle;
    if (ang == 180) { m->f.elec_field[0] = 0; m->f.elec_field[1] = 0; m->f.elec_field[2] = -Ev; }
    else {
        m->f.elec_field[0] = Ev * std::sin(ang / 180.0 * kPi);
        m->f.elec_field[1] = 0;
        m->f.elec_field[2] = Ev * std::cos(ang / 180.0 * kPi);
    }
    m->f.excitation_frequency = cond->excitation_frequency;
    m->f.omega = cond->excitation_frequency * 2.0 * kPi;
    if (cond->excitation_frequency != 0)
        for (int k = 0; k < 3; ++k) m->f.elec_field[k] *= std::sqrt(2);
    m->f.reduced_mag_field = cond->reduced_mag_field;
    double B = cond->reduced_mag_field * 1e-27 * m->N;
    m->f.cyclotron = kCharge * B / kMass;
    for (int k = 0; k < 3; ++k) m->f.accel[k] = -kCharge / kMass * m->f.elec_field[k];
    m->elecFieldAngle = ang;
    m->reducedElecField = cond->reduced_elec_field;
    m->cylSym = (ang == 180);
    m->nE = cond->n_electrons;
    *out = m;
    return EKMC_OK;
}

void orc_model_destroy(orc_model* m) { delete m; }
int orc_model_nprocesses(const orc_model* m) { return m->P; }
void orc_model_field(const orc_model* m, orc_field* out) { *out = m->f; }

int orc_model_tables(orc_model* m, double max_energy, double* interp, double* cumul,
                     double* total_freq, double* max_freq, double* energy_step) {
    if (max_energy > 0) interpolate_cross_sections(*m, max_energy);
    size_t n = (size_t)m->I * m->P;
    if (interp) std::memcpy(interp, m->interp.data(), n * sizeof(double));
    if (cumul) std::memcpy(cumul, m->cumul.data(), n * sizeof(double));
    if (total_freq) std::memcpy(total_freq, m->nutot.data(), m->I * sizeof(double));
    if (max_freq) std::memcpy(max_freq, m->numax.data(), m->I * sizeof(double));
    if (energy_step) *energy_step = m->step;
    return EKMC_OK;
}

double orc_max_accel_energy(const orc_model* m, double e0, double dt) { return max_accel_energy(*m, e0, dt); }

int orc_accelerate(const orc_model* m, int64_t n, const double* time, const double* dt, double* pos,
                   double* vel, double* energy, double* gain) {
    for (int64_t i = 0; i < n; ++i) {
        double g = 0;
        double eps = energy_of(&vel[3 * i]);
        accelerate(*m, time[i], dt[i], &pos[3 * i], &vel[3 * i], eps, g);
        energy[i] = eps;
        if (gain) gain[i] = g;
    }
    return EKMC_OK;
}

int orc_collide(const orc_model* m, int64_t n, double* vel, double* energy, const double* trial_freq,
                const double* variates, int32_t* process, double* ejected_vel, double* ejected_energy,
                double* energy_change) {
    int status = EKMC_OK;
    for (int64_t i = 0; i < n; ++i) {
        Variates var;
        var.tape = &variates[(size_t)i * EKMC_VARIATES_PER_EVENT];
        CollisionOut o = perform_collision(*m, &vel[3 * i], energy[i], trial_freq[i], var);
        if (o.zeroVelocity) status = EKMC_ERR_ZERO_VELOCITY;
        process[i] = o.pid;
        for (int k = 0; k < 3; ++k) ejected_vel[3 * i + k] = o.ejV[k];
        ejected_energy[i] = o.ejEps;
        energy_change[i] = o.dE;
    }
    return status;
}

int orc_choose_linear(const orc_model* m, double energy, double speed, double randF) {
    const int P = m->P, I = m->I;
    double over = energy / m->step;
    int k1 = (int)std::fmin(over, I - 1), k2 = (int)std::fmin(k1 + 1, I - 1);
    double w1 = (double)k2 - over;
    if (w1 < 0) w1 = 0.0;
    double w2 = 1.0 - w1;
    if (randF > w1 * m->nutot[k1] + w2 * m->nutot[k2]) return EKMC_NULL_COLLISION;
    double target = randF / m->N / speed;
    double sum = 0;
    int last = -1;
    for (int p = 0; p < P; ++p) {
        double val = w1 * m->interp[(size_t)k1 * P + p] + w2 * m->interp[(size_t)k2 * P + p];
        if (val != 0) last = p;
        sum += val;
        if (sum >= target && val != 0) return p;
    }
    return last;
}

void orc_philox(int64_t n, const uint32_t* counter, const uint32_t* key, uint32_t* out) {
    for (int64_t i = 0; i < n; ++i) philox4x32_10(&counter[4 * i], &key[2 * i], &out[4 * i]);
}

double orc_variate(uint64_t seed, uint64_t stream, uint32_t event, int slot) {
    return slot_variate(seed, stream, event, slot);
}

int64_t orc_advance_pool(const orc_model* mc, uint64_t seed, double trial, double tsync, int64_t n,
                         int64_t capacity, double* pos, double* vel, double* energy, double* time,
                         double* free_time, double* nu_e, uint64_t* stream, uint32_t* event,
                         int32_t* alive, uint64_t* counters, uint64_t* n_null) {
    const orc_model& m = *mc;
    std::vector<Electron> pool(n);
    for (int64_t i = 0; i < n; ++i) {
        Electron& e = pool[i];
        for (int k = 0; k < 3; ++k) { e.r[k] = pos[3 * i + k]; e.v[k] = vel[3 * i + k]; }
        e.eps = energy[i]; e.t = time[i]; e.tfree = free_time[i]; e.nue = nu_e[i];
        e.stream = stream[i]; e.event = event[i];
        alive[i] = 1;
    }
    std::vector<int> live(n, 1);
    bool any = true;
    while (any) {
        any = false;
        int64_t cur = (int64_t)pool.size();
        for (int64_t i = 0; i < cur; ++i) {
            if (!live[i]) continue;
            Electron& e = pool[i];
            uint64_t ps = e.stream;
            uint32_t pe = e.event;
            CollisionOut co;
            double gain = 0;
            int r = pass_step(m, seed, trial, tsync, e, co, gain);
            if (r == EKMC_NON_DEF || r == EKMC_PARTIAL_FLIGHT) continue;
            any = true;
            if (r == EKMC_NULL_COLLISION) { if (n_null) ++*n_null; continue; }
            if (counters) ++counters[r];
            if (m.type[r] == EKMC_ATTACHMENT) { live[i] = 0; continue; }
            if (m.type[r] == EKMC_IONIZATION && (int64_t)pool.size() < capacity) {
                Electron c;
                for (int k = 0; k < 3; ++k) { c.r[k] = pool[i].r[k]; c.v[k] = co.ejV[k]; }
                c.eps = co.ejEps; c.t = pool[i].t; c.tfree = kNonDef; c.nue = trial;
                c.stream = child_stream(seed, ps, pe); c.event = 0;
                pool.push_back(c);
                live.push_back(1);
            }
        }
    }
    for (int64_t i = 0; i < (int64_t)pool.size(); ++i) {
        const Electron& e = pool[i];
        for (int k = 0; k < 3; ++k) { pos[3 * i + k] = e.r[k]; vel[3 * i + k] = e.v[k]; }
        energy[i] = e.eps; time[i] = e.t; free_time[i] = e.tfree; nu_e[i] = e.nue;
        stream[i] = e.stream; event[i] = e.event; alive[i] = live[i];
    }
    return (int64_t)pool.size();
}

int orc_solve(orc_model* mp, int n_threads, ekmc_results* out, double* rates, double* eedf_cells,
              double* eedf) {
    orc_model& m = *mp;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    auto start = std::chrono::high_resolution_clock::now();
    evaluate_non_constant(m);
    m.nSamp = 1; m.nSync = 1; m.idx = 0; m.sT[0] = 0;
    calculate_mean(m);
    m.afterSS = 0;
    m.nIntPts = 0;
    const ekmc_numerics& q = m.num;
    const double maxAfter = q.max_collisions_after_ss * (double)m.nE;
    const double minBefore = q.min_collisions_before_ss * (double)m.nE;
    bool stoppedByMax = false;
    while ((!m.good && q.errors_to_be_checked) || m.nIntPts < q.n_integration_points ||
           m.totalIntTime / m.ssTime < q.n_integrated_ss_times || m.totalIntTime < q.integrated_absolute_time) {
        if ((double)m.afterSS >= maxAfter && m.nIntPts > 100) { stoppedByMax = true; break; }
        if (q.max_sync_points > 0 && m.nSync >= q.max_sync_points) break;
        electron_dynamics_until_sync(m);
        if (m.nSync % q.sync_over_sampling == 0) {
            ++m.nSamp;
            m.idx = m.nSamp - 1;
            if ((int)m.sT.size() <= m.idx) m.sT.resize(2 * m.nSamp, 0.0);
            m.sT[m.idx] = m.time;
            calculate_mean(m);
            if (m.maxElecEnergy > m.energyMaxElastic) return EKMC_ERR_ENERGY_OUT_OF_TABLE;
            int dec = (int)std::fmax(std::log(m.nSamp) / std::log(2.0) - 11, 6);
            int nSS = (int)std::pow(2, dec);
            if (m.ssTime != kNonDef) {
                ++m.nIntPts;
                m.afterSS = m.total - m.atSS;
                m.timeIntervalInteg = m.time - m.totalIntTime;
                m.totalIntTime = m.time - m.ssTime;
                time_dependent_distributions(m);
            } else if (m.nSamp >= 100 && m.nSamp % nSS == 0 && (double)m.total > minBefore) {
                check_steady_state(m);
            }
            dec = (int)std::fmax(std::log(m.nIntPts) / std::log(2.0) - 9, 7);
            int nErr = (int)std::pow(2, dec);
            if (m.ssTime != kNonDef && m.nIntPts > 200 && m.nIntPts % nErr == 0) check_statistical_errors(m);
        }
    }
    if (m.nIntPts < 50) return EKMC_ERR_STATISTICS;
    m.totalIntTime = m.time - m.ssTime;
    time_averages(m);
    check_statistical_errors(m);
    auto end = std::chrono::high_resolution_clock::now();
    std::memset(out, 0, sizeof(*out));
    out->mean_energy = m.avgE; out->mean_energy_error = m.avgEErr;
    for (int a = 0; a < 3; ++a) {
        out->flux_drift_velocity[a] = m.fluxV[a]; out->flux_drift_velocity_error[a] = m.fluxVErr[a];
        out->bulk_drift_velocity[a] = m.bulkV[a]; out->bulk_drift_velocity_error[a] = m.bulkVErr[a];
    }
    for (int a = 0; a < 9; ++a) {
        out->flux_diffusion[a] = m.fluxD[a]; out->flux_diffusion_error[a] = m.fluxDErr[a];
        out->bulk_diffusion[a] = m.bulkD[a]; out->bulk_diffusion_error[a] = m.bulkDErr[a];
    }
    double T = m.totalIntTime, NE = (double)m.nE;
    out->power_field = m.energyGainField / (NE * T * m.N);
    out->power_growth = m.energyGrowth / (NE * T * m.N);
    out->power_balance_rel_error = m.powerBalanceRelError;
    double ion = 0, att = 0;
    for (int p = 0; p < m.P; ++p) {
        double k = m.relDens[p] != 0 ? (double)m.counters[p] / (T * NE * m.relDens[p] * m.N) : 0;
        if (rates) rates[p] = k;
        if (m.type[p] == EKMC_IONIZATION) ion += m.relDens[p] * k;
        if (m.type[p] == EKMC_ATTACHMENT) att += m.relDens[p] * k;
    }
    out->total_ion_rate_coeff = ion;
    out->total_att_rate_coeff = att;
    out->steady_state_time = m.ssTime;
    out->final_time = m.time;
    out->total_integrated_time = m.totalIntTime;
    out->trial_collision_frequency = m.trial;
    out->max_elec_energy = m.maxElecEnergy;
    out->max_eedf_energy = m.maxEedfE;
    out->elapsed_time = std::chrono::duration<double>(end - start).count();
    out->real_collisions = m.total;
    out->null_collisions = m.nullc;
    out->real_collisions_at_ss = m.atSS;
    out->null_collisions_at_ss = m.nullAtSS;
    out->n_sampling_points = m.nSamp;
    out->n_integration_points = m.nIntPts;
    out->good_statistical_errors = m.good;
    out->stopped_by_max_collisions = stoppedByMax;
    if (eedf_cells && eedf) {
        double norm = 0;
        for (double c : m.eehSum) norm += c;
        for (int c = 0; c < m.nCells; ++c) {
            eedf_cells[c] = m.eedfCells[c];
            eedf[c] = m.eehSum[c] / (std::sqrt(m.eedfCells[c]) * norm * m.eedfStep);
        }
    }
    return EKMC_OK;
}

int orc_run_intervals(orc_model* mp, int n_threads, int n_sync, uint64_t* events, uint64_t* real,
                      double* seconds) {
    orc_model& m = *mp;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    evaluate_non_constant(m);
    m.nSamp = 1; m.nSync = 1; m.idx = 0; m.sT[0] = 0;
    calculate_mean(m);
    auto start = std::chrono::high_resolution_clock::now();
    for (int k = 0; k < n_sync; ++k) {
        electron_dynamics_until_sync(m);
        ++m.nSamp;
        m.idx = m.nSamp - 1;
        if ((int)m.sT.size() <= m.idx) m.sT.resize(2 * m.nSamp, 0.0);
        m.sT[m.idx] = m.time;
        calculate_mean(m);
    }
    auto end = std::chrono::high_resolution_clock::now();
    *events = m.total + m.nullc;
    *real = m.total;
    *seconds = std::chrono::duration<double>(end - start).count();
    return EKMC_OK;
}

int orc_series(const orc_model* m, int32_t* n_points, double* times, double* mean_energy,
               double* mean_velocity) {
    int n = std::min(*n_points, m->nSamp);
    for (int i = 0; i < n; ++i) {
        if (times) times[i] = m->sT[i];
        if (mean_energy) mean_energy[i] = m->sMeanE[i];
        if (mean_velocity) for (int a = 0; a < 3; ++a) mean_velocity[3 * i + a] = m->sMeanV[3 * i + a];
    }
    *n_points = m->nSamp;
    return EKMC_OK;
}

int orc_population_control(orc_model* mp, int64_t n, const int32_t* pid, const double* dE, double* pos,
                           double* energy, double* time, const double* ej_pos, const double* ej_energy,
                           uint64_t pop_seed, double* growth, uint64_t* counters) {
    orc_model& m = *mp;
    for (int64_t i = 0; i < n; ++i)
        if (pid[i] >= m.P || (pid[i] < 0 && pid[i] != EKMC_NULL_COLLISION && pid[i] != EKMC_PARTIAL_FLIGHT &&
                              pid[i] != EKMC_NON_DEF))
            return EKMC_ERR_INVALID_ARGUMENT;
    const uint64_t seed0 = m.num.seed;
    m.num.seed = pop_seed;
    m.popDraws = 0;
    m.nE = n;
    m.el.assign(n, Electron{});
    m.ej.assign(n, Electron{});
    m.pid.assign(pid, pid + n);
    m.dE.assign(dE, dE + n);
    m.gainF.assign(n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        Electron& e = m.el[i];
        Electron& j = m.ej[i];
        for (int k = 0; k < 3; ++k) { e.r[k] = pos[3 * i + k]; j.r[k] = ej_pos[3 * i + k]; }
        e.eps = energy[i];
        e.t = time[i];
        e.tfree = 1.0;
        e.nue = 1.0;
        e.stream = (uint64_t)i;
        j.eps = ej_energy[i];
        j.stream = (uint64_t)(n + i);
    }
    m.counters.assign(m.P, 0);
    m.gainP.assign(m.P, 0.0);
    m.lossP.assign(m.P, 0.0);
    m.energyGainField = m.energyGrowth = 0;
    m.total = m.nullc = 0;
    non_parallel_tasks(m);
    for (int64_t i = 0; i < n; ++i) {
        const Electron& e = m.el[i];
        for (int k = 0; k < 3; ++k) pos[3 * i + k] = e.r[k];
        energy[i] = e.eps;
        time[i] = e.t;
    }
    if (growth) *growth = m.energyGrowth;
    if (counters) for (int p = 0; p < m.P; ++p) counters[p] = m.counters[p];
    m.num.seed = seed0;
    return EKMC_OK;
}

int orc_check_steady_state(orc_model* mp, int32_t n, const double* times, const double* mean_energy,
                           double time, uint64_t total_collisions) {
    orc_model& m = *mp;
    if (n < 1) return -1;
    if (m.f.omega != 0) return -1;  // DC models only (no phase rows set up)
    m.nSamp = n;
    m.idx = n - 1;
    m.sT.assign(times, times + n);
    m.sMeanE.assign(mean_energy, mean_energy + n);
    m.time = time;
    m.total = total_collisions;
    m.ssTime = kNonDef;
    if (m.maxElecEnergy <= 0) m.maxElecEnergy = 1.0;
    if (m.nCells <= 0) m.nCells = m.num.n_energy_cells;
    m.nE = m.cond.n_electrons;  // maxCollisionsBeforeSteadyState is per electron
    m.el.assign(m.nE, Electron{});
    check_steady_state(m);
    return m.ssTime != kNonDef ? 1 : 0;
}

double orc_statistical_error(const double* x, int64_t n, double n_bins) {
    std::vector<double> a(x, x + n);
    return stat_error_strided(a, 1, 0, 0, (int)n, n_bins);
}

void orc_histogram(const double* x, int64_t n, const double* nodes, int32_t n_nodes, double* counts) {
    std::vector<double> nd(nodes, nodes + n_nodes);
    histogram_count(x, n, nd, counts);
}

}  // extern "C"
