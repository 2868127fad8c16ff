// common.h -- structures shared by the host engine and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace ekmc {

// Constant.h:6-23 (CODATA 2014, as the reference uses them)
constexpr double kBoltzmann = 1.38064852e-23;
constexpr double kCharge = 1.6021766208e-19;
constexpr double kMass = 9.10938356e-31;
constexpr double kNonDef = -123456789.0;
constexpr double kPi = 3.14159265358979323846;

constexpr int kBlock = 256;            // threads per block of every grid-wide kernel
constexpr int kWave = 64;              // CDNA wavefront
constexpr int kChunkDefault = 1024;    // main electrons per wave work unit (64 lanes x 16)
constexpr int kChunkMin = 256;         // smallest default work unit (small batches)
constexpr uint32_t kMaxLoopPerSync = 1u << 22;  // flight cap per electron per interval
constexpr int kEnvBins = 256;          // bins of the null-test envelope of the total frequency

// Field cases of accelerateElectron (BoltzmannMC.C:809-897)
enum FieldCase : int { FIELD_DC = 0, FIELD_AC = 1, FIELD_DC_B = 2, FIELD_ECR = 3, FIELD_AC_B = 4 };
// Flight-kernel class of a launch that takes both DC and DC+B conditions
// (the flight formula is chosen per chunk); a batch holding both shares one
// work queue, so the end of the launch idles once instead of twice.
constexpr int FLIGHT_DC_ANY = 5;
constexpr int kFlightClasses = 6;

// Philox counter purposes
enum Purpose : uint32_t { P_EVENT = 0, P_INIT = 1, P_POPCTL = 2, P_CHILD = 3, P_CLONE = 4 };

// Status bits raised by kernels (per condition)
enum StatusBits : uint32_t {
    ST_REGION_OVERFLOW = 1u,   // a chunk produced more newborns than its region holds
    ST_ZERO_VELOCITY = 2u,     // cart2sph of a null vector
    ST_LOOP_CAP = 4u,          // flight cap reached inside one interval
    ST_POPCTL = 8u,            // population control could not complete (marks < removals)
    ST_NONFINITE = 16u,        // an electron clock or energy became NaN/inf
};

// Constants of one process, one 64-byte record per process: a collision
// reads the record of the process it chose (one pointer for the kernels
// instead of one per field).
struct ProcRec {
    int32_t type;        // 0 conservative, 1 ionization, 2 attachment
    int32_t scat;        // angular scattering model
    double params[2];    // scattering parameters
    double energyLoss;   // [eV], negative for superelastic
    double tMass;        // target mass [kg]
    double rMass;        // reduced mass [kg]
    double thermalStd;   // sqrt(k Tg / M) [m/s]
    double wParam;       // [eV]
};
static_assert(sizeof(ProcRec) == 64, "process record is 64 bytes");

// Per-process constants, device resident.
struct DevProcesses {
    int P, G;
    int gasTempEffect;       // EKMC_GASTEMP_*
    int ionizOp;             // EKMC_IONIZ_*
    double sharingFactor;    // 0.5 equalSharing, 0 oneTakesAll (BoltzmannMC.h:262-270)
    int parkLanes;           // lanes waiting for a real collision that start the collision step
    const ProcRec* proc;     // [P]
    const int* gasFirst;     // [G]
    const int* gasLast;      // [G]
    const double* gasFrac;   // [G]
};

// One working condition as the kernels see it.
struct CondParams {
    // fields (BoltzmannMC.C:453-489)
    double E[3];
    double accel[3];
    double omega, cyclotron;
    int fieldCase;
    int I;                   // table rows
    int envShift;            // rows per null-test envelope bin = 1 << envShift
    // FP32 copies of the field constants for the mixed-form flights
    float fA[3];             // -e E / m_e
    float fE0, fE2;          // E_x, E_z
    float fW, fInvW;         // cyclotron frequency and its inverse
    float fvEx;              // e E_x / (m_e W): E x B drift speed
    float fAz;               // e E_z / m_e
    float fEmw;              // e / (m_e omega)
    float fInvOmega;         // 1 / omega
    double freq;             // omega / (2 pi) [Hz]
    double N;                // gas density [m^-3]
    double gasEnergy;        // 1.5 k Tg / e [eV]
    // per interval
    double trial;            // trial collision frequency [1/s]
    double trialInv;         // 1 / trial
    double tsync;            // synchronization time of this interval [s]
    double step;             // table energy step [eV]
    uint64_t seed;           // Philox key of this condition
    uint64_t syncIndex;      // interval counter (population-control keys)
    int64_t tableOffset;     // element offset of this condition's I x P tables
    // pool layout
    int64_t base;            // first main slot
    int64_t nMain;           // N_c
    int64_t regionBase;      // first newborn-region slot
    int32_t chunkBegin;      // first chunk of this condition
    int32_t nChunks;         // chunks of this condition
    int32_t active;          // 0: condition finished, kernels skip it
    int32_t pad;
};

// Per-condition population-control plan, written by device planning kernels.
struct PopPlan {
    int64_t holes;       // unfilled main slots after the local refill
    int64_t surplus;     // newborns still in regions
    int64_t alive;       // M
    int64_t remove;      // R = max(0, M - N)
    int64_t marks;       // K
    int64_t thrIndex;    // removal threshold: mark j goes iff (key_j, j) <= (thrKey, thrIndex)
    double markProb;     // p
    uint32_t status;
    uint32_t thrKey;
};

// Per-condition event record of one interval (summary of the flight and
// population-control kernels), in doubles: [0,P) counts, [P,2P) gains,
// [2P,3P) losses, then fieldGain, nNull, nReal, nEvents, nEscape, maxEnergy,
// growth, pad.
inline int event_record_size(int P) { return 3 * P + 8; }
enum EventTail { EV_FIELD = 0, EV_NULL = 1, EV_REAL = 2, EV_EVENTS = 3, EV_ESCAPE = 4, EV_MAXE = 5,
                 EV_GROWTH = 6, EV_PAD = 7 };

// Per-block moment record of the sampling kernel (FP64):
// sum eps, sum r[3], sum v[3], sum r r^T [9], sum r v^T [9], count, max eps
constexpr int kMomentRecord = 1 + 3 + 3 + 9 + 9 + 1 + 1;

}  // namespace ekmc
