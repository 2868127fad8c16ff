// engine.cpp -- host engine of the MI355X electron Monte Carlo path and the
// C ABI of include/ekmc.h.
//
// The host keeps what LoKI-MC's BoltzmannMC keeps between passes: the FP64
// collision tables and trial frequency (checkMaxCollisionFrequency,
// BoltzmannMC.C:716-763), the sampled series and the steady-state / error /
// stop logic of evaluateEEDF (BoltzmannMC.C:299-426, 1743-1893), the integer
// histogram sums and their re-binning (BoltzmannMC.C:1492-1548), and the time
// averages (BoltzmannMC.C:1484-1741, 1950-2266).  Electrons, flights,
// collisions, population control and all per-electron reductions live on the
// GPU (kernels.hip); one interval of every running condition is one batch of
// launches followed by one small device-to-host copy of per-condition sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ekmc.h"
#include "common.h"
#include "kernels.h"
#include "tables.h"

namespace ekmc {
namespace {

thread_local std::string g_error;

ekmc_status fail(ekmc_status st, const std::string& msg) {
    g_error = msg;
    return st;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return fail(EKMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));       \
    } while (0)

template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void**)&p, count * sizeof(T));
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

// Device staging of one event of the collision operator: velocity [3],
// energy, trial frequency, variates [11], ejected velocity [3], ejected
// energy, energy change (FP64) and the process id (int32).
constexpr size_t kCollideStageBytes = (3 + 1 + 1 + 11 + 3 + 1 + 1) * sizeof(double) + sizeof(int32_t);

// Host-side phases reported after the kernels by ekmc_get_kernel_stats.
enum HostPhase { H_PREPARE = 0, H_WAIT, H_POST, H_TABLES, H_FLUSH, H_LAUNCH, H_COUNT };
const char* kHostNames[H_COUNT] = {"host_prepare", "host_wait", "host_post", "host_tables", "host_hist_flush",
                                   "host_launch"};
using Clock = std::chrono::steady_clock;
inline double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// Page-locked host buffer (hipHostMalloc) for the per-interval copies.
template <typename T> struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        n = count;
        if (count == 0) return hipSuccess;
        return hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault);
    }
    void release() { if (p) (void)hipHostFree(p); p = nullptr; n = 0; }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

std::vector<double> linspaced(int n, double a, double b) {
    // Eigen's LinSpaced(n, a, b): a + i*(b-a)/(n-1), the last point exactly b
    std::vector<double> out(n);
    if (n == 1) { out[0] = b; return out; }
    double step = (b - a) / (n - 1);
    for (int i = 0; i < n; ++i) out[i] = (i == n - 1) ? b : a + i * step;
    return out;
}

// Electron energy [eV] of a stored velocity, in the arithmetic of the
// device's energy_of (the engine stores velocities, not energies).
inline double host_energy(double vx, double vy, double vz) {
#pragma clang fp contract(off)
    return 0.5 * kMass * (vx * vx + vy * vy + vz * vz) / kCharge;
}
inline double host_energy(float vx, float vy, float vz) {
#pragma clang fp contract(off)
    return (double)((float)(0.5 * kMass / kCharge) * (vx * vx + vy * vy + vz * vz));
}

// Per-condition host state: the loop variables of BoltzmannMC.
struct Cond {
    double workPerElectron = 1.0;  // collision events per main electron in the last interval
    ekmc_condition in;
    double N = 0, Tg = 0, gasEnergy = 0;
    double E[3] = {0, 0, 0}, accel[3] = {0, 0, 0}, omega = 0, cyc = 0;
    int fieldCase = FIELD_DC;
    bool cylSym = true;
    double angle = 180;
    Tables tab;
    double time = 0, tsync = 0, trial = 0, maxElecEnergy = 0, lastMaxE = 0, ssTime = kNonDef;
    int nSamp = 0, nSync = 0, nIntPts = 0, firstIntIdx = 0;
    double totalIntTime = 0, timeIntervalInteg = 0;
    uint64_t syncIndex = 0;
    std::vector<double> sT, sE, sR, sV, sFlux, sCov, sBulkV, sBulkD;
    std::vector<uint64_t> counts;
    std::vector<double> gainP, lossP;
    double fieldGain = 0, growth = 0;
    uint64_t total = 0, nullc = 0, atSS = 0, nullAtSS = 0, afterSS = 0, events = 0, escapes = 0;
    // distributions
    double maxEedfE = 0, eedfStep = 0, cosStep = 0, radStep = 0, axStep = 0, maxSpeed = 0;
    std::vector<double> eedfNodes, eedfCells, eehSum, eehPer, eahSum, evhSum, cosCells, radCells, axCells;
    double v3Step = 0;                   // 3D EVDF cell width (evdf3d_cells > 0)
    std::vector<double> v3Cells, ev3Sum;  // its cell centres per axis and counts
    int histRequest = 0;   // 1: histogram the current state before the next interval
    int histPhase = -1;
    int nPhases = 0;
    double phaseStep = 0;
    std::vector<double> perN, perE, perV, perBulkV, perFlux, perBulkD;
    // control
    bool finished = false, stoppedByMax = false, good = false;
    uint32_t status = 0;
    // averages
    double avgE = kNonDef, avgEErr = 0, fluxV[3] = {0}, fluxVErr[3] = {0}, fluxD[9] = {0}, fluxDErr[9] = {0};
    double bulkV[3] = {0}, bulkVErr[3] = {0}, bulkD[9] = {0}, bulkDErr[9] = {0};
    double powerBalanceRelError = 0;
    double elapsed = 0;
    std::chrono::high_resolution_clock::time_point start;
};

class EngineBase {
public:
    virtual ~EngineBase() = default;
    virtual ekmc_status create(const ekmc_mixture& mix, const ekmc_condition* conds, int n,
                               const ekmc_numerics& num) = 0;
    virtual ekmc_status initialize() = 0;
    virtual ekmc_status step(int nsync, int* running) = 0;
    virtual ekmc_status finalize() = 0;
    virtual ekmc_status get_electrons(int c, double* pos, double* vel, double* e, double* t) = 0;
    virtual ekmc_status op_accelerate(int c, int64_t n, const double* t, const double* dt, double* pos,
                                      double* vel, double* e, double* g) = 0;
    virtual ekmc_status op_tables(int c, double maxE, double* interp, double* cumul, double* tot, double* mx,
                                  double* step) = 0;
    virtual ekmc_status op_collide(int c, double maxE, int64_t n, double* vel, double* e, const double* nue,
                                   const double* var, int32_t* pid, double* ejv, double* eje, double* de) = 0;
    // The same operator with its device staging at `stage` (caller-provided
    // device memory of at least n * kCollideStageBytes bytes); allocates nothing.
    virtual ekmc_status op_collide_staged(int c, double maxE, int64_t n, double* vel, double* e, const double* nue,
                                          const double* var, int32_t* pid, double* ejv, double* eje, double* de,
                                          void* stage) = 0;
    int device = 0;  // the HIP device the engine was created on
    // A copy of one condition's electron pool taken after the flights of the
    // next interval and before its population control (ekmc_op_interval_pool).
    struct PoolSnap {
        int cond = -1;
        bool taken = false;
        double tsync = 0, trial = 0;
        std::vector<double> pos, vel, energy, time;
        std::vector<uint64_t> stream;
        std::vector<uint32_t> event;
    };
    PoolSnap* snap = nullptr;
    ProcessList pl;
    ekmc_numerics num;
    std::vector<Cond> cs;
    bool initialized = false, finalized = false;
    std::vector<std::pair<uint64_t, double>> kstat;  // per kernel id: launches, ms
    uint64_t totalEvents = 0, totalReal = 0;
};

}  // namespace
}  // namespace ekmc

struct ekmc_engine {
    std::unique_ptr<ekmc::EngineBase> impl;
};

namespace ekmc {
namespace {

// -------------------------------------------------------- statistics (host)
double mean_of(const std::vector<double>& a, int w, int col, int first, int n) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[(size_t)(first + i) * w + col];
    return s / n;
}
// MathFunctions::statisticalError (MathFunctions.C:211-231): 50 batch means
double stat_error(const std::vector<double>& a, int w, int col, int first, int n, double nBins) {
    if (n == 0) return 0;
    int nb = (int)nBins, np = (int)(n / nBins);
    double mean = mean_of(a, w, col, first, n), sum = 0;
    for (int b = 0; b < nb; ++b) sum += std::pow(mean_of(a, w, col, first + b * np, np) - mean, 2);
    return std::sqrt(sum) / nBins;
}

double max_accel_energy(const Cond& c, double initialEnergy, double deltaT) {
    // maximizationAccelerationEnergy (BoltzmannMC.C:765-802)
    double e = kCharge, me = kMass, e_me = e / me;
    double Ex0 = std::abs(c.E[0]), Ez0 = std::abs(c.E[2]);
    double Ex02 = Ex0 * Ex0, Ez02 = Ez0 * Ez0, E02 = Ex02 + Ez02, E0 = std::sqrt(E02);
    double v0 = std::sqrt(initialEnergy * e * 2.0 / me);
    double w = c.omega, W = c.cyc;
    double g = (E0 * v0 + 0.5 * e_me * E02 * deltaT) * deltaT;
    if (W == 0) {
        if (w != 0) g = std::fmin(g, 2.0 / w * (e_me * E02 / w + v0 * (Ex0 + Ez0)));
    } else if (w == 0) {
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
    return initialEnergy + g;
}

// =============================================================== Engine<V>
template <typename V> class Engine : public EngineBase {
public:
    hipStream_t stream = nullptr;
    int I = 10000, P = 0, nCond = 0;
    int64_t capacity = 0, totalMain = 0;
    int nChunks = 0, cus = 256, envShift = 0;
    std::vector<CondParams> hcp;
    std::vector<Chunk> hchunks;
    // device buffers
    DevBuf<RecA<V>> recA;  // electron records (FP64: their sampled part)
    DevBuf<RecB<V>> recB;  // FP64: the draw part of the records
    DevBuf<uint32_t> holeList, markKeys, status, queue;
    DevBuf<uint8_t> alive;
    DevBuf<V> tables;
    DevBuf<CondParams> dconds;
    DevBuf<Chunk> chunks;
    DevBuf<int32_t> holesC, surplusC, markC, flist, phaseRow, cylFlag;
    DevBuf<int64_t> holeOff, surOff, markOff, mainBase;
    DevBuf<double> tailC, momC, summary, thermalDev, grid;
    DevBuf<int32_t> gCond, gFirst, gCount, cGroup, cGroups;  // reduction groups (huge conditions)
    DevBuf<double> tailG, momG;
    int nGroups = 0;
    DevBuf<unsigned long long> accCount;
    DevBuf<long long> accGain, accLoss;
    DevBuf<PopPlan> plans;
    FlightLists fl = {};
    std::vector<int32_t> listActive;  // active flags the flight lists were built for
    int listAge = 0;                  // intervals since the flight lists were ordered
    // histogram grid and phase rows as last uploaded (staged in pinned memory)
    PinnedBuf<double> pinGrid;
    PinnedBuf<int32_t> pinRow;
    std::vector<double> upGrid;
    std::vector<int32_t> upRow;
    bool fusedHist = true;            // running conditions' histograms counted by the flight kernel
    bool mergeDC = true;              // DC and DC+B conditions in one flight launch
    DevBuf<ProcRec> prec;
    DevBuf<int> iproc;
    DevBuf<double> dproc;
    DevBuf<unsigned long long> hEedf, hPer, hEadf, hEvdf, hEvdf3;
    DevProcesses dpr;
    DevState<V> ds;
    DevWork dw;
    DevHist dh;
    int summarySize = 0, recSize = 0;
    int nCells = 1000, nCos = 100, nRad = 200, nAx = 200, nPhases = 0;
    int n3 = 0;  // cells per axis of the 3D EVDF (0: none)
    bool anyCyl = false, anyAC = false;
    PinnedBuf<double> hsum;         // per-condition summaries of the last interval
    PinnedBuf<CondParams> pinCond;  // staging of the per-interval condition upload
    PinnedBuf<double> pinNumax;     // running maximum frequency of a rebuilt table
    DevBuf<double> tsrcD, nutot64, numax64;
    DevBuf<int> tsrcI;
    DevTableSrc tsrc = {};
    DevBuf<SlotRange> ranges;
    KernelTimer timer;
    // the cycle after the flights as a HIP graph (captured at the first
    // interval, replayed at every later one)
    bool useGraph = true;
    hipGraph_t cycGraph = nullptr;
    hipGraphExec_t cycExec = nullptr;

    ~Engine() override {
        if (stream) (void)hipStreamSynchronize(stream);
        set_kernel_timer(nullptr);
        if (cycExec) (void)hipGraphExecDestroy(cycExec);
        if (cycGraph) (void)hipGraphDestroy(cycGraph);
        for (int k = 0; k < K_COUNT; ++k) {
            if (timer.start[k]) (void)hipEventDestroy(timer.start[k]);
            if (timer.stop[k]) (void)hipEventDestroy(timer.stop[k]);
        }
        recA.release(); recB.release(); holeList.release(); markKeys.release(); status.release(); queue.release();
        alive.release(); tables.release(); dconds.release(); chunks.release(); holesC.release();
        surplusC.release(); markC.release(); flist.release(); phaseRow.release(); cylFlag.release();
        holeOff.release(); surOff.release(); markOff.release(); mainBase.release(); tailC.release();
        momC.release(); summary.release(); thermalDev.release(); grid.release(); accCount.release();
        gCond.release(); gFirst.release(); gCount.release(); cGroup.release(); cGroups.release();
        tailG.release(); momG.release();
        accGain.release(); accLoss.release(); plans.release(); prec.release(); iproc.release(); dproc.release();
        hEedf.release(); hPer.release(); hEadf.release(); hEvdf.release(); hEvdf3.release(); ranges.release();
        hsum.release(); pinCond.release(); pinNumax.release(); tsrcD.release(); tsrcI.release();
        pinGrid.release(); pinRow.release();
        nutot64.release(); numax64.release();
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ------------------------------------------------------------ create
    ekmc_status create(const ekmc_mixture& mix, const ekmc_condition* conds, int n,
                       const ekmc_numerics& nm) override {
        num = nm;
        std::string err = build_process_list(mix, pl);
        if (!err.empty()) return fail(EKMC_ERR_INVALID_ARGUMENT, err);
        P = pl.P;
        if (P > 2048) return fail(EKMC_ERR_INVALID_ARGUMENT, "at most 2048 processes are supported");
        I = num.n_interp_points;
        if (I < 2) return fail(EKMC_ERR_INVALID_ARGUMENT, "nInterpPoints must be at least 2");
        if (num.sync_over_sampling < 1) return fail(EKMC_ERR_INVALID_ARGUMENT, "synchronizationOverSampling must be >= 1");
        if (!(num.sync_time_x_max_coll_freq > 0)) return fail(EKMC_ERR_INVALID_ARGUMENT, "synchronizationTimeXMaxCollisionFrequency must be > 0");
        nCond = n;
        nCells = num.n_energy_cells; nCos = num.n_cos_angle_cells;
        nRad = num.n_radial_velocity_cells; nAx = num.n_axial_velocity_cells;
        if (nCells < 1 || nCos < 1 || nRad < 1 || nAx < 1) return fail(EKMC_ERR_INVALID_ARGUMENT, "histogram sizes must be positive");
        if (num.park_lanes < 0 || num.park_lanes > kWave)
            return fail(EKMC_ERR_INVALID_ARGUMENT, "park_lanes must be 0 (default) or 1-64");
        if (num.env_coarsen < 0 || num.env_coarsen > 3) return fail(EKMC_ERR_INVALID_ARGUMENT, "env_coarsen must be 0-3");
        if (num.work_chunk < 0) return fail(EKMC_ERR_INVALID_ARGUMENT, "work_chunk must not be negative");
        if (num.evdf3d_cells < 0 || num.evdf3d_cells > 256)
            return fail(EKMC_ERR_INVALID_ARGUMENT, "evdf3d_cells must be 0-256");
        n3 = num.evdf3d_cells;
        // the flight kernel's LDS EEDF counts a chunk's main electrons in 16 bits
        if (num.work_chunk > 65535) return fail(EKMC_ERR_INVALID_ARGUMENT, "work_chunk must be at most 65535");
        int dev = 0;
        if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0) return fail(EKMC_ERR_NO_DEVICE, "no HIP device is visible");
        HIP_TRY(hipGetDevice(&device));
        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (int k = 0; k < K_COUNT; ++k) {
            HIP_TRY(hipEventCreate(&timer.start[k]));
            HIP_TRY(hipEventCreate(&timer.stop[k]));
            timer.armed[k] = false;
        }
        // EKMC_FUSED_HIST=0: every histogram by the histogram kernel
        const char* fh = std::getenv("EKMC_FUSED_HIST");
        fusedHist = !(fh && fh[0] == '0');
        const char* md = std::getenv("EKMC_MERGE_DC");
        mergeDC = !(md && md[0] == '0');
        // The cycle after the flights is replayed from a HIP graph for a batch
        // of conditions and launched kernel by kernel for one condition (the
        // graph measured 0.6 % faster per interval on the 64-condition scan
        // and 2.7 % slower on the one-condition engines of the same scan).
        // EKMC_GRAPH=1 or 0 forces either.
        const char* gr = std::getenv("EKMC_GRAPH");
        useGraph = gr && gr[0] ? gr[0] != '0' : n > 1;
        // HIP-event timing of the flight kernels; EKMC_KERNEL_TIMING=all times
        // every kernel, EKMC_KERNEL_TIMING=0 none
        const char* kt = std::getenv("EKMC_KERNEL_TIMING");
        timer.level = !kt ? 1 : (std::strcmp(kt, "all") == 0 ? 2 : (kt[0] == '0' ? 0 : 1));
        kstat.assign(K_COUNT + H_COUNT, {0, 0.0});
        cs.assign(n, Cond());
        int64_t sumN = 0;
        for (int c = 0; c < n; ++c) {
            const ekmc_condition& in = conds[c];
            if (in.n_electrons <= 0) return fail(EKMC_ERR_INVALID_ARGUMENT, "nElectrons must be positive");
            if (in.reduced_elec_field < 0 || in.reduced_mag_field < 0 || in.excitation_frequency < 0 ||
                in.gas_temperature <= 0 || in.gas_density <= 0)
                return fail(EKMC_ERR_INVALID_ARGUMENT, "working conditions must be non-negative (WorkingConditions.h:144-178)");
            if (in.elec_field_angle < 0 || in.elec_field_angle > 180)
                return fail(EKMC_ERR_INVALID_ARGUMENT, "elecFieldAngle must be in [0,180] degrees");
            Cond& C = cs[c];
            C.in = in;
            C.N = in.gas_density;
            C.Tg = in.gas_temperature;
            C.gasEnergy = 1.5 * kBoltzmann * C.Tg / kCharge;
            double Ev = in.reduced_elec_field * 1e-21 * C.N;
            C.angle = in.elec_field_angle;
            if (C.angle == 180) { C.E[0] = 0; C.E[1] = 0; C.E[2] = -Ev; }
            else { C.E[0] = Ev * std::sin(C.angle / 180.0 * kPi); C.E[1] = 0; C.E[2] = Ev * std::cos(C.angle / 180.0 * kPi); }
            C.omega = in.excitation_frequency * 2.0 * kPi;
            if (in.excitation_frequency != 0) for (double& v : C.E) v *= std::sqrt(2);
            C.cyc = kCharge * (in.reduced_mag_field * 1e-27 * C.N) / kMass;
            for (int k = 0; k < 3; ++k) C.accel[k] = -kCharge / kMass * C.E[k];
            if (in.reduced_mag_field == 0) C.fieldCase = in.excitation_frequency == 0 ? FIELD_DC : FIELD_AC;
            else if (in.excitation_frequency == 0) C.fieldCase = FIELD_DC_B;
            else if (std::abs(C.omega - C.cyc) / C.omega < 1E-6) C.fieldCase = FIELD_ECR;
            else C.fieldCase = FIELD_AC_B;
            C.cylSym = C.angle == 180;
            anyCyl |= C.cylSym;
            anyAC |= C.omega != 0;
            sumN += in.n_electrons;
        }
        nPhases = anyAC ? num.n_integration_phases : 0;
        if (anyAC && nPhases < 1) return fail(EKMC_ERR_INVALID_ARGUMENT, "nIntegrationPhases must be >= 1");
        // ---- work layout: chunks of ~work_chunk main electrons, one wavefront each
        int devId = 0;
        (void)hipGetDevice(&devId);
        hipDeviceProp_t prop;
        size_t ldsMax = 65536;
        if (hipGetDeviceProperties(&prop, devId) == hipSuccess) {
            if (prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
            if (prop.sharedMemPerBlock > 0) ldsMax = prop.sharedMemPerBlock;
        }
        // LDS of a flight-kernel block: the process records, the per-wave
        // tallies, histogram grid and envelope, and the record rings; the EEDF
        // is counted in LDS only where it fits
        if ((size_t)P * sizeof(ProcRec) + (size_t)(kBlock / kWave) * (3 * P * 8 + 64 + 2 * kEnvBins * sizeof(V)) +
                flight_ring_lds<V>() >
            ldsMax)
            return fail(EKMC_ERR_INVALID_ARGUMENT, "too many processes for the per-wave collision tallies in LDS");
        // Default work unit: 1024 main electrons per wave where that still
        // gives 12 waves per CU a chunk (the mixed flight kernel's residency,
        // 3 blocks of 4 waves), else the batch split over that many waves,
        // at least 256 (per-chunk work and newborn regions grow as chunks
        // shrink).  The FP64 kernel holds 8 waves per CU, so it gets about
        // 1.5 chunks per wave (its engines were tuned with this rule).  On one
        // MI355X a batch of more than 3.1 M electrons keeps 1024; the
        // one-condition engines of the 64-condition scan (1e4-1e6 electrons)
        // ran 24 % faster with 256 than with 1024.
        int64_t chunkTarget = num.work_chunk;
        if (chunkTarget == 0) {
            int64_t total = 0;
            for (int c = 0; c < n; ++c) total += conds[c].n_electrons;
            const int64_t waves = (int64_t)cus * 3 * (kBlock / kWave);
            chunkTarget = total >= waves * kChunkDefault
                              ? kChunkDefault
                              : std::max<int64_t>(kChunkMin, (total + waves - 1) / waves);
        }
        chunkTarget = std::max<int64_t>(kWave, (chunkTarget + kWave - 1) / kWave * kWave);
        envShift = 0;  // envelope bins of 2^envShift table rows, at most kEnvBins of them
        while (((I - 1) >> envShift) >= kEnvBins) ++envShift;
        envShift += num.env_coarsen;
        hcp.assign(n, CondParams());
        hchunks.clear();
        int64_t base = 0;
        for (int c = 0; c < n; ++c) {
            const int64_t Nc = conds[c].n_electrons;
            const int64_t nch = (Nc + chunkTarget - 1) / chunkTarget;
            if ((int64_t)hchunks.size() + nch > 0x7fffffffLL) return fail(EKMC_ERR_CAPACITY, "too many work chunks");
            CondParams& cp = hcp[c];
            std::memset(&cp, 0, sizeof(cp));
            for (int k = 0; k < 3; ++k) { cp.E[k] = cs[c].E[k]; cp.accel[k] = cs[c].accel[k]; }
            cp.omega = cs[c].omega;
            cp.cyclotron = cs[c].cyc;
            cp.fieldCase = cs[c].fieldCase;
            for (int k = 0; k < 3; ++k) cp.fA[k] = (float)cs[c].accel[k];
            cp.fE0 = (float)cs[c].E[0];
            cp.fE2 = (float)cs[c].E[2];
            cp.fW = (float)cs[c].cyc;
            cp.fInvW = cs[c].cyc != 0 ? (float)(1.0 / cs[c].cyc) : 0.0f;
            cp.fvEx = cs[c].cyc != 0 ? (float)(kCharge * cs[c].E[0] / (kMass * cs[c].cyc)) : 0.0f;
            cp.fAz = (float)(kCharge * cs[c].E[2] / kMass);
            cp.fEmw = cs[c].omega != 0 ? (float)(kCharge / (kMass * cs[c].omega)) : 0.0f;
            cp.fInvOmega = cs[c].omega != 0 ? (float)(1.0 / cs[c].omega) : 0.0f;
            cp.freq = cs[c].omega / (2.0 * kPi);
            cp.I = I;
            cp.N = cs[c].N;
            cp.gasEnergy = cs[c].gasEnergy;
            cp.seed = num.seed ^ ((uint64_t)c * 0x9E3779B97F4A7C15ull);
            cp.tableOffset = (int64_t)c * table_block_size(I, P);
            cp.envShift = envShift;
            cp.base = base;
            cp.nMain = Nc;
            cp.regionBase = base + Nc;
            cp.chunkBegin = (int32_t)hchunks.size();
            cp.nChunks = (int32_t)nch;
            cp.active = 1;
            // balanced chunk sizes; each chunk's newborn region holds twice its electrons
            int64_t first = base, region = base + Nc;
            for (int64_t j = 0; j < nch; ++j) {
                Chunk ch;
                ch.cond = c;
                ch.count = (int32_t)(Nc / nch + (j < Nc % nch ? 1 : 0));
                ch.first = first;
                ch.region = region;
                ch.regionCap = (int32_t)std::max<int64_t>(256, (2 * (int64_t)ch.count + kWave - 1) / kWave * kWave);
                ch.local = (int32_t)j;
                first += ch.count;
                region += ch.regionCap;
                hchunks.push_back(ch);
            }
            base = region;
        }
        nChunks = (int)hchunks.size();
        capacity = base;
        totalMain = sumN;
        if (capacity > 0xFFFFFFF0LL) return fail(EKMC_ERR_CAPACITY, "electron pool exceeds 2^32 slots");
        // ---- device allocations
        HIP_TRY(recA.alloc(capacity));
        if (split_record<V>()) HIP_TRY(recB.alloc(capacity));
        HIP_TRY(alive.alloc(capacity));
        HIP_TRY(holeList.alloc(capacity));
        HIP_TRY(markKeys.alloc(capacity));
        HIP_TRY(hipMemsetAsync(alive.p, 0, capacity, stream));
        HIP_TRY(tables.alloc((size_t)n * table_block_size(I, P)));
        HIP_TRY(dconds.alloc(n));
        HIP_TRY(chunks.alloc(nChunks));
        HIP_TRY(hipMemcpyAsync(chunks.p, hchunks.data(), nChunks * sizeof(Chunk), hipMemcpyHostToDevice, stream));
        HIP_TRY(holesC.alloc(nChunks)); HIP_TRY(surplusC.alloc(nChunks)); HIP_TRY(markC.alloc(nChunks));
        HIP_TRY(holeOff.alloc(nChunks)); HIP_TRY(surOff.alloc(nChunks)); HIP_TRY(markOff.alloc(nChunks));
        HIP_TRY(flist.alloc(nChunks));
        HIP_TRY(hipMemsetAsync(holesC.p, 0, nChunks * sizeof(int32_t), stream));
        HIP_TRY(hipMemsetAsync(surplusC.p, 0, nChunks * sizeof(int32_t), stream));
        HIP_TRY(hipMemsetAsync(markC.p, 0, nChunks * sizeof(int32_t), stream));
        recSize = event_record_size(P);
        summarySize = recSize + kMomentRecord + 8;
        HIP_TRY(tailC.alloc((size_t)nChunks * CT_SIZE));
        HIP_TRY(momC.alloc((size_t)nChunks * kMomentRecord));
        HIP_TRY(hipMemsetAsync(tailC.p, 0, (size_t)nChunks * CT_SIZE * sizeof(double), stream));
        // reduction groups of at most kFinishGroup work units, when a condition
        // has more units than one block reduces quickly (one condition of 1e8
        // electrons: 98,000 units, 1.3 ms in one block)
        {
            int maxCh = 0;
            for (int c = 0; c < n; ++c) maxCh = std::max(maxCh, (int)hcp[c].nChunks);
            if (maxCh > kFinishDirect) {
                std::vector<int32_t> gc, gf, gn, cg(n), cn(n);
                for (int c = 0; c < n; ++c) {
                    cg[c] = (int32_t)gc.size();
                    for (int j = 0; j < hcp[c].nChunks; j += kFinishGroup) {
                        gc.push_back(c);
                        gf.push_back(hcp[c].chunkBegin + j);
                        gn.push_back(std::min(kFinishGroup, (int)hcp[c].nChunks - j));
                    }
                    cn[c] = (int32_t)gc.size() - cg[c];
                }
                nGroups = (int)gc.size();
                HIP_TRY(gCond.alloc(nGroups)); HIP_TRY(gFirst.alloc(nGroups)); HIP_TRY(gCount.alloc(nGroups));
                HIP_TRY(cGroup.alloc(n)); HIP_TRY(cGroups.alloc(n));
                HIP_TRY(tailG.alloc((size_t)nGroups * CT_SIZE)); HIP_TRY(momG.alloc((size_t)nGroups * kMomentRecord));
                HIP_TRY(hipMemcpyAsync(gCond.p, gc.data(), nGroups * sizeof(int32_t), hipMemcpyHostToDevice, stream));
                HIP_TRY(hipMemcpyAsync(gFirst.p, gf.data(), nGroups * sizeof(int32_t), hipMemcpyHostToDevice, stream));
                HIP_TRY(hipMemcpyAsync(gCount.p, gn.data(), nGroups * sizeof(int32_t), hipMemcpyHostToDevice, stream));
                HIP_TRY(hipMemcpyAsync(cGroup.p, cg.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
                HIP_TRY(hipMemcpyAsync(cGroups.p, cn.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
                HIP_TRY(hipStreamSynchronize(stream));
            }
        }
        HIP_TRY(summary.alloc((size_t)n * summarySize));
        HIP_TRY(accCount.alloc((size_t)n * P)); HIP_TRY(accGain.alloc((size_t)n * P)); HIP_TRY(accLoss.alloc((size_t)n * P));
        HIP_TRY(hipMemsetAsync(accCount.p, 0, (size_t)n * P * 8, stream));
        HIP_TRY(hipMemsetAsync(accGain.p, 0, (size_t)n * P * 8, stream));
        HIP_TRY(hipMemsetAsync(accLoss.p, 0, (size_t)n * P * 8, stream));
        HIP_TRY(queue.alloc(8));
        HIP_TRY(hipMemsetAsync(queue.p, 0, 8 * sizeof(uint32_t), stream));
        HIP_TRY(plans.alloc(n));
        HIP_TRY(status.alloc(n));
        HIP_TRY(hipMemsetAsync(status.p, 0, n * sizeof(uint32_t), stream));
        HIP_TRY(hipMemsetAsync(plans.p, 0, n * sizeof(PopPlan), stream));
        // processes: one record per process, plus the per-gas ranges
        // thermal standard deviation uses the gas temperature, shared by all conditions
        double Tg0 = cs[0].Tg;
        for (int c = 1; c < n; ++c)
            if (cs[c].Tg != Tg0 && num.gas_temperature_effect != EKMC_GASTEMP_FALSE)
                return fail(EKMC_ERR_INVALID_ARGUMENT, "batched conditions must share the gas temperature when thermal motion is on");
        std::vector<ProcRec> hrec(P);
        for (int p = 0; p < P; ++p) {
            ProcRec& r = hrec[p];
            r.type = pl.type[p];
            r.scat = pl.scat[p];
            r.params[0] = pl.params[2 * p];
            r.params[1] = pl.params[2 * p + 1];
            r.energyLoss = pl.energyLoss[p];
            r.tMass = pl.tMass[p];
            r.rMass = pl.rMass[p];
            r.thermalStd = std::sqrt(kBoltzmann * Tg0 / pl.tMass[p]);
            r.wParam = pl.wParam[p];
        }
        std::vector<int> hi(pl.gasFirst);
        hi.insert(hi.end(), pl.gasLast.begin(), pl.gasLast.end());
        HIP_TRY(prec.alloc(P));
        HIP_TRY(iproc.alloc(hi.size()));
        HIP_TRY(dproc.alloc(pl.gasFrac.size()));
        HIP_TRY(hipMemcpyAsync(prec.p, hrec.data(), P * sizeof(ProcRec), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(iproc.p, hi.data(), hi.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dproc.p, pl.gasFrac.data(), pl.gasFrac.size() * sizeof(double), hipMemcpyHostToDevice,
                               stream));
        dpr.P = P; dpr.G = pl.G;
        dpr.gasTempEffect = num.gas_temperature_effect;
        dpr.ionizOp = num.ionization_operator;
        dpr.sharingFactor = num.ionization_operator == EKMC_IONIZ_ONE_TAKES_ALL ? 0.0 : 0.5;
        dpr.parkLanes = num.park_lanes ? num.park_lanes : 32;
        dpr.proc = prec.p;
        dpr.gasFirst = iproc.p; dpr.gasLast = iproc.p + pl.G; dpr.gasFrac = dproc.p;
        // collision-table sources: prepared cross sections and interpolation constants
        {
            std::vector<double> td, te, ts;
            std::vector<int> ti(3 * P);
            std::vector<int> off(P, 0);
            for (int p = 0; p < P; ++p) {
                if (pl.isSuper[p]) continue;
                off[p] = (int)te.size();
                te.insert(te.end(), pl.energy[p].begin(), pl.energy[p].end());
                ts.insert(ts.end(), pl.sigma[p].begin(), pl.sigma[p].end());
            }
            for (int p = 0; p < P; ++p) {
                const int q = pl.isSuper[p] ? p - 1 : p;
                ti[p] = off[q];
                ti[P + p] = off[q] + (int)pl.energy[q].size();
                ti[2 * P + p] = pl.isSuper[p];
            }
            const size_t nPts = te.size();
            td.reserve(2 * nPts + 5 * P);
            td.insert(td.end(), te.begin(), te.end());
            td.insert(td.end(), ts.begin(), ts.end());
            for (int p = 0; p < P; ++p) td.push_back(pl.minE[p]);
            for (int p = 0; p < P; ++p) td.push_back(pl.maxE[p]);
            for (int p = 0; p < P; ++p) td.push_back(pl.relDens[p]);
            for (int p = 0; p < P; ++p) td.push_back(pl.statW[p]);
            for (int p = 0; p < P; ++p) td.push_back(pl.isSuper[p] ? pl.minE[p - 1] : 0.0);
            HIP_TRY(tsrcD.alloc(td.size()));
            HIP_TRY(tsrcI.alloc(ti.size()));
            HIP_TRY(hipMemcpy(tsrcD.p, td.data(), td.size() * sizeof(double), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(tsrcI.p, ti.data(), ti.size() * sizeof(int), hipMemcpyHostToDevice));
            tsrc.P = P; tsrc.I = I;
            tsrc.energy = tsrcD.p; tsrc.sigma = tsrcD.p + nPts;
            const double* pd = tsrcD.p + 2 * nPts;
            tsrc.minE = pd; tsrc.maxE = pd + P; tsrc.relDens = pd + 2 * P; tsrc.statW = pd + 3 * P;
            tsrc.thrPartner = pd + 4 * P;
            tsrc.srcBegin = tsrcI.p; tsrc.srcEnd = tsrcI.p + P; tsrc.isSuper = tsrcI.p + 2 * P;
            HIP_TRY(nutot64.alloc(I));
            HIP_TRY(numax64.alloc(I));
            HIP_TRY(pinNumax.alloc(I));
        }
        // histograms
        HIP_TRY(hEedf.alloc((size_t)n * nCells));
        if (anyAC) HIP_TRY(hPer.alloc((size_t)n * nPhases * nCells));
        if (anyCyl) { HIP_TRY(hEadf.alloc((size_t)n * nCells * nCos)); HIP_TRY(hEvdf.alloc((size_t)n * nRad * nAx)); }
        if (n3 > 0) HIP_TRY(hEvdf3.alloc((size_t)n * n3 * n3 * n3));
        // the histogram sums accumulate from zero (device memory of an earlier
        // engine may be handed out again with its contents)
        for (DevBuf<unsigned long long>* b : {&hEedf, &hPer, &hEadf, &hEvdf, &hEvdf3})
            if (b->p) HIP_TRY(hipMemsetAsync(b->p, 0, b->n * sizeof(unsigned long long), stream));
        HIP_TRY(grid.alloc((size_t)n * 6));
        HIP_TRY(phaseRow.alloc(n));
        HIP_TRY(cylFlag.alloc(n));
        std::vector<int32_t> cyl(n);
        for (int c = 0; c < n; ++c) cyl[c] = cs[c].cylSym ? 1 : 0;
        HIP_TRY(hipMemcpyAsync(cylFlag.p, cyl.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        ds = DevState<V>{recA.p, recB.p, alive.p};
        dw.conds = dconds.p; dw.chunks = chunks.p; dw.holesC = holesC.p; dw.surplusC = surplusC.p;
        dw.markC = markC.p; dw.holeOff = holeOff.p; dw.surOff = surOff.p; dw.markOff = markOff.p;
        dw.holeList = holeList.p; dw.markKeys = markKeys.p; dw.tailC = tailC.p; dw.momC = momC.p; dw.accCount = accCount.p;
        dw.accGain = accGain.p; dw.accLoss = accLoss.p; dw.plans = plans.p;
        dw.status = status.p; dw.queue = queue.p; dw.summary = summary.p; dw.summarySize = summarySize;
        dw.recSize = recSize; dw.nChunks = nChunks; dw.nCond = n;
        dw.groupCond = gCond.p; dw.groupFirst = gFirst.p; dw.groupCount = gCount.p;
        dw.condGroup = cGroup.p; dw.condGroups = cGroups.p; dw.tailG = tailG.p; dw.momG = momG.p;
        dw.nGroups = nGroups;
        dh.eedf = hEedf.p; dh.eedfPer = hPer.p; dh.eadf = hEadf.p; dh.evdf = hEvdf.p; dh.grid = grid.p;
        dh.evdf3 = hEvdf3.p; dh.n3 = n3;
        dh.phaseRow = phaseRow.p; dh.cyl = cylFlag.p;
        dh.nCells = nCells; dh.nCos = nCos; dh.nRad = nRad; dh.nAx = nAx; dh.nPhases = nPhases;
        // the EEDF of a histogram taken at load is counted in LDS when that
        // keeps the flight kernel's blocks per CU
        dh.histLds = 0;
        if (nCells <= 2048) {
            const bool thermal = num.gas_temperature_effect != EKMC_GASTEMP_FALSE;
            DevHist withLds = dh, noLds = dh;
            withLds.histLds = nCells;
            noLds.histLds = 0;
            if (flight_blocks_per_cu<V>(FIELD_DC, thermal, P, withLds) >= flight_blocks_per_cu<V>(FIELD_DC, thermal, P, noLds))
                dh.histLds = nCells;
        }
        // otherwise the histogram kernel counts the sampled state: counting at
        // load with HBM atomics contends on the few busy EEDF cells (measured
        // 3.5x slower flights on the 64-condition scan)
        if (dh.histLds < nCells) fusedHist = false;
        // the 3D EVDF is counted by the histogram kernel only
        if (n3 > 0) fusedHist = false;
        // first main electron of each condition, for the init kernel
        std::vector<int64_t> mb(n);
        std::vector<double> td(n);
        int64_t off = 0;
        for (int c = 0; c < n; ++c) {
            mb[c] = off;
            off += conds[c].n_electrons;
            td[c] = std::sqrt(kBoltzmann * num.initial_elec_temp_over_gas_temp * cs[c].Tg / kMass);
        }
        HIP_TRY(mainBase.alloc(n));
        HIP_TRY(thermalDev.alloc(n));
        HIP_TRY(hipMemcpyAsync(mainBase.p, mb.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(thermalDev.p, td.data(), n * sizeof(double), hipMemcpyHostToDevice, stream));
        HIP_TRY(hsum.alloc((size_t)n * summarySize));
        std::memset(hsum.p, 0, hsum.n * sizeof(double));
        HIP_TRY(pinCond.alloc(n));
        HIP_TRY(pinGrid.alloc((size_t)n * 6));
        HIP_TRY(pinRow.alloc(n));
        // histogram blocks: ranges of at most 8192 main slots of one condition
        std::vector<SlotRange> hr;
        for (int c = 0; c < n; ++c)
            for (int64_t o = 0; o < hcp[c].nMain; o += 8192) {
                SlotRange r;
                r.cond = c;
                r.first = hcp[c].base + o;
                r.count = (int32_t)std::min<int64_t>(8192, hcp[c].nMain - o);
                hr.push_back(r);
            }
        HIP_TRY(ranges.alloc(hr.size()));
        HIP_TRY(hipMemcpy(ranges.p, hr.data(), hr.size() * sizeof(SlotRange), hipMemcpyHostToDevice));
        dh.ranges = ranges.p;
        dh.nRanges = (int)hr.size();
        HIP_TRY(hipStreamSynchronize(stream));
        return EKMC_OK;
    }

    // ---------------------------------------------------- device helpers
    // interpolateCrossSections (BoltzmannMC.C:561-615) on the device, into the
    // condition's tables; the running maximum of the total frequency comes
    // back for the trial-frequency check.  A rebuild to the same energy is
    // skipped (BoltzmannMC.C:567-569).
    ekmc_status interpolate(int c, double maxEnergy, bool force = false) {
        Tables& tb = cs[c].tab;
        if (maxEnergy == tb.maxEnergy && !force) return EKMC_OK;
        const auto t0 = Clock::now();
        tb.I = I;
        tb.P = P;
        tb.maxEnergy = maxEnergy;
        tb.step = maxEnergy / (double)(I - 1);
        HIP_TRY(launch_tables<V>(tsrc, cs[c].N, tb.step, envShift, tables.p + hcp[c].tableOffset, nutot64.p,
                                 numax64.p, stream));
        HIP_TRY(hipMemcpyAsync(pinNumax.p, numax64.p, (size_t)I * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        tb.numax.assign(pinNumax.p, pinNumax.p + I);
        host_add(H_TABLES, ms_since(t0));
        return EKMC_OK;
    }

    // checkMaxCollisionFrequency (BoltzmannMC.C:716-763) with the ensemble
    // maximum energy at the start of the interval.
    ekmc_status check_max_collision_frequency(int c) {
        Cond& C = cs[c];
        double thermalC = 0;
        if (num.gas_temperature_effect == EKMC_GASTEMP_TRUE || num.gas_temperature_effect == EKMC_GASTEMP_SMART)
            thermalC = 10.0 * C.gasEnergy;
        bool updated = true;
        while (updated) {
            updated = false;
            double maxEnergy = max_accel_energy(C, C.lastMaxE, 10.0 / C.trial) + thermalC;
            if (maxEnergy > C.tab.maxEnergy || 2.5 * maxEnergy < C.tab.maxEnergy) {
                ekmc_status st = interpolate(c, 2.0 * maxEnergy < pl.energyMaxElastic ? 2.0 * maxEnergy
                                                                                    : pl.energyMaxElastic);
                if (st != EKMC_OK) return st;
            }
            int k = (int)std::fmin(std::ceil(maxEnergy / C.tab.step), I - 1);
            if (C.trial < C.tab.numax[k]) { updated = true; C.trial *= 1.1; }
        }
        return EKMC_OK;
    }

    // The histogram grid and phase rows for the next kernels, uploaded only
    // when they differ from the last upload (the stream was synchronized
    // since that upload, so the pinned staging may be rewritten).
    ekmc_status upload_hist_grid(const std::vector<double>& g, const std::vector<int32_t>& r) {
        if (g == upGrid && r == upRow) return EKMC_OK;
        std::memcpy(pinGrid.p, g.data(), g.size() * sizeof(double));
        std::memcpy(pinRow.p, r.data(), r.size() * sizeof(int32_t));
        HIP_TRY(hipMemcpyAsync(grid.p, pinGrid.p, g.size() * sizeof(double), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(phaseRow.p, pinRow.p, r.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        upGrid = g;
        upRow = r;
        return EKMC_OK;
    }

    void host_add(int ph, double ms) {
        kstat[K_COUNT + ph].first += 1;
        kstat[K_COUNT + ph].second += ms;
    }

    ekmc_status sync_timers() {
        for (int k = 0; k < K_COUNT; ++k) {
            if (!timer.armed[k]) continue;
            float ms = 0;
            if (hipEventElapsedTime(&ms, timer.start[k], timer.stop[k]) == hipSuccess) {
                kstat[k].first += 1;
                kstat[k].second += ms;
            }
            timer.armed[k] = false;
        }
        return EKMC_OK;
    }

    // Flight-kernel work lists of the running conditions, per field case.
    std::vector<int32_t> hlist;
    ekmc_status rebuild_lists() {
        std::vector<int32_t> act(nCond);
        for (int c = 0; c < nCond; ++c) act[c] = cs[c].finished ? 0 : 1;
        // rebuilt when the running set changes, and every 256 intervals for
        // the order (below)
        if (act == listActive && ++listAge < 256) return EKMC_OK;
        listAge = 0;
        listActive = act;
        hlist.clear();
        fl = FlightLists{};
        const bool thermal = num.gas_temperature_effect != EKMC_GASTEMP_FALSE;
        // DC and DC+B conditions share one launch when both are running
        bool anyDC = false, anyDCB = false;
        for (int c = 0; c < nCond; ++c) {
            anyDC |= act[c] && cs[c].fieldCase == FIELD_DC;
            anyDCB |= act[c] && cs[c].fieldCase == FIELD_DC_B;
        }
        const bool merged = anyDC && anyDCB && mergeDC;
        for (int fc = 0; fc < kFlightClasses; ++fc) {
            const size_t begin = hlist.size();
            for (int c = 0; c < nCond; ++c) {
                if (!act[c]) continue;
                const int cls = merged && (cs[c].fieldCase == FIELD_DC || cs[c].fieldCase == FIELD_DC_B) ? FLIGHT_DC_ANY
                                                                                                      : cs[c].fieldCase;
                if (cls != fc) continue;
                for (int j = 0; j < hcp[c].nChunks; ++j) hlist.push_back(hcp[c].chunkBegin + j);
            }
            // more work first (main electrons times the condition's events per
            // electron in its last interval), so that the last chunks taken
            // are the short ones and the waves finish together; the
            // processing order does not change any result (flight kernel per
            // interval on the 64-condition scan: 0.896 ms ordered by size,
            // 0.890 ms by expected work)
            std::stable_sort(hlist.begin() + begin, hlist.end(), [&](int32_t a, int32_t b) {
                return hchunks[a].count * cs[hchunks[a].cond].workPerElectron >
                       hchunks[b].count * cs[hchunks[b].cond].workPerElectron;
            });
            const int32_t count = (int32_t)(hlist.size() - begin);
            fl.list[fc] = flist.p + begin;
            fl.count[fc] = count;
            if (count) {
                const int64_t waves = (int64_t)flight_blocks_per_cu<V>(fc, thermal, P, dh) * cus;
                const int64_t need = (count + kBlock / kWave - 1) / (kBlock / kWave);
                fl.grid[fc] = (int32_t)std::max<int64_t>(1, std::min<int64_t>(need, waves));
            }
        }
        if (!hlist.empty())
            HIP_TRY(hipMemcpyAsync(flist.p, hlist.data(), hlist.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return EKMC_OK;
    }

    // Fixed-order reduction of the chunk records and the download of the
    // per-condition summaries.
    hipError_t launch_reduce() {
        hipError_t e = launch_finish(dw, P, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hsum.p, summary.p, hsum.n * sizeof(double), hipMemcpyDeviceToHost, stream);
        return e;
    }

    // The cycle that follows the flights -- population control, sampling,
    // the reduction and the summary download -- is the same launches with the
    // same arguments every interval (the device state and work descriptors
    // are fixed at creation; what changes lives in device memory), so it is
    // captured once into a HIP graph and replayed.  It holds kernels and one
    // copy only, no memset nodes.  Per-kernel timing (EKMC_KERNEL_TIMING=all)
    // records events between the kernels and runs them one by one.
    ekmc_status launch_cycle() {
        if (!useGraph || timer.level >= 2) {
            HIP_TRY(launch_popctl<V>(ds, dw, stream));
            HIP_TRY(launch_sample<V>(ds, dw, stream));
            HIP_TRY(launch_reduce());
            return EKMC_OK;
        }
        if (!cycExec) {
            HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = launch_popctl<V>(ds, dw, stream);
            if (e == hipSuccess) e = launch_sample<V>(ds, dw, stream);
            if (e == hipSuccess) e = launch_reduce();
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(stream, &g);
            if (e == hipSuccess) e = ec;
            if (e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                HIP_TRY(e);
            }
            cycGraph = g;
            HIP_TRY(hipGraphInstantiate(&cycExec, cycGraph, nullptr, nullptr, 0));
        }
        HIP_TRY(hipGraphLaunch(cycExec, stream));
        return EKMC_OK;
    }

    ekmc_status wait_summary() {
        const auto t0 = Clock::now();
        HIP_TRY(hipStreamSynchronize(stream));
        host_add(H_WAIT, ms_since(t0));
        sync_timers();
        return EKMC_OK;
    }

    // ------------------------------------------------------- sampling (host)
    void calc_mean(Cond& C, const double* m) {
        // calculateMeanDataForSwarmParams (BoltzmannMC.C:1410-1482)
        int i = C.nSamp - 1;
        auto grow = [&](std::vector<double>& v, int w) { if ((int)v.size() < (i + 1) * w) v.resize((size_t)2 * (i + 1) * w, 0.0); };
        grow(C.sT, 1); grow(C.sE, 1); grow(C.sR, 3); grow(C.sV, 3); grow(C.sFlux, 9); grow(C.sCov, 9);
        grow(C.sBulkV, 3); grow(C.sBulkD, 9);
        C.sT[i] = C.time;
        const double n = m[25];
        C.maxElecEnergy = std::fmax(C.maxElecEnergy, m[26]);
        C.sE[i] = m[0] / n;
        for (int a = 0; a < 3; ++a) { C.sR[3 * i + a] = m[1 + a] / n; C.sV[3 * i + a] = m[4 + a] / n; }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                C.sCov[9 * i + 3 * a + b] = m[7 + 3 * a + b] / n - C.sR[3 * i + a] * C.sR[3 * i + b];
                C.sFlux[9 * i + 3 * a + b] = m[16 + 3 * a + b] / n - C.sR[3 * i + a] * C.sV[3 * i + b];
            }
        if (i != 0) {
            double dt = C.sT[i] - C.sT[i - 1];
            for (int a = 0; a < 3; ++a) C.sBulkV[3 * i + a] = (C.sR[3 * i + a] - C.sR[3 * (i - 1) + a]) / dt;
            for (int a = 0; a < 9; ++a) C.sBulkD[9 * i + a] = 0.5 * (C.sCov[9 * i + a] - C.sCov[9 * (i - 1) + a]) / dt;
        } else {
            for (int a = 0; a < 3; ++a) C.sBulkV[a] = C.sV[a];
            for (int a = 0; a < 9; ++a) C.sBulkD[a] = C.sFlux[a];
        }
        if (C.omega != 0 && C.ssTime != kNonDef) {
            int ph = phase_index(C);
            C.perN[ph] += 1.0;
            C.perE[ph] += C.sE[i];
            for (int a = 0; a < 3; ++a) { C.perV[3 * ph + a] += C.sV[3 * i + a]; C.perBulkV[3 * ph + a] += C.sBulkV[3 * i + a]; }
            for (int a = 0; a < 9; ++a) { C.perFlux[9 * ph + a] += C.sFlux[9 * i + a]; C.perBulkD[9 * ph + a] += C.sBulkD[9 * i + a]; }
        }
    }

    int phase_index(const Cond& C) const {
        double phase = std::fmod(C.omega * C.time, 2.0 * kPi);
        return (int)std::fmin(std::floor(phase / C.phaseStep), nPhases - 1);
    }

    // Download the device histogram sums of condition c into the FP64 host
    // sums (with the grid they were filled on) and zero them on the device.
    ekmc_status flush_hist(int c) {
        const auto t0 = Clock::now();
        struct Done { Engine* e; Clock::time_point t; ~Done() { e->host_add(H_FLUSH, ms_since(t)); } } done{this, t0};
        Cond& C = cs[c];
        std::vector<unsigned long long> tmp;
        auto pull = [&](DevBuf<unsigned long long>& b, size_t off, size_t n, std::vector<double>& dst) -> ekmc_status {
            tmp.resize(n);
            HIP_TRY(hipMemcpyAsync(tmp.data(), b.p + off, n * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipMemsetAsync(b.p + off, 0, n * sizeof(unsigned long long), stream));
            HIP_TRY(hipStreamSynchronize(stream));
            if (dst.size() < n) dst.assign(n, 0.0);
            for (size_t i = 0; i < n; ++i) dst[i] += (double)tmp[i];
            return EKMC_OK;
        };
        ekmc_status st;
        if ((st = pull(hEedf, (size_t)c * nCells, nCells, C.eehSum)) != EKMC_OK) return st;
        if (C.omega != 0 && hPer.p)
            if ((st = pull(hPer, (size_t)c * nPhases * nCells, (size_t)nPhases * nCells, C.eehPer)) != EKMC_OK) return st;
        if (C.cylSym && hEadf.p) {
            if ((st = pull(hEadf, (size_t)c * nCells * nCos, (size_t)nCells * nCos, C.eahSum)) != EKMC_OK) return st;
            if ((st = pull(hEvdf, (size_t)c * nRad * nAx, (size_t)nRad * nAx, C.evhSum)) != EKMC_OK) return st;
        }
        if (n3 > 0 && hEvdf3.p) {
            const size_t cube = (size_t)n3 * n3 * n3;
            if ((st = pull(hEvdf3, (size_t)c * cube, cube, C.ev3Sum)) != EKMC_OK) return st;
        }
        return EKMC_OK;
    }

    void set_eedf_grid(Cond& C) {
        C.maxEedfE = 1.2 * C.maxElecEnergy;
        C.eedfNodes = linspaced(nCells + 1, 0.0, C.maxEedfE);
        C.eedfStep = C.eedfNodes[1];
        C.eedfCells = linspaced(nCells, C.eedfStep / 2.0, C.maxEedfE - C.eedfStep / 2.0);
    }

    // getTimeDependDistributions (BoltzmannMC.C:1492-1572): re-bin when the
    // maximum energy outgrew the grid, then ask for this sample's histogram.
    ekmc_status request_histogram(int c) {
        Cond& C = cs[c];
        if (C.maxElecEnergy > C.maxEedfE) {
            ekmc_status st = flush_hist(c);
            if (st != EKMC_OK) return st;
            double oldStep = C.eedfStep;
            std::vector<double> oldNodes = C.eedfNodes, oldSum = C.eehSum, oldPer = C.eehPer, oldAh = C.eahSum;
            set_eedf_grid(C);
            std::fill(C.eehSum.begin(), C.eehSum.end(), 0.0);
            std::fill(C.eehPer.begin(), C.eehPer.end(), 0.0);
            std::fill(C.eahSum.begin(), C.eahSum.end(), 0.0);
            for (int o = 0; o < nCells; ++o) {
                double l = oldNodes[o], r = oldNodes[o + 1];
                int nl = (int)(l / C.eedfStep), nr = (int)(r / C.eedfStep);
                double lf = nl == nr ? 1.0 : (C.eedfNodes[nr] - l) / oldStep;
                auto split = [&](std::vector<double>& dst, const std::vector<double>& src, size_t dOff, size_t sOff, size_t stride) {
                    if (src.empty()) return;
                    dst[dOff + (size_t)nl * stride] += lf * src[sOff + (size_t)o * stride];
                    if (nl != nr) dst[dOff + (size_t)nr * stride] += (1.0 - lf) * src[sOff + (size_t)o * stride];
                };
                split(C.eehSum, oldSum, 0, 0, 1);
                for (int ph = 0; ph < (int)(oldPer.size() / nCells); ++ph) split(C.eehPer, oldPer, (size_t)ph * nCells, (size_t)ph * nCells, 1);
                if (!oldAh.empty())
                    for (int k = 0; k < nCos; ++k) split(C.eahSum, oldAh, k, k, nCos);
            }
        }
        C.histRequest = 1;
        C.histPhase = C.omega != 0 ? phase_index(C) : -1;
        return EKMC_OK;
    }

    void reset_at_steady_state(Cond& C) {
        // checkSteadyState, BoltzmannMC.C:1843-1892
        C.firstIntIdx = C.nSamp - 1;
        C.nIntPts = 1;
        C.ssTime = C.time;
        C.atSS = C.total;
        C.nullAtSS = C.nullc;
        C.fieldGain = 0;
        std::fill(C.counts.begin(), C.counts.end(), 0);
        std::fill(C.gainP.begin(), C.gainP.end(), 0.0);
        std::fill(C.lossP.begin(), C.lossP.end(), 0.0);
        C.growth = 0;
        C.totalIntTime = 0;
        set_eedf_grid(C);
        C.eehSum.assign(nCells, 0.0);
        C.cosStep = 0; C.radStep = 0; C.axStep = 0;
        std::vector<double> cosNodes = linspaced(nCos + 1, -1.0, 1.0);
        C.cosStep = cosNodes[1] + 1.0;
        C.cosCells = linspaced(nCos, -1.0 + C.cosStep / 2.0, 1.0 - C.cosStep / 2.0);
        C.maxSpeed = std::sqrt(2.0 * C.maxEedfE * kCharge / kMass);
        std::vector<double> rn = linspaced(nRad + 1, 0, C.maxSpeed), an = linspaced(nAx + 1, -C.maxSpeed, C.maxSpeed);
        C.radStep = rn[1];
        C.radCells = linspaced(nRad, C.radStep / 2.0, C.maxSpeed - C.radStep / 2.0);
        C.axStep = an[1] + C.maxSpeed;
        C.axCells = linspaced(nAx, -C.maxSpeed + C.axStep / 2.0, C.maxSpeed - C.axStep / 2.0);
        if (C.cylSym) { C.eahSum.assign((size_t)nCells * nCos, 0.0); C.evhSum.assign((size_t)nRad * nAx, 0.0); }
        if (C.omega != 0) C.eehPer.assign((size_t)nPhases * nCells, 0.0);
        if (n3 > 0) {
            // the 3D EVDF grid, like the 2D one, is fixed at the steady state
            C.v3Step = 2.0 * C.maxSpeed / n3;
            C.v3Cells = linspaced(n3, -C.maxSpeed + C.v3Step / 2.0, C.maxSpeed - C.v3Step / 2.0);
            C.ev3Sum.assign((size_t)n3 * n3 * n3, 0.0);
        }
    }

    void check_steady_state(Cond& C) {
        double t1 = 0.5 * C.time, t2 = 0.75 * C.time, a1 = 0, a2 = 0, rs = 0;
        int n1 = 0, n2 = 0;
        for (int i = 0; i < C.nSamp; ++i) {
            double st = C.sT[i];
            if (st >= t1 && st <= t2) { a1 += C.sE[i]; ++n1; }
            else if (st > t2) { double v = C.sE[i]; a2 += v; rs += v * v; ++n2; }
        }
        a1 /= n1; a2 /= n2;
        rs = std::sqrt((rs / n2 - a2 * a2) / n2);
        double maxBefore = num.max_collisions_before_ss * (double)C.in.n_electrons;
        if ((a1 >= a2 && rs < 0.01) || (double)C.total >= maxBefore) reset_at_steady_state(C);
    }

    void time_averages(Cond& C) {
        int f = C.firstIntIdx, n = C.nIntPts;
        C.avgE = mean_of(C.sE, 1, 0, f, n);
        C.avgEErr = stat_error(C.sE, 1, 0, f, n, 50);
        for (int a = 0; a < 3; ++a) {
            C.fluxV[a] = mean_of(C.sV, 3, a, f, n); C.fluxVErr[a] = stat_error(C.sV, 3, a, f, n, 50);
            C.bulkV[a] = mean_of(C.sBulkV, 3, a, f, n); C.bulkVErr[a] = stat_error(C.sBulkV, 3, a, f, n, 50);
        }
        for (int a = 0; a < 9; ++a) {
            C.fluxD[a] = mean_of(C.sFlux, 9, a, f, n); C.fluxDErr[a] = stat_error(C.sFlux, 9, a, f, n, 50);
            C.bulkD[a] = mean_of(C.sBulkD, 9, a, f, n); C.bulkDErr[a] = stat_error(C.sBulkD, 9, a, f, n, 50);
        }
    }

    void check_power_balance(Cond& C) {
        double g = C.fieldGain, l = 0;
        if (C.growth > 0) g += C.growth; else l += C.growth;
        for (int p = 0; p < P; ++p) { g += C.gainP[p]; l += C.lossP[p]; }
        C.powerBalanceRelError = std::abs(g + l) / g;
    }

    void check_statistical_errors(Cond& C) {
        // BoltzmannMC.C:1743-1767
        time_averages(C);
        check_power_balance(C);
        const ekmc_numerics& q = num;
        double nf = C.fluxV[0] * C.fluxV[0] + C.fluxV[1] * C.fluxV[1] + C.fluxV[2] * C.fluxV[2];
        double nb = C.bulkV[0] * C.bulkV[0] + C.bulkV[1] * C.bulkV[1] + C.bulkV[2] * C.bulkV[2];
        double ef = (std::abs(C.fluxV[0]) * C.fluxVErr[0] + std::abs(C.fluxV[1]) * C.fluxVErr[1] + std::abs(C.fluxV[2]) * C.fluxVErr[2]) / nf;
        double eb = (std::abs(C.bulkV[0]) * C.bulkVErr[0] + std::abs(C.bulkV[1]) * C.bulkVErr[1] + std::abs(C.bulkV[2]) * C.bulkVErr[2]) / nb;
        if (C.avgEErr / C.avgE <= q.rel_error_mean_energy && ef <= q.rel_error_flux_drift_velocity &&
            C.fluxDErr[0] / C.fluxD[0] <= q.rel_error_flux_diffusion_coeffs &&
            C.fluxDErr[4] / C.fluxD[4] <= q.rel_error_flux_diffusion_coeffs &&
            C.fluxDErr[8] / C.fluxD[8] <= q.rel_error_flux_diffusion_coeffs && eb <= q.rel_error_bulk_drift_velocity &&
            C.bulkDErr[0] / C.bulkD[0] <= q.rel_error_bulk_diffusion_coeffs &&
            C.bulkDErr[4] / C.bulkD[4] <= q.rel_error_bulk_diffusion_coeffs &&
            C.bulkDErr[8] / C.bulkD[8] <= q.rel_error_bulk_diffusion_coeffs &&
            C.powerBalanceRelError <= q.rel_error_power_balance)
            C.good = true;
    }

    bool keep_running(const Cond& C) const {
        // loop condition of evaluateEEDF (BoltzmannMC.C:320-325)
        const ekmc_numerics& q = num;
        bool cont = (!C.good && q.errors_to_be_checked) || C.nIntPts < q.n_integration_points ||
                    C.totalIntTime / C.ssTime < q.n_integrated_ss_times || C.totalIntTime < q.integrated_absolute_time;
        return cont;
    }

    // ------------------------------------------------------------ initialize
    ekmc_status initialize() override {
        if (initialized) return fail(EKMC_ERR_STATE, "engine already initialized");
        for (int c = 0; c < nCond; ++c) {
            Cond& C = cs[c];
            C.start = std::chrono::high_resolution_clock::now();
            C.counts.assign(P, 0); C.gainP.assign(P, 0.0); C.lossP.assign(P, 0.0);
            C.nPhases = C.omega != 0 ? nPhases : 0;
            if (C.nPhases) {
                C.phaseStep = 2.0 * kPi / nPhases;
                C.perN.assign(nPhases, 0.0); C.perE.assign(nPhases, 0.0); C.perV.assign(3 * nPhases, 0.0);
                C.perBulkV.assign(3 * nPhases, 0.0); C.perFlux.assign(9 * nPhases, 0.0); C.perBulkD.assign(9 * nPhases, 0.0);
            }
            hcp[c].trial = 1.0;
            hcp[c].trialInv = 1.0;
            hcp[c].tsync = 0;
        }
        HIP_TRY(hipMemcpyAsync(dconds.p, hcp.data(), nCond * sizeof(CondParams), hipMemcpyHostToDevice, stream));
        set_kernel_timer(&timer);
        HIP_TRY(launch_init<V>(ds, dw, totalMain, mainBase.p, thermalDev.p, stream));
        HIP_TRY(launch_sample<V>(ds, dw, stream));
        HIP_TRY(launch_reduce());
        ekmc_status st = wait_summary();
        if (st != EKMC_OK) return st;
        for (int c = 0; c < nCond; ++c) {
            Cond& C = cs[c];
            const double* m = &hsum[(size_t)c * summarySize + recSize];
            C.lastMaxE = m[26];
            C.maxElecEnergy = m[26];
            if ((st = interpolate(c, 2.0 * C.lastMaxE)) != EKMC_OK) return st;
            C.trial = C.tab.numax[I - 1];
            C.nSamp = 1; C.nSync = 1;
            C.time = 0;
            calc_mean(C, m);
            C.afterSS = 0;
            C.nIntPts = 0;
        }
        initialized = true;
        return EKMC_OK;
    }

    // ------------------------------------------------------------------ step
    ekmc_status step(int nsync, int* running) override {
        if (!initialized) return fail(EKMC_ERR_STATE, "ekmc_initialize has not been called");
        set_kernel_timer(&timer);
        for (int s = 0; s < nsync; ++s) {
            const auto tPrep = Clock::now();
            int active = 0;
            bool anyHist = false, anyHistKernel = false;
            std::vector<double> hgrid((size_t)nCond * 6, 0.0);
            std::vector<int32_t> hrow(nCond, -1);
            for (int c = 0; c < nCond; ++c) {
                Cond& C = cs[c];
                hcp[c].active = C.finished ? 0 : 1;
                double* gr = &hgrid[(size_t)c * 6];
                gr[0] = C.eedfStep; gr[1] = C.cosStep; gr[2] = C.radStep; gr[3] = C.axStep; gr[4] = C.maxSpeed;
                if (C.histRequest) {
                    // a running condition's flight kernel counts the state as it
                    // loads it (flag 2); a finished one needs the histogram kernel
                    gr[5] = (C.finished || !fusedHist) ? 1 : 2;
                    anyHist = true;
                    anyHistKernel |= C.finished || !fusedHist;
                    hrow[c] = C.histPhase;
                    C.histRequest = 0;
                }
                if (C.finished) continue;
                ++active;
                {
                    ekmc_status st = check_max_collision_frequency(c);
                    if (st != EKMC_OK) return st;
                }
                double dts = num.sync_time_x_max_coll_freq / C.trial;
                C.tsync = C.time + dts;
                ++C.nSync;
                ++C.syncIndex;
                hcp[c].trial = C.trial;
                hcp[c].trialInv = 1.0 / C.trial;
                hcp[c].tsync = C.tsync;
                hcp[c].step = C.tab.step;
                hcp[c].syncIndex = C.syncIndex;
            }
            if (active == 0 && !anyHist) break;
            {
                // histograms of the state sampled at the end of the previous interval
                ekmc_status st = upload_hist_grid(hgrid, hrow);
                if (st != EKMC_OK) return st;
                if (anyHistKernel) HIP_TRY(launch_hist<V>(ds, dh, stream));
            }
            if (active == 0) { HIP_TRY(hipStreamSynchronize(stream)); break; }
            {
                ekmc_status st = rebuild_lists();
                if (st != EKMC_OK) return st;
            }
            host_add(H_PREPARE, ms_since(tPrep));
            const auto tLaunch = Clock::now();
            std::memcpy(pinCond.p, hcp.data(), nCond * sizeof(CondParams));
            HIP_TRY(hipMemcpyAsync(dconds.p, pinCond.p, nCond * sizeof(CondParams), hipMemcpyHostToDevice, stream));
            HIP_TRY(launch_flights<V>(ds, dw, dpr, tables.p, dh, fl, stream));
            if (snap && !snap->taken && !cs[snap->cond].finished) {
                ekmc_status st = take_snapshot(*snap);
                if (st != EKMC_OK) return st;
            }
            ekmc_status st = launch_cycle();
            if (st != EKMC_OK) return st;
            host_add(H_LAUNCH, ms_since(tLaunch));
            if ((st = wait_summary()) != EKMC_OK) return st;
            const auto tPost = Clock::now();
            for (int c = 0; c < nCond; ++c) {
                if (cs[c].finished) continue;
                st = post_interval(c);
                if (st != EKMC_OK) return st;
            }
            host_add(H_POST, ms_since(tPost));
        }
        int run = 0;
        for (auto& C : cs) run += C.finished ? 0 : 1;
        if (running) *running = run;
        return EKMC_OK;
    }

    // Every alive electron of condition c's slots (main slots, then the
    // newborn regions of its chunks), in slot order.
    ekmc_status take_snapshot(PoolSnap& sn) {
        const int c = sn.cond;
        const int64_t begin = hcp[c].base, end = c + 1 < nCond ? hcp[c + 1].base : capacity;
        const size_t n = (size_t)(end - begin);
        std::vector<Rec<V>> h(n);
        std::vector<RecA<V>> ha(n);
        std::vector<RecB<V>> hb(split_record<V>() ? n : 0);
        std::vector<uint8_t> al(n);
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipMemcpy(ha.data(), recA.p + begin, n * sizeof(RecA<V>), hipMemcpyDeviceToHost));
        if (split_record<V>())
            HIP_TRY(hipMemcpy(hb.data(), recB.p + begin, n * sizeof(RecB<V>), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            h[i].a = ha[i];
            if constexpr (split_record<V>()) h[i].b = hb[i];
        }
        HIP_TRY(hipMemcpy(al.data(), alive.p + begin, n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            if (!al[i]) continue;
            const RecA<V>& q = ha[i];
            sn.pos.insert(sn.pos.end(), {q.x, q.y, q.z});
            sn.vel.insert(sn.vel.end(), {(double)q.vx, (double)q.vy, (double)q.vz});
            sn.energy.push_back(host_energy(q.vx, q.vy, q.vz));
            sn.time.push_back(q.t);
            sn.stream.push_back(draw_part(h[i]).stream);
            sn.event.push_back(draw_part(h[i]).event);
        }
        sn.tsync = cs[c].tsync;
        sn.trial = cs[c].trial;
        sn.taken = true;
        return EKMC_OK;
    }

    ekmc_status post_interval(int c) {
        Cond& C = cs[c];
        const double* ev = &hsum[(size_t)c * summarySize];
        const double* m = ev + recSize;
        const double* tail = m + kMomentRecord;
        uint32_t stbits = (uint32_t)tail[6];
        if (stbits & ST_ZERO_VELOCITY) return fail(EKMC_ERR_ZERO_VELOCITY, "Trying to use the function 'cart2sph' for a vector with null norm!");
        if (stbits & ST_REGION_OVERFLOW) return fail(EKMC_ERR_CAPACITY, "newborn region of a thread overflowed");
        if (stbits & ST_POPCTL) return fail(EKMC_ERR_CAPACITY, "population control could not draw an exact sample");
        if (stbits & ST_LOOP_CAP) return fail(EKMC_ERR_CAPACITY, "flight cap reached within one interval");
        if (stbits & ST_NONFINITE) return fail(EKMC_ERR_STATE, "an electron time or energy became non-finite");
        // nonParallelCollisionTasks bookkeeping (BoltzmannMC.C:1291-1339);
        // the counts arrive as exact non-negative integers in doubles
        for (int p = 0; p < P; ++p) {
            C.counts[p] += (uint64_t)ev[p];
            C.gainP[p] += ev[P + p];
            C.lossP[p] += ev[2 * P + p];
        }
        C.fieldGain += ev[3 * P + EV_FIELD];
        C.growth += ev[3 * P + EV_GROWTH];
        uint64_t real = (uint64_t)ev[3 * P + EV_REAL], nul = (uint64_t)ev[3 * P + EV_NULL];
        C.total += real;
        C.nullc += nul;
        C.events += (uint64_t)ev[3 * P + EV_EVENTS];
        C.workPerElectron = C.in.n_electrons > 0 ? ev[3 * P + EV_EVENTS] / (double)C.in.n_electrons : 1.0;
        C.escapes += (uint64_t)ev[3 * P + EV_ESCAPE];
        totalEvents += real + nul;
        totalReal += real;
        C.lastMaxE = m[26];
        C.time = C.tsync;
        if (C.nSync % num.sync_over_sampling == 0) {
            ++C.nSamp;
            calc_mean(C, m);
            if (C.maxElecEnergy > pl.energyMaxElastic) {
                char buf[256];
                std::snprintf(buf, sizeof buf, "The energy of the electrons reached %f eV, while at least one of the elastic cross sections are defined only until %f eV.", C.maxElecEnergy, pl.energyMaxElastic);
                return fail(EKMC_ERR_ENERGY_OUT_OF_TABLE, buf);
            }
            // the reference's cadence expressions, evaluated only where they
            // can exceed their floor (below 2^17 samples dec is 6)
            int dec = C.nSamp < (1 << 17) ? 6 : (int)std::fmax(std::log(C.nSamp) / std::log(2.0) - 11, 6);
            int nSS = 1 << dec;
            if (C.ssTime != kNonDef) {
                ++C.nIntPts;
                C.afterSS = C.total - C.atSS;
                C.timeIntervalInteg = C.time - C.totalIntTime;
                C.totalIntTime = C.time - C.ssTime;
                ekmc_status st = request_histogram(c);
                if (st != EKMC_OK) return st;
            } else if (C.nSamp >= 100 && C.nSamp % nSS == 0 &&
                       (double)C.total > num.min_collisions_before_ss * (double)C.in.n_electrons) {
                check_steady_state(C);
                if (C.ssTime != kNonDef) {
                    C.histRequest = 1;
                    C.histPhase = C.omega != 0 ? phase_index(C) : -1;
                }
            }
            dec = C.nIntPts < (1 << 16) ? 7 : (int)std::fmax(std::log(C.nIntPts) / std::log(2.0) - 9, 7);
            int nErr = 1 << dec;
            if (C.ssTime != kNonDef && C.nIntPts > 200 && C.nIntPts % nErr == 0) check_statistical_errors(C);
        }
        bool stop = !keep_running(C);
        if (!stop && (double)C.afterSS >= num.max_collisions_after_ss * (double)C.in.n_electrons && C.nIntPts > 100) {
            stop = true;
            C.stoppedByMax = true;
        }
        if (!stop && num.max_sync_points > 0 && C.nSync >= num.max_sync_points) stop = true;
        if (stop) {
            C.finished = true;
            C.elapsed = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - C.start).count();
        }
        return EKMC_OK;
    }

    // -------------------------------------------------------------- finalize
    ekmc_status finalize() override {
        if (!initialized) return fail(EKMC_ERR_STATE, "ekmc_initialize has not been called");
        // pending histograms of the last samples
        bool any = false;
        std::vector<double> hgrid((size_t)nCond * 6, 0.0);
        std::vector<int32_t> hrow(nCond, -1);
        for (int c = 0; c < nCond; ++c) {
            Cond& C = cs[c];
            double* gr = &hgrid[(size_t)c * 6];
            gr[0] = C.eedfStep; gr[1] = C.cosStep; gr[2] = C.radStep; gr[3] = C.axStep; gr[4] = C.maxSpeed;
            if (C.histRequest) { gr[5] = 1; any = true; hrow[c] = C.histPhase; C.histRequest = 0; }
        }
        {
            ekmc_status st = upload_hist_grid(hgrid, hrow);
            if (st != EKMC_OK) return st;
            if (any) HIP_TRY(launch_hist<V>(ds, dh, stream));
        }
        HIP_TRY(hipStreamSynchronize(stream));
        sync_timers();
        for (int c = 0; c < nCond; ++c) {
            Cond& C = cs[c];
            if (C.ssTime == kNonDef || C.nIntPts < 50)
                return fail(EKMC_ERR_STATISTICS, "Error in 'statisticalError'. nBins is larger than nElements.");
            ekmc_status st = flush_hist(c);
            if (st != EKMC_OK) return st;
            C.totalIntTime = C.time - C.ssTime;
            time_averages(C);
            check_statistical_errors(C);
            if (!C.finished) {
                C.finished = true;
                C.elapsed = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - C.start).count();
            }
        }
        finalized = true;
        return EKMC_OK;
    }

    ekmc_status get_electrons(int c, double* pos, double* vel, double* e, double* tt) override {
        const CondParams& cp = hcp[c];
        const int64_t n = cp.nMain;
        std::vector<RecA<V>> h(n);
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipMemcpy(h.data(), recA.p + cp.base, n * sizeof(RecA<V>), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) {
            const RecA<V>& q = h[i];
            if (pos) { pos[3 * i] = q.x; pos[3 * i + 1] = q.y; pos[3 * i + 2] = q.z; }
            if (vel) { vel[3 * i] = q.vx; vel[3 * i + 1] = q.vy; vel[3 * i + 2] = q.vz; }
            if (e) e[i] = host_energy(q.vx, q.vy, q.vz);
            if (tt) tt[i] = q.t;
        }
        return EKMC_OK;
    }

    ekmc_status op_accelerate(int c, int64_t n, const double* tm, const double* dt, double* pos, double* vel,
                              double* e, double* g) override {
        DevBuf<double> b;
        HIP_TRY(b.alloc((size_t)n * 10));
        double *dT = b.p, *dDt = b.p + n, *dP = b.p + 2 * n, *dV = b.p + 5 * n, *dE = b.p + 8 * n, *dG = b.p + 9 * n;
        HIP_TRY(hipMemcpyAsync(dconds.p + c, &hcp[c], sizeof(CondParams), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dT, tm, n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dDt, dt, n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dP, pos, 3 * n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dV, vel, 3 * n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(launch_op_accelerate<V>(dconds.p + c, n, dT, dDt, dP, dV, dE, dG, stream));
        HIP_TRY(hipMemcpyAsync(pos, dP, 3 * n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(vel, dV, 3 * n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(e, dE, n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(g, dG, n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        b.release();
        return EKMC_OK;
    }

    ekmc_status op_tables(int c, double maxE, double* interp, double* cumul, double* tot, double* mx,
                          double* step) override {
        if (!(maxE > 0)) return fail(EKMC_ERR_INVALID_ARGUMENT, "max_energy must be positive");
        ekmc_status st = interpolate(c, maxE, true);
        if (st != EKMC_OK) return st;
        const size_t n = (size_t)I * P;
        std::vector<V> h(2 * n + I);
        HIP_TRY(hipMemcpy(h.data(), tables.p + hcp[c].tableOffset, h.size() * sizeof(V), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            if (cumul) cumul[i] = (double)h[i];
            if (interp) interp[i] = (double)h[n + i];
        }
        for (int i = 0; i < I; ++i) {
            if (tot) tot[i] = (double)h[2 * n + i];
            if (mx) mx[i] = cs[c].tab.numax[i];
        }
        if (step) *step = cs[c].tab.step;
        return EKMC_OK;
    }

    ekmc_status op_collide(int c, double maxE, int64_t n, double* vel, double* e, const double* nuev,
                           const double* var, int32_t* pid, double* ejv, double* eje, double* de) override {
        if (maxE > 0) {
            ekmc_status st = interpolate(c, maxE);
            if (st != EKMC_OK) return st;
        } else if (cs[c].tab.maxEnergy <= 0) {
            return fail(EKMC_ERR_STATE, "no tables built for this condition; pass max_energy > 0");
        }
        DevBuf<unsigned char> b;
        HIP_TRY(b.alloc((size_t)n * kCollideStageBytes));
        ekmc_status st = collide_on(c, n, vel, e, nuev, var, pid, ejv, eje, de, b.p);
        b.release();
        return st;
    }

    ekmc_status op_collide_staged(int c, double maxE, int64_t n, double* vel, double* e, const double* nuev,
                                  const double* var, int32_t* pid, double* ejv, double* eje, double* de,
                                  void* stage) override {
        if (maxE > 0) {
            ekmc_status st = interpolate(c, maxE);
            if (st != EKMC_OK) return st;
        } else if (cs[c].tab.maxEnergy <= 0) {
            return fail(EKMC_ERR_STATE, "no tables built for this condition; pass max_energy > 0");
        }
        return collide_on(c, n, vel, e, nuev, var, pid, ejv, eje, de, stage);
    }

    // Stage the inputs at `stage` (n * kCollideStageBytes bytes of device
    // memory: 21 FP64 values and one int32 per event, every one written by
    // the copies or the kernel), collide, copy the outputs back.
    ekmc_status collide_on(int c, int64_t n, double* vel, double* e, const double* nuev, const double* var,
                           int32_t* pid, double* ejv, double* eje, double* de, void* stage) {
        CondParams cp = hcp[c];
        cp.step = cs[c].tab.step;
        double *dV = (double*)stage, *dE = dV + 3 * n, *dN = dE + n, *dVar = dN + n, *dEjv = dVar + 11 * n,
               *dEje = dEjv + 3 * n, *dDe = dEje + n;
        int32_t* dpp = (int32_t*)(dDe + n);
        HIP_TRY(hipMemcpyAsync(dconds.p + c, &cp, sizeof(CondParams), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dV, vel, 3 * n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dE, e, n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dN, nuev, n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(dVar, var, 11 * n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(launch_op_collide<V>(dconds.p + c, dpr, tables.p, n, dV, dE, dN, dVar, dpp, dEjv, dEje, dDe, stream));
        HIP_TRY(hipMemcpyAsync(vel, dV, 3 * n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(e, dE, n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(pid, dpp, n * 4, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(ejv, dEjv, 3 * n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(eje, dEje, n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipMemcpyAsync(de, dDe, n * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return EKMC_OK;
    }
};

void fill_results(const EngineBase& E, int c, ekmc_results* r) {
    const Cond& C = E.cs[c];
    std::memset(r, 0, sizeof(*r));
    r->mean_energy = C.avgE; r->mean_energy_error = C.avgEErr;
    for (int a = 0; a < 3; ++a) {
        r->flux_drift_velocity[a] = C.fluxV[a]; r->flux_drift_velocity_error[a] = C.fluxVErr[a];
        r->bulk_drift_velocity[a] = C.bulkV[a]; r->bulk_drift_velocity_error[a] = C.bulkVErr[a];
    }
    for (int a = 0; a < 9; ++a) {
        r->flux_diffusion[a] = C.fluxD[a]; r->flux_diffusion_error[a] = C.fluxDErr[a];
        r->bulk_diffusion[a] = C.bulkD[a]; r->bulk_diffusion_error[a] = C.bulkDErr[a];
    }
    // rotation about y by -angle into the frame where E is along z
    // (evaluateSwarmParameters, BoltzmannMC.C:2064-2140)
    double ra = -C.angle / 180.0 * kPi;
    double R[9] = {std::cos(ra), 0, std::sin(ra), 0, 1, 0, -std::sin(ra), 0, std::cos(ra)};
    double A[9];
    for (int k = 0; k < 9; ++k) A[k] = std::abs(R[k]);
    auto rotv = [&](const double* M, const double* v, double* out) {
        for (int a = 0; a < 3; ++a) out[a] = M[3 * a] * v[0] + M[3 * a + 1] * v[1] + M[3 * a + 2] * v[2];
    };
    auto rotm = [&](const double* M, const double* D, double* out) {
        double tmp[9];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += M[3 * a + k] * D[3 * k + b];
                tmp[3 * a + b] = s;
            }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += tmp[3 * a + k] * M[3 * b + k];
                out[3 * a + b] = s;
            }
    };
    rotv(R, C.fluxV, r->rotated_flux_drift_velocity); rotv(A, C.fluxVErr, r->rotated_flux_drift_velocity_error);
    rotv(R, C.bulkV, r->rotated_bulk_drift_velocity); rotv(A, C.bulkVErr, r->rotated_bulk_drift_velocity_error);
    rotm(R, C.fluxD, r->rotated_flux_diffusion); rotm(A, C.fluxDErr, r->rotated_flux_diffusion_error);
    rotm(R, C.bulkD, r->rotated_bulk_diffusion); rotm(A, C.bulkDErr, r->rotated_bulk_diffusion_error);
    const double N = C.N, NE = (double)C.in.n_electrons, T = C.totalIntTime;
    r->power_field = C.fieldGain / (NE * T * N);
    r->power_growth = C.growth / (NE * T * N);
    r->power_balance_rel_error = C.powerBalanceRelError;
    double ion = 0, att = 0;
    for (int p = 0; p < E.pl.P; ++p) {
        double k = E.pl.relDens[p] != 0 ? (double)C.counts[p] / (T * NE * E.pl.relDens[p] * N) : 0;
        if (E.pl.type[p] == EKMC_IONIZATION) ion += E.pl.relDens[p] * k;
        if (E.pl.type[p] == EKMC_ATTACHMENT) att += E.pl.relDens[p] * k;
    }
    r->total_ion_rate_coeff = ion;
    r->total_att_rate_coeff = att;
    double ENSI = C.in.reduced_elec_field * 1e-21;
    if (C.omega == 0 && ENSI > 0) {
        r->flux_red_mobility = std::abs(r->rotated_flux_drift_velocity[2]) / ENSI;
        r->bulk_red_mobility = std::abs(r->rotated_bulk_drift_velocity[2]) / ENSI;
    }
    r->flux_red_transv_diff = N * (r->rotated_flux_diffusion[0] + r->rotated_flux_diffusion[4]) / 2.0;
    r->flux_red_long_diff = N * r->rotated_flux_diffusion[8];
    r->bulk_red_transv_diff = N * (r->rotated_bulk_diffusion[0] + r->rotated_bulk_diffusion[4]) / 2.0;
    r->bulk_red_long_diff = N * r->rotated_bulk_diffusion[8];
    r->steady_state_time = C.ssTime;
    r->final_time = C.time;
    r->total_integrated_time = C.totalIntTime;
    r->trial_collision_frequency = C.trial;
    r->max_elec_energy = C.maxElecEnergy;
    r->max_eedf_energy = C.maxEedfE;
    r->elapsed_time = C.elapsed;
    r->real_collisions = C.total;
    r->null_collisions = C.nullc;
    r->real_collisions_at_ss = C.atSS;
    r->null_collisions_at_ss = C.nullAtSS;
    r->escape_events = C.escapes;
    r->n_sampling_points = C.nSamp;
    r->n_integration_points = C.nIntPts;
    r->good_statistical_errors = C.good;
    r->stopped_by_max_collisions = C.stoppedByMax;
}

}  // namespace
}  // namespace ekmc

// ============================================= allocation-free operator
// Staging of ekmc_op_collide_noalloc: one static device array per GPU, held
// until the process ends.  In host code `scratch` only names the symbol; its
// device address on the current GPU comes from hipGetSymbolAddress, once per
// device, and only that address is handed to copies and kernels.
namespace ekmc {
namespace {
constexpr size_t kScratchBytes = (size_t)EKMC_NOALLOC_MAX_EVENTS * kCollideStageBytes;
static_assert(kScratchBytes <= ((size_t)8 << 20), "the static staging stays modest");
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void* g_scratchPtr[kMaxDevices] = {};

// Device address of `scratch` on device `dev`, which must be current.
ekmc_status scratch_address(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(EKMC_ERR_INVALID_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(EKMC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return EKMC_OK;
}
}  // namespace
}  // namespace ekmc

// ================================================================== C ABI
using namespace ekmc;

extern "C" {

void ekmc_numerics_default(ekmc_numerics* n) {
    std::memset(n, 0, sizeof(*n));
    n->precision = EKMC_PRECISION_FP64;
    n->gas_temperature_effect = EKMC_GASTEMP_FALSE;
    n->ionization_operator = EKMC_IONIZ_EQUAL_SHARING;
    n->n_interp_points = 10000;
    n->initial_elec_temp_over_gas_temp = 0.01;
    n->min_collisions_before_ss = 0;
    n->max_collisions_before_ss = 1e100;
    n->max_collisions_after_ss = 1e100;
    n->n_energy_cells = 1000;
    n->n_cos_angle_cells = 100;
    n->n_radial_velocity_cells = 200;
    n->n_axial_velocity_cells = 200;
    n->errors_to_be_checked = 0;
    n->n_integration_phases = 100;
    n->rel_error_mean_energy = n->rel_error_flux_drift_velocity = n->rel_error_bulk_drift_velocity = 1e100;
    n->rel_error_flux_diffusion_coeffs = n->rel_error_bulk_diffusion_coeffs = n->rel_error_power_balance = 1e100;
    n->n_integration_points = 200;
    n->n_integrated_ss_times = 0;
    n->integrated_absolute_time = 0;
    n->sync_time_x_max_coll_freq = 1;
    n->sync_over_sampling = 1;
    n->max_sync_points = 0;
    n->seed = 0;
    n->work_chunk = 0;
    n->park_lanes = 0;
    n->env_coarsen = 0;
    n->evdf3d_cells = 0;
}

ekmc_status ekmc_create(const ekmc_mixture* mixture, const ekmc_condition* conditions, int32_t n_cond,
                        const ekmc_numerics* numerics, ekmc_engine** out) {
    if (!mixture || !conditions || !numerics || !out || n_cond <= 0)
        return fail(EKMC_ERR_INVALID_ARGUMENT, "null argument or no conditions");
    *out = nullptr;
    std::unique_ptr<ekmc_engine> e(new ekmc_engine());
    if (numerics->precision == EKMC_PRECISION_FP64) e->impl.reset(new Engine<double>());
    else if (numerics->precision == EKMC_PRECISION_MIXED) e->impl.reset(new Engine<float>());
    else return fail(EKMC_ERR_INVALID_ARGUMENT, "unknown precision");
    ekmc_status st = e->impl->create(*mixture, conditions, n_cond, *numerics);
    if (st != EKMC_OK) return st;
    *out = e.release();
    return EKMC_OK;
}

void ekmc_destroy(ekmc_engine* engine) { delete engine; }

// One autotuner trial: a fresh engine of the batch with `nm`, `warmup`
// untimed and `steps` timed intervals; *rate = collision events per second.
static ekmc_status autotune_trial(const ekmc_mixture* mixture, const ekmc_condition* conditions, int32_t n_cond,
                                  ekmc_numerics nm, int32_t warmup, int32_t steps, double* rate) {
    // no condition may stop inside the trial
    nm.n_integration_points = 1e300;
    nm.max_collisions_after_ss = 1e300;
    nm.max_sync_points = 0;
    ekmc_engine* raw = nullptr;
    ekmc_status st = ekmc_create(mixture, conditions, n_cond, &nm, &raw);
    if (st != EKMC_OK) return st;
    std::unique_ptr<ekmc_engine> e(raw);
    int running = 0;
    if ((st = e->impl->initialize()) != EKMC_OK) return st;
    if (warmup > 0 && (st = e->impl->step(warmup, &running)) != EKMC_OK) return st;
    const uint64_t ev0 = e->impl->totalEvents;
    const auto t0 = Clock::now();
    if ((st = e->impl->step(steps, &running)) != EKMC_OK) return st;
    const double seconds = ms_since(t0) * 1e-3;
    *rate = seconds > 0 ? (double)(e->impl->totalEvents - ev0) / seconds : 0.0;
    return EKMC_OK;
}

ekmc_status ekmc_autotune_launch(const ekmc_mixture* mixture, const ekmc_condition* conditions, int32_t n_cond,
                                 const ekmc_numerics* numerics, int32_t warmup, int32_t steps, ekmc_numerics* best,
                                 double* best_rate) {
    if (!mixture || !conditions || !numerics || !best || n_cond <= 0)
        return fail(EKMC_ERR_INVALID_ARGUMENT, "null argument or no conditions");
    if (warmup < 0 || steps < 1) return fail(EKMC_ERR_INVALID_ARGUMENT, "autotune needs warmup >= 0, steps >= 1");
    ekmc_numerics cur = *numerics;
    cur.park_lanes = 0;
    cur.env_coarsen = 0;
    double rate = 0;
    ekmc_status st = ekmc_autotune(mixture, conditions, n_cond, &cur, nullptr, 0, warmup, steps, &cur.work_chunk,
                                   nullptr);
    if (st != EKMC_OK) return st;
    if ((st = autotune_trial(mixture, conditions, n_cond, cur, warmup, steps, &rate)) != EKMC_OK) return st;
    // coordinate descent: the lane threshold at the tuned work unit, then the
    // envelope coarseness at both
    static const int32_t kPark[] = {16, 24, 40, 48};
    ekmc_numerics trial = cur;
    for (int32_t p : kPark) {
        trial.park_lanes = p;
        double r = 0;
        if ((st = autotune_trial(mixture, conditions, n_cond, trial, warmup, steps, &r)) != EKMC_OK) return st;
        if (r > rate) { rate = r; cur.park_lanes = p; }
    }
    trial = cur;
    for (int32_t ec = 1; ec <= 2; ++ec) {
        trial.env_coarsen = ec;
        double r = 0;
        if ((st = autotune_trial(mixture, conditions, n_cond, trial, warmup, steps, &r)) != EKMC_OK) return st;
        if (r > rate) { rate = r; cur.env_coarsen = ec; }
    }
    if (cur.park_lanes == 0) cur.park_lanes = 32;
    *best = cur;
    if (best_rate) *best_rate = rate;
    return EKMC_OK;
}

ekmc_status ekmc_autotune(const ekmc_mixture* mixture, const ekmc_condition* conditions, int32_t n_cond,
                          const ekmc_numerics* numerics, const int32_t* candidates, int32_t n_candidates,
                          int32_t warmup, int32_t steps, int32_t* best_work_chunk, double* rates) {
    static const int32_t kCandidates[] = {256, 512, 768, 1024, 1536, 2048};
    if (!mixture || !conditions || !numerics || !best_work_chunk || n_cond <= 0)
        return fail(EKMC_ERR_INVALID_ARGUMENT, "null argument or no conditions");
    if (warmup < 0 || steps < 1 || n_candidates < 0)
        return fail(EKMC_ERR_INVALID_ARGUMENT, "autotune needs warmup >= 0, steps >= 1, n_candidates >= 0");
    if (!candidates || n_candidates == 0) {
        candidates = kCandidates;
        n_candidates = (int32_t)(sizeof(kCandidates) / sizeof(kCandidates[0]));
    }
    for (int32_t i = 0; i < n_candidates; ++i)
        if (candidates[i] <= 0) return fail(EKMC_ERR_INVALID_ARGUMENT, "work-unit candidates must be positive");
    double best = -1;
    int32_t bestChunk = 0;
    for (int32_t i = 0; i < n_candidates; ++i) {
        ekmc_numerics nm = *numerics;
        nm.work_chunk = candidates[i];
        double rate = 0;
        ekmc_status st = autotune_trial(mixture, conditions, n_cond, nm, warmup, steps, &rate);
        if (st != EKMC_OK) return st;
        if (rates) rates[i] = rate;
        if (rate > best) {
            best = rate;
            bestChunk = candidates[i];
        }
    }
    *best_work_chunk = bestChunk;
    return EKMC_OK;
}

ekmc_status ekmc_initialize(ekmc_engine* e) {
    if (!e) return fail(EKMC_ERR_INVALID_ARGUMENT, "null engine");
    return e->impl->initialize();
}

ekmc_status ekmc_step(ekmc_engine* e, int32_t n_sync, int32_t* n_running) {
    if (!e) return fail(EKMC_ERR_INVALID_ARGUMENT, "null engine");
    return e->impl->step(n_sync, n_running);
}

ekmc_status ekmc_finalize(ekmc_engine* e) {
    if (!e) return fail(EKMC_ERR_INVALID_ARGUMENT, "null engine");
    return e->impl->finalize();
}

ekmc_status ekmc_solve(ekmc_engine* e) {
    if (!e) return fail(EKMC_ERR_INVALID_ARGUMENT, "null engine");
    ekmc_status st = e->impl->initialize();
    if (st != EKMC_OK) return st;
    int running = 1;
    while (running > 0) {
        st = e->impl->step(64, &running);
        if (st != EKMC_OK) return st;
    }
    return e->impl->finalize();
}

ekmc_status ekmc_get_results(const ekmc_engine* e, int32_t cond, ekmc_results* out) {
    if (!e || !out || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    if (!e->impl->finalized) return fail(EKMC_ERR_STATE, "results are available after ekmc_finalize/ekmc_solve");
    fill_results(*e->impl, cond, out);
    return EKMC_OK;
}

ekmc_status ekmc_get_process_data(const ekmc_engine* e, int32_t cond, double* rates, uint64_t* counts, double* gain,
                                  double* loss) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const EngineBase& E = *e->impl;
    const Cond& C = E.cs[cond];
    const double N = C.N, NE = (double)C.in.n_electrons, T = C.totalIntTime;
    for (int p = 0; p < E.pl.P; ++p) {
        if (rates) rates[p] = E.pl.relDens[p] != 0 && T > 0 ? (double)C.counts[p] / (T * NE * E.pl.relDens[p] * N) : 0;
        if (counts) counts[p] = C.counts[p];
        if (gain) gain[p] = T > 0 ? C.gainP[p] / (NE * T * N) : 0;
        if (loss) loss[p] = T > 0 ? C.lossP[p] / (NE * T * N) : 0;
    }
    return EKMC_OK;
}

ekmc_status ekmc_get_eedf(const ekmc_engine* e, int32_t cond, double* cells, double* eedf) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    if (!e->impl->finalized) return fail(EKMC_ERR_STATE, "the eedf is available after ekmc_finalize");
    const Cond& C = e->impl->cs[cond];
    double norm = 0;
    for (double v : C.eehSum) norm += v;
    for (size_t i = 0; i < C.eedfCells.size(); ++i) {
        if (cells) cells[i] = C.eedfCells[i];
        if (eedf) eedf[i] = C.eehSum[i] / (std::sqrt(C.eedfCells[i]) * norm * C.eedfStep);
    }
    return EKMC_OK;
}

ekmc_status ekmc_get_eedf_periodic(const ekmc_engine* e, int32_t cond, double* per, double* meanE, double* fluxV) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const Cond& C = e->impl->cs[cond];
    if (C.omega == 0) return fail(EKMC_ERR_STATE, "condition has no AC field");
    const int nc = (int)C.eedfCells.size(), np = C.nPhases;
    for (int ph = 0; ph < np; ++ph) {
        double norm = 0;
        for (int i = 0; i < nc; ++i) norm += C.eehPer[(size_t)ph * nc + i];
        for (int i = 0; i < nc; ++i)
            if (per) per[(size_t)ph * nc + i] = C.eehPer[(size_t)ph * nc + i] / (std::sqrt(C.eedfCells[i]) * norm * C.eedfStep);
        if (meanE) meanE[ph] = C.perE[ph] / C.perN[ph];
        if (fluxV) for (int a = 0; a < 3; ++a) fluxV[3 * ph + a] = C.perV[3 * ph + a] / C.perN[ph];
    }
    return EKMC_OK;
}

ekmc_status ekmc_get_swarm_periodic(const ekmc_engine* e, int32_t cond, double* phases, double* meanE, double* fluxV,
                                    double* bulkV, double* fluxD, double* bulkD) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const Cond& C = e->impl->cs[cond];
    if (C.omega == 0) return fail(EKMC_ERR_STATE, "condition has no AC field");
    const int np = C.nPhases;
    // bin centres: np points from step/2 to 2 pi - step/2, the last exact
    const double lo = 0.5 * C.phaseStep, hi = 2.0 * kPi - 0.5 * C.phaseStep;
    const double step = np > 1 ? (hi - lo) / (double)(np - 1) : 0.0;
    for (int ph = 0; ph < np; ++ph) {
        const double n = C.perN[ph];
        if (phases) phases[ph] = ph == np - 1 ? hi : lo + (double)ph * step;
        if (meanE) meanE[ph] = C.perE[ph] / n;
        for (int a = 0; a < 3; ++a) {
            if (fluxV) fluxV[3 * ph + a] = C.perV[3 * ph + a] / n;
            if (bulkV) bulkV[3 * ph + a] = C.perBulkV[3 * ph + a] / n;
        }
        for (int a = 0; a < 9; ++a) {
            if (fluxD) fluxD[9 * ph + a] = C.perFlux[9 * ph + a] / n;
            if (bulkD) bulkD[9 * ph + a] = C.perBulkD[9 * ph + a] / n;
        }
    }
    return EKMC_OK;
}

ekmc_status ekmc_get_evdf3d(const ekmc_engine* e, int32_t cond, double* axis, double* evdf3d) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const int n3 = e->impl->num.evdf3d_cells;
    if (n3 == 0) return fail(EKMC_ERR_STATE, "the 3D EVDF was not requested (evdf3d_cells is 0)");
    if (!e->impl->finalized) return fail(EKMC_ERR_STATE, "distributions are available after ekmc_finalize");
    const Cond& C = e->impl->cs[cond];
    if (C.v3Cells.size() != (size_t)n3 || C.ev3Sum.size() != (size_t)n3 * n3 * n3)
        return fail(EKMC_ERR_STATE, "the condition has no sample after its steady state");
    double tot = 0;
    for (double v : C.ev3Sum) tot += v;
    const double vol = C.v3Step * C.v3Step * C.v3Step;
    if (axis) for (int i = 0; i < n3; ++i) axis[i] = C.v3Cells[i];
    if (evdf3d)
        for (size_t i = 0; i < C.ev3Sum.size(); ++i) evdf3d[i] = tot > 0 ? C.ev3Sum[i] / (tot * vol) : 0.0;
    return EKMC_OK;
}

ekmc_status ekmc_get_evdf(const ekmc_engine* e, int32_t cond, double* radial, double* axial, double* evdf, double* efadf,
                          double* esadf) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const Cond& C = e->impl->cs[cond];
    if (!C.cylSym) return fail(EKMC_ERR_STATE, "condition is not cylindrically symmetric");
    if (!e->impl->finalized) return fail(EKMC_ERR_STATE, "distributions are available after ekmc_finalize");
    const int nR = (int)C.radCells.size(), nA = (int)C.axCells.size(), nc = (int)C.eedfCells.size();
    const int nCos = (int)C.cosCells.size();
    // getTimeAverageDistributions (BoltzmannMC.C:1574-1605)
    double denom = 0;
    for (int r = 0; r < nR; ++r) {
        double row = 0;
        for (int a = 0; a < nA; ++a) row += C.evhSum[(size_t)r * nA + a];
        denom += row * C.axStep * 2.0 * kPi * C.radCells[r] * C.radStep;
    }
    for (int r = 0; r < nR; ++r) {
        if (radial) radial[r] = C.radCells[r];
        for (int a = 0; a < nA; ++a)
            if (evdf) evdf[(size_t)r * nA + a] = C.evhSum[(size_t)r * nA + a] / denom;
    }
    if (axial) for (int a = 0; a < nA; ++a) axial[a] = C.axCells[a];
    double norm = 0;
    for (double v : C.eehSum) norm += v;
    for (int i = 0; i < nc; ++i) {
        double f1 = 0, f2 = 0;
        for (int k = 0; k < nCos; ++k) {
            double ead = 2.0 * C.eahSum[(size_t)i * nCos + k] / (std::sqrt(C.eedfCells[i]) * norm * C.eedfStep * C.cosStep);
            f1 += ead * C.cosCells[k];
            f2 += ead * 0.5 * (3.0 * C.cosCells[k] * C.cosCells[k] - 1);
        }
        if (efadf) efadf[i] = 1.5 * f1 * C.cosStep;
        if (esadf) esadf[i] = 2.5 * f2 * C.cosStep;
    }
    return EKMC_OK;
}

ekmc_status ekmc_get_series(const ekmc_engine* e, int32_t cond, int32_t* n_points, double* times, double* meanE,
                            double* meanR, double* meanV, double* flux, double* bulkV, double* bulkD) {
    if (!e || !n_points || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const Cond& C = e->impl->cs[cond];
    int n = std::min(*n_points, C.nSamp);
    for (int i = 0; i < n; ++i) {
        if (times) times[i] = C.sT[i];
        if (meanE) meanE[i] = C.sE[i];
        for (int a = 0; a < 3; ++a) {
            if (meanR) meanR[3 * i + a] = C.sR[3 * i + a];
            if (meanV) meanV[3 * i + a] = C.sV[3 * i + a];
            if (bulkV) bulkV[3 * i + a] = C.sBulkV[3 * i + a];
        }
        for (int a = 0; a < 9; ++a) {
            if (flux) flux[9 * i + a] = C.sFlux[9 * i + a];
            if (bulkD) bulkD[9 * i + a] = C.sBulkD[9 * i + a];
        }
    }
    *n_points = C.nSamp;
    return EKMC_OK;
}

ekmc_status ekmc_get_position_covariance(const ekmc_engine* e, int32_t cond, int32_t* n_points, double* cov) {
    if (!e || !n_points || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const Cond& C = e->impl->cs[cond];
    const int n = std::min(*n_points, C.nSamp);
    if (cov)
        for (int i = 0; i < 9 * n; ++i) cov[i] = C.sCov[i];
    *n_points = C.nSamp;
    return EKMC_OK;
}

ekmc_status ekmc_get_electrons(const ekmc_engine* e, int32_t cond, double* pos, double* vel, double* energy,
                               double* time) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    return e->impl->get_electrons(cond, pos, vel, energy, time);
}

ekmc_status ekmc_get_status(const ekmc_engine* e, int32_t cond, ekmc_cond_status* out) {
    if (!e || !out || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const Cond& C = e->impl->cs[cond];
    std::memset(out, 0, sizeof(*out));
    out->time = C.time;
    out->trial_collision_frequency = C.trial;
    out->steady_state_time = C.ssTime;
    out->last_mean_energy = C.nSamp > 0 ? C.sE[C.nSamp - 1] : 0;
    out->max_elec_energy = C.maxElecEnergy;
    out->table_max_energy = C.tab.maxEnergy;
    out->real_collisions = C.total;
    out->null_collisions = C.nullc;
    out->n_sampling_points = C.nSamp;
    out->n_integration_points = C.nIntPts;
    out->finished = C.finished ? 1 : 0;
    out->escape_events = C.escapes;
    return EKMC_OK;
}

ekmc_status ekmc_get_kernel_stats(const ekmc_engine* e, ekmc_kernel_stat* stats, int32_t* n_stats) {
    if (!e || !n_stats) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    const int total = K_COUNT + H_COUNT;
    int n = std::min<int>(*n_stats, total);
    for (int k = 0; k < n && stats; ++k) {
        std::memset(&stats[k], 0, sizeof(stats[k]));
        std::snprintf(stats[k].name, sizeof(stats[k].name), "%s", k < K_COUNT ? kKernelNames[k] : kHostNames[k - K_COUNT]);
        stats[k].launches = e->impl->kstat[k].first;
        stats[k].total_ms = e->impl->kstat[k].second;
    }
    *n_stats = total;
    return EKMC_OK;
}

ekmc_status ekmc_reset_kernel_stats(ekmc_engine* e) {
    if (!e) return fail(EKMC_ERR_INVALID_ARGUMENT, "null engine");
    for (auto& k : e->impl->kstat) k = {0, 0.0};
    return EKMC_OK;
}

ekmc_status ekmc_get_event_count(const ekmc_engine* e, uint64_t* events, uint64_t* real) {
    if (!e) return fail(EKMC_ERR_INVALID_ARGUMENT, "null engine");
    if (events) *events = e->impl->totalEvents;
    if (real) *real = e->impl->totalReal;
    return EKMC_OK;
}

const char* ekmc_last_error(void) { return g_error.c_str(); }

ekmc_status ekmc_build_tables(const ekmc_mixture* mixture, double gas_density, int32_t n_interp, double max_energy,
                              double* interp, double* cumul, double* total_freq, double* max_freq, double* energy_step) {
    if (!mixture || n_interp < 2 || !(max_energy > 0)) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    ProcessList pl;
    std::string err = build_process_list(*mixture, pl);
    if (!err.empty()) return fail(EKMC_ERR_INVALID_ARGUMENT, err);
    Tables t;
    build_tables(pl, gas_density, n_interp, max_energy, t);
    if (interp) std::memcpy(interp, t.interp.data(), t.interp.size() * 8);
    if (cumul) std::memcpy(cumul, t.cumul.data(), t.cumul.size() * 8);
    if (total_freq) std::memcpy(total_freq, t.nutot.data(), t.nutot.size() * 8);
    if (max_freq) std::memcpy(max_freq, t.numax.data(), t.numax.size() * 8);
    if (energy_step) *energy_step = t.step;
    return EKMC_OK;
}

ekmc_status ekmc_op_accelerate(ekmc_engine* e, int32_t cond, int64_t n, const double* time, const double* dt,
                               double* pos, double* vel, double* energy, double* gain) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size() || n < 0) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    if (n == 0) return EKMC_OK;
    return e->impl->op_accelerate(cond, n, time, dt, pos, vel, energy, gain);
}

ekmc_status ekmc_op_collide(ekmc_engine* e, int32_t cond, double max_energy, int64_t n, double* vel, double* energy,
                            const double* trial_freq, const double* variates, int32_t* process, double* ejected_vel,
                            double* ejected_energy, double* energy_change) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size() || n < 0) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    if (n == 0) return EKMC_OK;
    return e->impl->op_collide(cond, max_energy, n, vel, energy, trial_freq, variates, process, ejected_vel,
                               ejected_energy, energy_change);
}

ekmc_status ekmc_op_collide_noalloc(ekmc_engine* e, int32_t cond, double max_energy, int64_t n, double* vel,
                                    double* energy, const double* trial_freq, const double* variates,
                                    int32_t* process, double* ejected_vel, double* ejected_energy,
                                    double* energy_change) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size() || n < 0) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    if (n > EKMC_NOALLOC_MAX_EVENTS)
        return fail(EKMC_ERR_CAPACITY, "ekmc_op_collide_noalloc takes at most " +
                                           std::to_string(EKMC_NOALLOC_MAX_EVENTS) + " events per call, got " +
                                           std::to_string(n));
    if (n == 0) return EKMC_OK;
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    const int dev = e->impl->device;
    if (cur != dev) HIP_TRY(hipSetDevice(dev));
    void* stage = nullptr;
    ekmc_status st = scratch_address(dev, &stage);
    if (st == EKMC_OK)
        st = e->impl->op_collide_staged(cond, max_energy, n, vel, energy, trial_freq, variates, process, ejected_vel,
                                        ejected_energy, energy_change, stage);
    if (cur != dev) (void)hipSetDevice(cur);
    return st;
}

ekmc_status ekmc_op_tables(ekmc_engine* e, int32_t cond, double max_energy, double* interp, double* cumul,
                           double* total_freq, double* max_freq, double* energy_step) {
    if (!e || cond < 0 || cond >= (int)e->impl->cs.size()) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    return e->impl->op_tables(cond, max_energy, interp, cumul, total_freq, max_freq, energy_step);
}

ekmc_status ekmc_op_interval_pool(ekmc_engine* e, int32_t cond, int64_t* n_pool, double* pos, double* vel,
                                  double* energy, double* time, uint64_t* stream, uint32_t* event, double* t_sync,
                                  double* trial_freq) {
    if (!e || !n_pool || cond < 0 || cond >= (int)e->impl->cs.size() || *n_pool < 0)
        return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    EngineBase& E = *e->impl;
    if (!E.initialized || E.finalized) return fail(EKMC_ERR_STATE, "the engine must be initialized and not finalized");
    if (E.cs[cond].finished) return fail(EKMC_ERR_STATE, "the condition has finished");
    EngineBase::PoolSnap sn;
    sn.cond = cond;
    E.snap = &sn;
    int running = 0;
    ekmc_status st = E.step(1, &running);
    E.snap = nullptr;
    if (st != EKMC_OK) return st;
    if (!sn.taken) return fail(EKMC_ERR_STATE, "no flights were run for the condition");
    const int64_t m = (int64_t)sn.energy.size();
    const int64_t cap = *n_pool;
    *n_pool = m;
    if (t_sync) *t_sync = sn.tsync;
    if (trial_freq) *trial_freq = sn.trial;
    if (cap < m) return fail(EKMC_ERR_CAPACITY, "n_pool is smaller than the pool");
    if (pos) std::memcpy(pos, sn.pos.data(), (size_t)m * 3 * sizeof(double));
    if (vel) std::memcpy(vel, sn.vel.data(), (size_t)m * 3 * sizeof(double));
    if (energy) std::memcpy(energy, sn.energy.data(), (size_t)m * sizeof(double));
    if (time) std::memcpy(time, sn.time.data(), (size_t)m * sizeof(double));
    if (stream) std::memcpy(stream, sn.stream.data(), (size_t)m * sizeof(uint64_t));
    if (event) std::memcpy(event, sn.event.data(), (size_t)m * sizeof(uint32_t));
    return EKMC_OK;
}

ekmc_status ekmc_op_philox(int64_t n, const uint32_t* counter, const uint32_t* key, uint32_t* out) {
    if (n < 0 || (n > 0 && (!counter || !key || !out))) return fail(EKMC_ERR_INVALID_ARGUMENT, "bad argument");
    if (n == 0) return EKMC_OK;
    int dev = 0;
    if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0) return fail(EKMC_ERR_NO_DEVICE, "no HIP device is visible");
    uint32_t* b = nullptr;
    HIP_TRY(hipMalloc((void**)&b, (size_t)n * 40));
    HIP_TRY(hipMemcpy(b, counter, n * 16, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b + 4 * n, key, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(launch_op_philox(n, b, b + 4 * n, b + 6 * n, nullptr));
    HIP_TRY(hipMemcpy(out, b + 6 * n, n * 16, hipMemcpyDeviceToHost));
    (void)hipFree(b);
    return EKMC_OK;
}

}  // extern "C"
