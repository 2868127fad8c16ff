// kernels.h -- device data layout and host-callable launchers of the gfx950
// kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "common.h"

namespace ekmc {

// One electron in HBM.  Mixed form: one 64-byte record per pool slot, a
// whole HBM sector, so that storing an electron never writes part of a
// sector.  FP64 form: two records per slot in two arrays, RecA (64 bytes:
// position, clock, velocity, pending free time: the state an ensemble sample
// reads) and RecB (32 bytes: trial frequency of the free time, event counter,
// Philox stream), so that the sampling kernel reads 64 of the 96 bytes.  (The
// same split of the mixed record, 48 + 16 bytes, made the sampling kernel 34
// us faster per interval on the 64-condition scan and the flight kernel 46 us
// slower: a 48-byte record shares sectors with its neighbours.)  The flight
// kernel's ring loads both parts, piece by piece, into the layout of Rec,
// which is RecA followed by RecB.  The energy is not stored: it is 0.5 m_e
// |v|^2 / e of the stored velocity, which is what the last flight of every
// electron computed.
template <typename V> constexpr bool split_record() { return sizeof(V) == 8; }
template <typename V> struct RecA;
template <> struct __attribute__((aligned(16))) RecA<float> {
    double x, y, z, t;          // position [m], clock [s]
    float vx, vy, vz, tfree;    // velocity [m/s], pending free time [s] (< 0: none)
    float nue;                  // trial frequency its free time was drawn with
    uint32_t event;             // collision counter (Philox counter word)
    uint64_t stream;            // Philox stream of this electron
};
template <> struct __attribute__((aligned(16))) RecA<double> {
    double x, y, z, t;
    double vx, vy, vz, tfree;
};
template <typename V> struct RecB;
template <> struct __attribute__((aligned(16))) RecB<float> {};  // not used: the mixed record is whole
template <> struct __attribute__((aligned(16))) RecB<double> {
    double nue;
    uint32_t event, pad0;
    uint64_t stream, pad1;
};
template <typename V> struct Rec;
template <> struct Rec<float> {
    RecA<float> a;
};
template <> struct Rec<double> {
    RecA<double> a;
    RecB<double> b;
};
static_assert(sizeof(RecA<float>) == 64 && sizeof(Rec<float>) == 64, "mixed record is 64 bytes");
static_assert(sizeof(RecA<double>) == 64 && sizeof(RecB<double>) == 32 && sizeof(Rec<double>) == 96,
              "FP64 record is 64 + 32 bytes");

// The part of a record that holds its trial frequency, event counter and
// stream.
template <typename V> __host__ __device__ inline const auto& draw_part(const Rec<V>& q) {
    if constexpr (split_record<V>()) return q.b;
    else return q.a;
}
template <typename V> __host__ __device__ inline auto& draw_part(Rec<V>& q) {
    if constexpr (split_record<V>()) return q.b;
    else return q.a;
}

template <typename V> struct DevState {
    RecA<V>* ra;      // [pool capacity]
    RecB<V>* rb;      // [pool capacity] FP64 form only
    uint8_t* alive;   // [pool capacity] 0 dead/empty, 1 alive, 2 clone made by this interval's
                      // population control (back to 1 in the sampling kernel)
};

// A wave's unit of work: up to `count` main electrons of one condition plus
// the newborn region their ionizations fill.  The chunks of a condition are
// contiguous in the chunk array and in the pool.
struct Chunk {
    int32_t cond;
    int32_t count;        // main electrons [first, first + count)
    int64_t first;        // first main slot; the chunk's hole list lives at holeList[first ..)
    int64_t region;       // first newborn slot
    int32_t regionCap;    // newborn capacity
    int32_t local;        // index of the chunk within its condition
};

// Per-chunk scalar tallies (FP64, reduced in chunk order).
enum ChunkTail { CT_FIELD = 0, CT_NULL, CT_REAL, CT_EVENTS, CT_ESCAPE, CT_MAXE, CT_GROWTH, CT_PAD, CT_SIZE };

// Device work buffers shared by the interval kernels.
struct DevWork {
    const CondParams* conds;      // [n_cond]
    const Chunk* chunks;          // [n_chunks]
    int32_t* holesC;              // [n_chunks] holes left after the chunk's own refill
    int32_t* surplusC;            // [n_chunks] newborns left in the chunk's region
    int32_t* markC;               // [n_chunks]
    int64_t* holeOff;             // [n_chunks] exclusive offsets within the condition
    int64_t* surOff;
    int64_t* markOff;
    uint32_t* holeList;           // [pool capacity]; a chunk's holes at holeList[first ..)
    uint32_t* markKeys;           // [pool capacity]; removal keys of a condition's marks at [base ..)
    double* tailC;                // [n_chunks * CT_SIZE]
    double* momC;                 // [n_chunks * kMomentRecord]
    unsigned long long* accCount; // [n_cond * P] real collisions per process
    long long* accGain;           // [n_cond * P] fixed-point energy gains (2^-24 eV)
    long long* accLoss;           // [n_cond * P]
    PopPlan* plans;               // [n_cond]
    uint32_t* status;             // [n_cond]
    uint32_t* queue;              // [8] work-queue counters of the flight kernels
    double* summary;              // [n_cond * summarySize]
    int summarySize;
    int recSize;                  // 3P + 8
    int nChunks;
    int nCond;
    // Reduction groups, for batches with a condition of more than
    // kFinishDirect work units: fixed ranges of at most kFinishGroup units
    // of one condition, summed by one block each before the per-condition
    // reduction (nGroups 0: the per-condition reduction reads the units).
    const int32_t* groupCond;     // [nGroups] condition of the group
    const int32_t* groupFirst;    // [nGroups] first work unit (global index)
    const int32_t* groupCount;    // [nGroups] work units in the group
    const int32_t* condGroup;     // [n_cond] first group of the condition
    const int32_t* condGroups;    // [n_cond] groups of the condition
    double* tailG;                // [nGroups * CT_SIZE]
    double* momG;                 // [nGroups * kMomentRecord]
    int nGroups;
};
constexpr int kFinishDirect = 4096;  // work units a condition reduces directly
constexpr int kFinishGroup = 2048;   // work units per reduction group

// A contiguous range of main slots of one condition (histogram blocks).
struct SlotRange {
    int32_t cond;
    int32_t count;
    int64_t first;
};

// Histogram accumulators (integer counts summed over samples) and the grid of
// each condition.  grid[6c..6c+6) = {eedf step, cos step, radial step, axial
// step, max speed, fill flag}; the fill flag is 1 for the histogram kernel, 2
// for the flight kernel (counting main electrons as it loads them), 0 for no
// histogram; phaseRow[c] is the AC phase row or -1.
struct DevHist {
    unsigned long long* eedf;     // [n_cond * nCells]
    unsigned long long* eedfPer;  // [n_cond * nPhases * nCells] or null
    unsigned long long* eadf;     // [n_cond * nCells * nCos] or null
    unsigned long long* evdf;     // [n_cond * nRad * nAx] or null
    unsigned long long* evdf3;    // [n_cond * n3^3] (v_x, v_y, v_z) or null
    const double* grid;
    const int32_t* phaseRow;
    const int32_t* cyl;           // [n_cond] 1 = cylindrically symmetric
    const SlotRange* ranges;      // [nRanges] main-slot ranges, one block each
    int nRanges;
    int nCells, nCos, nRad, nAx, nPhases;
    int n3;                       // cells per axis of the 3D EVDF (0: none)
    int histLds;                  // EEDF cells the flight kernel counts in LDS (0: in HBM)
};

// Flight-kernel work lists, one per field case (or FLIGHT_DC_ANY): chunk ids
// in processing order.
struct FlightLists {
    const int32_t* list[kFlightClasses];
    int32_t count[kFlightClasses];
    int32_t grid[kFlightClasses];  // persistent blocks per launch
};

// Sources of the collision tables (interpolateCrossSections,
// BoltzmannMC.C:561-615): each process's prepared integral cross section
// (a superelastic process reads its inelastic partner's) and the constants of
// the interpolation.
struct DevTableSrc {
    int P, I;
    const double* energy;      // concatenated prepared tables
    const double* sigma;
    const int* srcBegin;       // [P] first point of the table process p reads
    const int* srcEnd;         // [P] one past its last point
    const int* isSuper;        // [P]
    const double* minE;        // [P]
    const double* maxE;        // [P]
    const double* relDens;     // [P]
    const double* statW;       // [P]
    const double* thrPartner;  // [P] threshold of the inelastic partner (superelastic only)
};

// Build condition tables on the device for gas density N and energy step
// `step` (= max energy / (I-1)): `tables` receives cumul [I*P], interp [I*P],
// nutot [I], then the null-test envelope envHi [kEnvBins], envLo [kEnvBins]
// over bins of 1 << envShift rows, in V; nutot64 [I] and numax64 [I]
// (running maximum) in FP64.
template <typename V>
hipError_t launch_tables(const DevTableSrc& src, double N, double step, int envShift, V* tables,
                         double* nutot64, double* numax64, hipStream_t st);
// Elements of one condition's table block.
inline int64_t table_block_size(int I, int P) { return (int64_t)I * (2 * P + 1) + 2 * kEnvBins; }

// mainBase[c]: index of condition c's first electron among all main electrons
template <typename V>
hipError_t launch_init(const DevState<V>& s, const DevWork& w, int64_t totalMain,
                       const int64_t* mainBase, const double* thermalDev, hipStream_t st);
// The flights of one interval: one launch per flight class with work.
template <typename V>
hipError_t launch_flights(const DevState<V>& s, const DevWork& w, const DevProcesses& pr,
                          const V* tables, const DevHist& h, const FlightLists& fl, hipStream_t st);
// Population control after the flights: the same launches every interval.
template <typename V>
hipError_t launch_popctl(const DevState<V>& s, const DevWork& w, hipStream_t st);
template <typename V>
hipError_t launch_sample(const DevState<V>& s, const DevWork& w, hipStream_t st);
template <typename V>
hipError_t launch_hist(const DevState<V>& s, const DevHist& h, hipStream_t st);
hipError_t launch_finish(const DevWork& w, int P, hipStream_t st);
// Resident blocks per CU of the flight kernel of field case (or class) fc.
// (histogram configuration h: the EEDF in LDS when h.histLds >= h.nCells)
template <typename V>
int flight_blocks_per_cu(int fc, bool thermal, int P, const DevHist& h);
// Record ring of the flight kernel: ring_segs() segments of kRingSeg main
// records per wave (8 KB per wave for the mixed record, 12 KB for FP64; both
// still fit the blocks per CU their registers allow).
constexpr int kRingSeg = 32;
template <typename V> constexpr int ring_segs() { return 4; }
// Static LDS of one flight-kernel block: the record rings of its waves.
template <typename V> constexpr size_t flight_ring_lds() {
    return (size_t)ring_segs<V>() * (kBlock / kWave) * kRingSeg * sizeof(Rec<V>);
}

template <typename V>
hipError_t launch_op_accelerate(const CondParams* cond, int64_t n, const double* time, const double* dt,
                                double* pos, double* vel, double* energy, double* gain, hipStream_t st);
template <typename V>
hipError_t launch_op_collide(const CondParams* cond, const DevProcesses& pr, const V* tables, int64_t n,
                             double* vel, double* energy, const double* nue, const double* variates,
                             int32_t* pid, double* ejv, double* eje, double* de, hipStream_t st);
hipError_t launch_op_philox(int64_t n, const uint32_t* ctr, const uint32_t* key, uint32_t* out,
                            hipStream_t st);

// kernel names for timing (index matches the event slots used by the engine)
enum KernelId { K_INIT = 0, K_ADVANCE, K_PLAN, K_MARKS, K_APPLY, K_PAIRS, K_MOVE, K_CLONE, K_SAMPLE,
                K_HIST, K_REDUCE, K_COUNT };
extern const char* kKernelNames[K_COUNT];
// Per-kernel HIP events recorded by the launchers when set.
struct KernelTimer {
    hipEvent_t start[K_COUNT], stop[K_COUNT];
    bool armed[K_COUNT];
    int level;  // 0 none, 1 flight kernels, 2 every kernel
};
void set_kernel_timer(KernelTimer* t);

}  // namespace ekmc
