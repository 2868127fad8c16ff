// kernels.h -- host-callable launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "common.h"

namespace ekmc {

// SoA electron pool.  V is double (plain form) or float (mixed form).
template <typename V> struct DevState {
    double *x, *y, *z, *t;
    V *vx, *vy, *vz, *eps, *tfree, *nue;
    uint64_t* stream;
    uint32_t* event;
    uint8_t* alive;
};

// Device work buffers shared by the interval kernels.
struct DevWork {
    const CondParams* conds;      // [n_cond]
    const int32_t* blockCond;     // [n_blocks] condition of each work block
    double* evPart;               // [n_blocks * event_record_size(P)]
    double* momPart;              // [n_blocks * kMomentRecord]
    int32_t* holesT;              // [n_threads]
    int32_t* surplusT;            // [n_threads]
    int32_t* markT;               // [n_threads]
    int64_t* holeOff;             // [n_threads] exclusive offsets within the condition
    int64_t* surOff;              // [n_threads]
    int64_t* markOff;             // [n_threads]
    uint32_t* holeList;           // [pool capacity], strided per thread over main slots
    PopPlan* plans;               // [n_cond]
    uint32_t* skipSet;            // [n_cond * kMaxSkip]
    uint32_t* status;             // [n_cond]
    double* summary;              // [n_cond * summary_size]
    int summarySize;
    int recSize;
    int nBlocks;
    int nCond;
};

// Histogram accumulators (integer counts summed over samples) and the grid of
// each condition.  grid[6c..6c+6) = {eedf step, cos step, radial step, axial
// step, max speed, fill flag}; phaseRow[c] is the AC phase row or -1.
struct DevHist {
    unsigned long long* eedf;     // [n_cond * nCells]
    unsigned long long* eedfPer;  // [n_cond * nPhases * nCells] or null
    unsigned long long* eadf;     // [n_cond * nCells * nCos] or null
    unsigned long long* evdf;     // [n_cond * nRad * nAx] or null
    const double* grid;
    const int32_t* phaseRow;
    const int32_t* cyl;           // [n_cond] 1 = cylindrically symmetric
    int nCells, nCos, nRad, nAx, nPhases;
};

template <typename V>
hipError_t launch_init(const DevState<V>& s, const DevWork& w, int64_t totalMain,
                       const int64_t* mainBase, const double* thermalDev, const int32_t* slotCond,
                       hipStream_t st);
template <typename V>
hipError_t launch_interval(const DevState<V>& s, const DevWork& w, const DevProcesses& pr,
                           const V* tables, uint32_t caseMask, hipStream_t st);
template <typename V>
hipError_t launch_sample(const DevState<V>& s, const DevWork& w, hipStream_t st);
template <typename V>
hipError_t launch_hist(const DevState<V>& s, const DevWork& w, const DevHist& h, hipStream_t st);
hipError_t launch_reduce(const DevWork& w, int P, hipStream_t st);

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
// Per-kernel HIP events recorded by launch_interval/launch_sample when set.
struct KernelTimer {
    hipEvent_t start[K_COUNT], stop[K_COUNT];
    bool armed[K_COUNT];
    bool enabled;
};
void set_kernel_timer(KernelTimer* t);

}  // namespace ekmc
