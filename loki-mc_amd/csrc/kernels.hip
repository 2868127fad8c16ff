// kernels.hip -- gfx950 kernels of the electron Monte Carlo engine.
//
// One synchronization interval (BoltzmannMC.C:617-688 plus the population
// control of :1282-1408) runs as this kernel sequence over all active
// conditions at once:
//   k_advance      each thread owns a fixed set of electrons of one condition
//                  (main slots tl, tl+T, ...) and a private newborn region; it
//                  flies every owned electron, and every electron ejected on
//                  the way, to t_sync with the electron held in registers,
//                  counts collisions into LDS, and refills its own holes
//                  (attached electrons) from its own newborns.
//   k_plan         per condition: scans of per-thread holes/surplus, the
//                  population M, the number R of removals or C of clones.
//   k_count_marks / k_plan_marks / k_apply_marks
//                  exact uniform removal of R electrons: Bernoulli marks keyed
//                  by electron identity, then a Floyd sample of marks to drop.
//   k_plan_pairs / k_move / k_clone
//                  move surplus newborns into holes; clone uniformly chosen
//                  electrons into the holes that remain (rejection, as the
//                  reference does at BoltzmannMC.C:1368-1371).
//   k_sample       moments and histograms of the N electrons at t_sync
//                  (BoltzmannMC.C:1410-1482, 1492-1572).
//   k_reduce       fixed-order reduction of the per-block partial records.
// Every reduction is in a fixed order (wave butterflies, then waves, then
// blocks in index order) or on integers, so runs are bitwise reproducible.
#include "kernels.h"
#include "device_physics.h"

namespace ekmc {

const char* kKernelNames[K_COUNT] = {"init", "advance", "plan", "count_marks", "apply_marks",
                                     "plan_pairs", "move", "clone", "sample", "hist", "reduce"};

static KernelTimer* g_timer = nullptr;
void set_kernel_timer(KernelTimer* t) { g_timer = t; }

static inline void timer_start(int k, hipStream_t st) {
    if (g_timer && g_timer->enabled) { (void)hipEventRecord(g_timer->start[k], st); }
}
static inline void timer_stop(int k, hipStream_t st) {
    if (g_timer && g_timer->enabled) { (void)hipEventRecord(g_timer->stop[k], st); g_timer->armed[k] = true; }
}

constexpr double kFix = 16777216.0;  // 2^24: fixed-point scale of per-block energy sums [eV]

// ------------------------------------------------------------ state helpers
template <typename V> __device__ __forceinline__ Electron<V> load_e(const DevState<V>& s, int64_t i) {
    Electron<V> e;
    e.x = s.x[i]; e.y = s.y[i]; e.z = s.z[i]; e.t = s.t[i];
    e.vx = s.vx[i]; e.vy = s.vy[i]; e.vz = s.vz[i]; e.eps = s.eps[i];
    e.tfree = s.tfree[i]; e.nue = s.nue[i]; e.stream = s.stream[i]; e.event = s.event[i];
    return e;
}
template <typename V> __device__ __forceinline__ void store_e(const DevState<V>& s, int64_t i, const Electron<V>& e) {
    s.x[i] = e.x; s.y[i] = e.y; s.z[i] = e.z; s.t[i] = e.t;
    s.vx[i] = e.vx; s.vy[i] = e.vy; s.vz[i] = e.vz; s.eps[i] = e.eps;
    s.tfree[i] = e.tfree; s.nue[i] = e.nue; s.stream[i] = e.stream; s.event[i] = e.event;
}
template <typename V> __device__ __forceinline__ void copy_e(const DevState<V>& s, int64_t src, int64_t dst) {
    store_e(s, dst, load_e(s, src));
    s.alive[dst] = 1;
}

// ------------------------------------------------------- block reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
// Sum over a kBlock-thread block in a fixed order; result valid in thread 0.
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / kWave; ++i) r += red[i];
    __syncthreads();
    return r;
}
__device__ double block_max(double v, double* red) {
    v = wave_max(v);
    int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / kWave; ++i) r = fmax(r, red[i]);
    __syncthreads();
    return r;
}

// Exclusive scan of `in[0,n)` (int32) into `out` (int64) by one block of
// `blockDim.x` threads; returns the total in every thread.
__device__ int64_t block_exclusive_scan(const int32_t* in, int64_t* out, int64_t n, int64_t* lds) {
    const int B = blockDim.x;
    int64_t running = 0;
    for (int64_t base = 0; base < n; base += B) {
        int64_t i = base + threadIdx.x;
        int64_t v = i < n ? in[i] : 0;
        lds[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < B; o <<= 1) {  // Hillis-Steele inclusive scan
            int64_t add = threadIdx.x >= o ? lds[threadIdx.x - o] : 0;
            __syncthreads();
            lds[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < n) out[i] = running + lds[threadIdx.x] - v;
        int64_t tot = lds[B - 1];
        __syncthreads();
        running += tot;
    }
    return running;
}

// Position of a per-thread list element: the thread t whose exclusive offset
// range [off[t], off[t+1]) contains i (largest t with off[t] <= i).
__device__ __forceinline__ int64_t owner_of(const int64_t* off, int64_t T, int64_t i) {
    int64_t lo = 0, hi = T - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ double u53(uint64_t seed, uint64_t stream, uint32_t ev, uint32_t purpose, uint32_t call) {
    uint32_t w[4];
    philox_call(seed, stream, ev, purpose, call, w);
    uint64_t x = ((((uint64_t)w[0]) << 32) | w[1]) >> 11;
    return (double)x * (1.0 / 9007199254740992.0);
}

// ----------------------------------------------------------------- init
// Maxwellian at initialElecTempOverGasTemp * Tg (BoltzmannMC.C:491-516):
// three normals from four uniforms (0,1] of Philox calls 0 and 1 of stream i.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_init(DevState<V> s, const CondParams* conds, int64_t totalMain,
                                                 const int64_t* mainBase, const double* thermalDev,
                                                 const int32_t* slotCond) {
    int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= totalMain) return;
    int c = slotCond[g];
    const CondParams& cp = conds[c];
    int64_t i = g - mainBase[c];
    int64_t slot = cp.base + i;
    uint32_t w[4], w2[4];
    philox_call(cp.seed, (uint64_t)i, 0, P_INIT, 0, w);
    philox_call(cp.seed, (uint64_t)i, 0, P_INIT, 1, w2);
    const double k = 1.0 / 9007199254740992.0;
    double r1 = ((double)(((((uint64_t)w[0]) << 32) | w[1]) >> 11) + 1.0) * k;
    double r2 = ((double)(((((uint64_t)w[2]) << 32) | w[3]) >> 11) + 1.0) * k;
    double r3 = ((double)(((((uint64_t)w2[0]) << 32) | w2[1]) >> 11) + 1.0) * k;
    double r4 = ((double)(((((uint64_t)w2[2]) << 32) | w2[3]) >> 11) + 1.0) * k;
    double aux1 = sqrt(-2.0 * log(r1)), aux2 = 2.0 * kPi * r2;
    double n0 = aux1 * cos(aux2), n1 = aux1 * sin(aux2), n2 = sqrt(-2.0 * log(r3)) * cos(2.0 * kPi * r4);
    double dev = thermalDev[c];
    double vx = n0 * dev, vy = n1 * dev, vz = n2 * dev;
    s.x[slot] = 0; s.y[slot] = 0; s.z[slot] = 0; s.t[slot] = 0;
    s.vx[slot] = (V)vx; s.vy[slot] = (V)vy; s.vz[slot] = (V)vz;
    s.eps[slot] = (V)energy_of(vx, vy, vz);
    s.tfree[slot] = (V)kNonDef;
    s.nue[slot] = (V)cp.trial;
    s.stream[slot] = (uint64_t)i;
    s.event[slot] = 0;
    s.alive[slot] = 1;
}

// --------------------------------------------------------------- advance
struct AdvanceTally {
    double fieldGain, maxE, nNull, nReal, nEvents, nEscape;
    uint32_t status;
};

// Fly one electron to t_sync (the passes of BoltzmannMC.C:630-681 for one
// electron); ejected electrons are queued in the thread's newborn region.
// Returns true when the electron was attached.
template <typename V, int FC, bool THERMAL>
__device__ __forceinline__ bool run_electron(const CondParams& cp, const DevProcesses& pr, const V* cumul,
                                             const V* interp, const V* nutot, const DevState<V>& s,
                                             Electron<V>& e, int64_t region, int64_t q, int& nq,
                                             unsigned long long* ldsCount, long long* ldsGain,
                                             long long* ldsLoss, AdvanceTally& tl) {
    uint32_t iter = 0;
    while (true) {
        CollOut<V> o;
        double gain = 0;
        int st = interval_step<V, FC, THERMAL>(cp, pr, cumul, interp, nutot, e, o, gain);
        tl.fieldGain += gain;
        if (st) return false;
        if (!isfinite(e.t) || !isfinite((double)e.eps)) { tl.status |= ST_NONFINITE; e.t = cp.tsync; return false; }
        tl.nEvents += 1;
        if (o.escape) tl.nEscape += 1;
        tl.status |= o.flags;
        if (o.pid < 0) {
            tl.nNull += 1;
        } else {
            tl.nReal += 1;
            atomicAdd(&ldsCount[o.pid], 1ull);
            long long fx = __double2ll_rn(o.dE * kFix);
            if (o.dE >= 0) atomicAdd((unsigned long long*)&ldsGain[o.pid], (unsigned long long)fx);
            else atomicAdd((unsigned long long*)&ldsLoss[o.pid], (unsigned long long)fx);
            int type = pr.type[o.pid];
            if (type == 1) {
                if (nq < q) {
                    Electron<V> n;
                    n.x = e.x; n.y = e.y; n.z = e.z; n.t = e.t;
                    n.vx = o.ejv[0]; n.vy = o.ejv[1]; n.vz = o.ejv[2]; n.eps = o.ejEps;
                    n.tfree = (V)kNonDef; n.nue = (V)cp.trial;
                    n.stream = child_stream(cp.seed, e.stream, e.event - 1);
                    n.event = 0;
                    store_e(s, region + nq, n);
                    s.alive[region + nq] = 1;
                    ++nq;
                } else {
                    tl.status |= ST_REGION_OVERFLOW;
                }
            } else if (type == 2) {
                return true;
            }
        }
        if (++iter >= kMaxLoopPerSync) { tl.status |= ST_LOOP_CAP; return false; }
    }
}

template <typename V, int FC, bool THERMAL>
__global__ __launch_bounds__(kBlock) void k_advance(DevState<V> s, DevWork w, DevProcesses pr, const V* tables) {
    extern __shared__ unsigned char smem[];
    const int b = blockIdx.x;
    const int c = w.blockCond[b];
    const CondParams cp = w.conds[c];
    if (cp.fieldCase != FC) return;  // another instance flies this block
    const int P = pr.P;
    unsigned long long* ldsCount = (unsigned long long*)smem;      // [P]
    long long* ldsGain = (long long*)(ldsCount + P);                // [P]
    long long* ldsLoss = ldsGain + P;                               // [P]
    double* red = (double*)(ldsLoss + P);                           // [kBlock/kWave]
    double* part = w.evPart + (size_t)b * w.recSize;
    for (int i = threadIdx.x; i < 3 * P; i += kBlock) ((long long*)smem)[i] = 0;
    __syncthreads();
    if (!cp.active) {
        for (int i = threadIdx.x; i < w.recSize; i += kBlock) part[i] = 0;
        return;
    }
    const int64_t T = cp.threads;
    const int64_t tl = (int64_t)(b - cp.blockBegin) * kBlock + threadIdx.x;
    const int64_t gt = cp.threadBegin + tl;
    const int64_t I = cp.I;
    const V* cumul = tables + cp.tableOffset;
    const V* interp = cumul + I * P;
    const V* nutot = interp + I * P;
    const int64_t q = cp.regionPerThread;
    const int64_t region = cp.regionBase + tl * q;

    AdvanceTally ta = {0, 0, 0, 0, 0, 0, 0};
    int nq = 0, holes = 0;
    for (int64_t li = tl; li < cp.nMain; li += T) {
        int64_t slot = cp.base + li;
        if (!s.alive[slot]) { w.holeList[cp.base + tl + (int64_t)holes * T] = (uint32_t)slot; ++holes; continue; }
        Electron<V> e = load_e(s, slot);
        bool dead = run_electron<V, FC, THERMAL>(cp, pr, cumul, interp, nutot, s, e, region, q, nq, ldsCount,
                                                 ldsGain, ldsLoss, ta);
        if (dead) {
            s.alive[slot] = 0;
            w.holeList[cp.base + tl + (int64_t)holes * T] = (uint32_t)slot;
            ++holes;
        } else {
            store_e(s, slot, e);
            s.alive[slot] = 1;
            ta.maxE = fmax(ta.maxE, (double)e.eps);
        }
    }
    for (int j = 0; j < nq; ++j) {  // own newborns, including theirs
        Electron<V> e = load_e(s, region + j);
        bool dead = run_electron<V, FC, THERMAL>(cp, pr, cumul, interp, nutot, s, e, region, q, nq, ldsCount,
                                                 ldsGain, ldsLoss, ta);
        if (dead) s.alive[region + j] = 0;
        else { store_e(s, region + j, e); ta.maxE = fmax(ta.maxE, (double)e.eps); }
    }
    // refill own holes from own newborns, compact the rest to the region front
    int hi = 0, wr = 0;
    for (int j = 0; j < nq; ++j) {
        int64_t src = region + j;
        if (!s.alive[src]) continue;
        if (hi < holes) {
            int64_t dst = w.holeList[cp.base + tl + (int64_t)hi * T];
            copy_e(s, src, dst);
            s.alive[src] = 0;
            ++hi;
        } else {
            if (wr != j) { copy_e(s, src, region + wr); s.alive[src] = 0; }
            ++wr;
        }
    }
    for (int h = hi; h < holes; ++h)
        w.holeList[cp.base + tl + (int64_t)(h - hi) * T] = w.holeList[cp.base + tl + (int64_t)h * T];
    w.holesT[gt] = holes - hi;
    w.surplusT[gt] = wr;
    if (ta.status) atomicOr(&w.status[c], ta.status);

    // block partial record: counts, gains, losses, then the scalar tail
    double sums[5] = {ta.fieldGain, ta.nNull, ta.nReal, ta.nEvents, ta.nEscape};
    double tail[5];
    for (int k = 0; k < 5; ++k) tail[k] = block_sum(sums[k], red);
    double mx = block_max(ta.maxE, red);
    __syncthreads();
    for (int p = threadIdx.x; p < P; p += kBlock) {
        part[p] = (double)ldsCount[p];
        part[P + p] = (double)ldsGain[p] / kFix;
        part[2 * P + p] = (double)ldsLoss[p] / kFix;
    }
    if (threadIdx.x == 0) {
        part[3 * P + EV_FIELD] = tail[0];
        part[3 * P + EV_NULL] = tail[1];
        part[3 * P + EV_REAL] = tail[2];
        part[3 * P + EV_EVENTS] = tail[3];
        part[3 * P + EV_ESCAPE] = tail[4];
        part[3 * P + EV_MAXE] = mx;
        part[3 * P + EV_GROWTH] = 0;
        part[3 * P + EV_PAD] = 0;
    }
}

// ------------------------------------------------------ population control
__global__ __launch_bounds__(1024) void k_plan(DevWork w) {
    __shared__ int64_t lds[1024];
    const int c = blockIdx.x;
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    int64_t H = block_exclusive_scan(w.holesT + cp.threadBegin, w.holeOff + cp.threadBegin, cp.threads, lds);
    int64_t S = block_exclusive_scan(w.surplusT + cp.threadBegin, w.surOff + cp.threadBegin, cp.threads, lds);
    if (threadIdx.x == 0) {
        PopPlan pl = {};
        pl.holes = H;
        pl.surplus = S;
        pl.alive = cp.nMain - H + S;
        pl.remove = S > H ? S - H : 0;
        double R = (double)pl.remove;
        pl.markProb = pl.remove > 0 ? fmin(1.0, (R + 8.0 * sqrt(R) + 20.0) / (double)pl.alive) : 0.0;
        w.plans[c] = pl;
    }
}

// Bernoulli mark of an alive electron (identity = its stream id)
__device__ __forceinline__ bool is_marked(const CondParams& cp, double p, uint64_t stream) {
    return u53(cp.seed, stream, (uint32_t)cp.syncIndex, P_POPCTL, (uint32_t)(cp.syncIndex >> 32) & 0xffffu) < p;
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_count_marks(DevState<V> s, DevWork w) {
    const int b = blockIdx.x;
    const int c = w.blockCond[b];
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    const int64_t T = cp.threads, tl = (int64_t)(b - cp.blockBegin) * kBlock + threadIdx.x, gt = cp.threadBegin + tl;
    const PopPlan pl = w.plans[c];
    if (pl.remove == 0) { w.markT[gt] = 0; return; }
    int m = 0;
    for (int64_t li = tl; li < cp.nMain; li += T) {
        int64_t slot = cp.base + li;
        if (s.alive[slot] && is_marked(cp, pl.markProb, s.stream[slot])) ++m;
    }
    int64_t region = cp.regionBase + tl * cp.regionPerThread;
    for (int j = 0; j < w.surplusT[gt]; ++j)
        if (is_marked(cp, pl.markProb, s.stream[region + j])) ++m;
    w.markT[gt] = m;
}

// Scan of marks and Floyd's sample of the K - R marks that are kept alive.
__global__ __launch_bounds__(1024) void k_plan_marks(DevWork w) {
    __shared__ int64_t lds[1024];
    __shared__ uint32_t hashset[2 * kMaxSkip];
    __shared__ uint32_t chosen[kMaxSkip];
    const int c = blockIdx.x;
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    PopPlan pl = w.plans[c];
    if (pl.remove == 0) return;
    int64_t K = block_exclusive_scan(w.markT + cp.threadBegin, w.markOff + cp.threadBegin, cp.threads, lds);
    int64_t m = K - pl.remove;
    bool bad = K < pl.remove || m > kMaxSkip;
    if (threadIdx.x == 0) {
        pl.marks = K;
        pl.skips = bad ? 0 : m;
        if (bad) pl.status |= ST_POPCTL;
        w.plans[c] = pl;
        if (bad) atomicOr(&w.status[c], (uint32_t)ST_POPCTL);
    }
    if (bad || m == 0) return;
    for (int i = threadIdx.x; i < 2 * kMaxSkip; i += blockDim.x) hashset[i] = 0xffffffffu;
    __syncthreads();
    if (threadIdx.x == 0) {
        // Floyd: for j = K-m .. K-1, t ~ U{0..j}; insert t, or j if t is taken
        int n = 0;
        const uint32_t mask = 2 * kMaxSkip - 1;
        for (int64_t j = K - m; j < K; ++j) {
            double u = u53(cp.seed, (uint64_t)j, (uint32_t)cp.syncIndex, P_POPCTL, 0x8000u | ((uint32_t)(cp.syncIndex >> 32) & 0x7fffu));
            uint32_t t = (uint32_t)fmin(floor(u * (double)(j + 1)), (double)j);
            uint32_t pick = t;
            for (int pass = 0; pass < 2; ++pass) {
                uint32_t h = (pick * 2654435761u) & mask;
                bool found = false;
                while (hashset[h] != 0xffffffffu) {
                    if (hashset[h] == pick) { found = true; break; }
                    h = (h + 1) & mask;
                }
                if (!found) { hashset[h] = pick; chosen[n++] = pick; break; }
                pick = (uint32_t)j;  // t taken: take j (never taken before)
            }
        }
    }
    __syncthreads();
    // bitonic sort of the m chosen indices (padded to a power of two)
    int n2 = 1;
    while (n2 < m) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x)
        if (i >= m) chosen[i] = 0xffffffffu;
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                int l = i ^ j;
                if (l > i) {
                    bool up = (i & k) == 0;
                    uint32_t a = chosen[i], bb = chosen[l];
                    if ((a > bb) == up) { chosen[i] = bb; chosen[l] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) w.skipSet[(size_t)c * kMaxSkip + i] = chosen[i];
}

__device__ __forceinline__ bool in_sorted(const uint32_t* a, int64_t n, uint32_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == v;
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_apply_marks(DevState<V> s, DevWork w) {
    __shared__ double red[kBlock / kWave];
    const int b = blockIdx.x;
    const int c = w.blockCond[b];
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    const PopPlan pl = w.plans[c];
    if (pl.remove == 0 || (pl.status & ST_POPCTL)) return;
    const int64_t T = cp.threads, tl = (int64_t)(b - cp.blockBegin) * kBlock + threadIdx.x, gt = cp.threadBegin + tl;
    const uint32_t* skip = w.skipSet + (size_t)c * kMaxSkip;
    int64_t gi = w.markOff[gt];
    double growth = 0;
    int holes = w.holesT[gt];
    for (int64_t li = tl; li < cp.nMain; li += T) {
        int64_t slot = cp.base + li;
        if (!s.alive[slot] || !is_marked(cp, pl.markProb, s.stream[slot])) continue;
        if (!in_sorted(skip, pl.skips, (uint32_t)gi)) {
            s.alive[slot] = 0;
            growth -= (double)s.eps[slot];
            w.holeList[cp.base + tl + (int64_t)holes * T] = (uint32_t)slot;
            ++holes;
        }
        ++gi;
    }
    int64_t region = cp.regionBase + tl * cp.regionPerThread;
    int sur = w.surplusT[gt];
    for (int j = 0; j < sur; ++j) {
        if (!is_marked(cp, pl.markProb, s.stream[region + j])) continue;
        if (!in_sorted(skip, pl.skips, (uint32_t)gi)) {
            s.alive[region + j] = 0;
            growth -= (double)s.eps[region + j];
        }
        ++gi;
    }
    // refill own holes from own surviving newborns, compact the rest
    int hi = 0, wr = 0;
    for (int j = 0; j < sur; ++j) {
        int64_t src = region + j;
        if (!s.alive[src]) continue;
        if (hi < holes) {
            int64_t dst = w.holeList[cp.base + tl + (int64_t)hi * T];
            copy_e(s, src, dst);
            s.alive[src] = 0;
            ++hi;
        } else {
            if (wr != j) { copy_e(s, src, region + wr); s.alive[src] = 0; }
            ++wr;
        }
    }
    for (int h = hi; h < holes; ++h)
        w.holeList[cp.base + tl + (int64_t)(h - hi) * T] = w.holeList[cp.base + tl + (int64_t)h * T];
    w.holesT[gt] = holes - hi;
    w.surplusT[gt] = wr;
    double g = block_sum(growth, red);
    if (threadIdx.x == 0) w.evPart[(size_t)b * w.recSize + (w.recSize - 8) + EV_GROWTH] += g;
}

__global__ __launch_bounds__(1024) void k_plan_pairs(DevWork w) {
    __shared__ int64_t lds[1024];
    const int c = blockIdx.x;
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    PopPlan pl = w.plans[c];
    if (pl.holes == 0 && pl.surplus == 0) return;
    int64_t H = block_exclusive_scan(w.holesT + cp.threadBegin, w.holeOff + cp.threadBegin, cp.threads, lds);
    int64_t S = block_exclusive_scan(w.surplusT + cp.threadBegin, w.surOff + cp.threadBegin, cp.threads, lds);
    if (threadIdx.x == 0) {
        pl.holes = H;
        pl.surplus = S;
        w.plans[c] = pl;
    }
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_move(DevState<V> s, DevWork w) {
    const int b = blockIdx.x;
    const int c = w.blockCond[b];
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    const PopPlan pl = w.plans[c];
    const int64_t n = pl.holes < pl.surplus ? pl.holes : pl.surplus;
    const int64_t T = cp.threads, tl = (int64_t)(b - cp.blockBegin) * kBlock + threadIdx.x;
    const int64_t* hoff = w.holeOff + cp.threadBegin;
    const int64_t* soff = w.surOff + cp.threadBegin;
    for (int64_t i = tl; i < n; i += T) {
        int64_t th = owner_of(hoff, T, i);
        int64_t dst = w.holeList[cp.base + th + (i - hoff[th]) * T];
        int64_t ts = owner_of(soff, T, i);
        int64_t src = cp.regionBase + ts * cp.regionPerThread + (i - soff[ts]);
        copy_e(s, src, dst);
        s.alive[src] = 0;
    }
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_clone(DevState<V> s, DevWork w) {
    __shared__ double red[kBlock / kWave];
    const int b = blockIdx.x;
    const int c = w.blockCond[b];
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    const PopPlan pl = w.plans[c];
    const int64_t T = cp.threads, tl = (int64_t)(b - cp.blockBegin) * kBlock + threadIdx.x;
    const int64_t* hoff = w.holeOff + cp.threadBegin;
    double growth = 0;
    uint32_t status = 0;
    for (int64_t i = pl.surplus + tl; i < pl.holes; i += T) {
        int64_t th = owner_of(hoff, T, i);
        int64_t dst = w.holeList[cp.base + th + (i - hoff[th]) * T];
        // uniform alive source by rejection over the main slots (BoltzmannMC.C:1368-1371)
        int64_t src = -1;
        for (uint32_t k = 0; k < (1u << 20); ++k) {
            double u = u53(cp.seed, (uint64_t)i, (uint32_t)cp.syncIndex, P_CLONE, k & 0xffffu);
            int64_t cand = (int64_t)fmin(floor(u * (double)cp.nMain), (double)(cp.nMain - 1));
            if (s.alive[cp.base + cand] == 1) { src = cp.base + cand; break; }
        }
        if (src < 0) { status |= ST_POPCTL; continue; }
        Electron<V> e = load_e(s, src);
        growth += (double)e.eps;
        uint32_t wds[4];
        philox_call(cp.seed ^ 0x5851F42D4C957F2Dull, e.stream, (uint32_t)dst, P_CLONE,
                    (uint32_t)(cp.syncIndex & 0xffffu), wds);
        e.stream = ((uint64_t)wds[1] << 32) | wds[0];
        e.event = 0;
        store_e(s, dst, e);
        s.alive[dst] = 2;  // alive, but not a clone source in this pass (BoltzmannMC.C:1374)
    }
    if (status) atomicOr(&w.status[c], status);
    double g = block_sum(growth, red);
    if (threadIdx.x == 0) w.evPart[(size_t)b * w.recSize + (w.recSize - 8) + EV_GROWTH] += g;
}

// ----------------------------------------------------------------- sample
// Ensemble sums of one sampling instant (calculateMeanDataForSwarmParams,
// BoltzmannMC.C:1410-1482) in FP64: eps, r, v, r r^T, r v^T, count, max eps.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_sample(DevState<V> s, DevWork w) {
    __shared__ double red[kBlock / kWave];
    const int b = blockIdx.x;
    const int c = w.blockCond[b];
    const CondParams cp = w.conds[c];
    double* part = w.momPart + (size_t)b * kMomentRecord;
    if (!cp.active) {
        if (threadIdx.x < kMomentRecord) part[threadIdx.x] = 0;
        return;
    }
    const int64_t T = cp.threads, tl = (int64_t)(b - cp.blockBegin) * kBlock + threadIdx.x;
    double m[kMomentRecord];
    for (int k = 0; k < kMomentRecord; ++k) m[k] = 0;
    for (int64_t li = tl; li < cp.nMain; li += T) {
        int64_t slot = cp.base + li;
        double r[3] = {s.x[slot], s.y[slot], s.z[slot]};
        double v[3] = {(double)s.vx[slot], (double)s.vy[slot], (double)s.vz[slot]};
        double e = (double)s.eps[slot];
        m[0] += e;
        for (int a = 0; a < 3; ++a) { m[1 + a] += r[a]; m[4 + a] += v[a]; }
        for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb) { m[7 + 3 * a + bb] += r[a] * r[bb]; m[16 + 3 * a + bb] += r[a] * v[bb]; }
        m[25] += 1;
        m[26] = fmax(m[26], e);
    }
    for (int k = 0; k < kMomentRecord - 1; ++k) {
        double t = block_sum(m[k], red);
        if (threadIdx.x == 0) part[k] = t;
    }
    double mx = block_max(m[26], red);
    if (threadIdx.x == 0) part[26] = mx;
}

// Histograms of one sampling instant (getTimeDependDistributions,
// BoltzmannMC.C:1551-1571; MathFunctions::histogramCount/histogram2DCount,
// MathFunctions.C:61-81,106-127), added to integer accumulators.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_hist(DevState<V> s, DevWork w, DevHist h) {
    extern __shared__ unsigned char smem[];
    uint32_t* lhist = (uint32_t*)smem;  // [nCells]
    const int b = blockIdx.x;
    const int c = w.blockCond[b];
    const CondParams cp = w.conds[c];
    const double* g = h.grid + 6 * (size_t)c;
    if (g[5] == 0) return;
    for (int i = threadIdx.x; i < h.nCells; i += kBlock) lhist[i] = 0;
    __syncthreads();
    const int64_t T = cp.threads, tl = (int64_t)(b - cp.blockBegin) * kBlock + threadIdx.x;
    const double eStep = g[0], cStep = g[1], rStep = g[2], aStep = g[3], maxSpeed = g[4];
    const bool cyl = h.cyl[c] != 0 && h.eadf;
    for (int64_t li = tl; li < cp.nMain; li += T) {
        int64_t slot = cp.base + li;
        double e = (double)s.eps[slot];
        int cell = (int)(e / eStep);
        if (cell >= 0 && cell < h.nCells) atomicAdd(&lhist[cell], 1u);
        if (cyl) {
            double vx = s.vx[slot], vy = s.vy[slot], vz = s.vz[slot];
            double cosA = vz / sqrt(vx * vx + vy * vy + vz * vz);
            int cy = (int)((cosA + 1.0) / cStep);
            if (cell < h.nCells && cell >= 0 && cy < h.nCos && cy >= 0)
                atomicAdd(&h.eadf[((size_t)c * h.nCells + cell) * h.nCos + cy], 1ull);
            int rx = (int)(sqrt(vx * vx + vy * vy) / rStep);
            int az = (int)((vz + maxSpeed) / aStep);
            if (rx < h.nRad && rx >= 0 && az < h.nAx && az >= 0)
                atomicAdd(&h.evdf[((size_t)c * h.nRad + rx) * h.nAx + az], 1ull);
        }
    }
    __syncthreads();
    const int row = h.phaseRow[c];
    for (int i = threadIdx.x; i < h.nCells; i += kBlock) {
        uint32_t n = lhist[i];
        if (!n) continue;
        atomicAdd(&h.eedf[(size_t)c * h.nCells + i], (unsigned long long)n);
        if (row >= 0 && h.eedfPer)
            atomicAdd(&h.eedfPer[((size_t)c * h.nPhases + row) * h.nCells + i], (unsigned long long)n);
    }
}

// ----------------------------------------------------------------- reduce
// summary[c] = [ event sums (recSize) | moment sums (kMomentRecord) | plan (8) | status ]
__global__ __launch_bounds__(256) void k_reduce(DevWork w, int P) {
    const int c = blockIdx.x;
    const CondParams cp = w.conds[c];
    double* out = w.summary + (size_t)c * w.summarySize;
    if (!cp.active) return;
    const int R = w.recSize;
    for (int f = threadIdx.x; f < R + kMomentRecord; f += blockDim.x) {
        double acc = 0;
        if (f < R) {
            bool isMax = f == 3 * P + EV_MAXE;
            for (int bb = 0; bb < cp.blocks; ++bb) {
                double v = w.evPart[(size_t)(cp.blockBegin + bb) * R + f];
                acc = isMax ? fmax(acc, v) : acc + v;
            }
        } else {
            int k = f - R;
            for (int bb = 0; bb < cp.blocks; ++bb) {
                double v = w.momPart[(size_t)(cp.blockBegin + bb) * kMomentRecord + k];
                acc = (k == 26) ? fmax(acc, v) : acc + v;
            }
        }
        out[f] = acc;
    }
    if (threadIdx.x == 0) {
        const PopPlan pl = w.plans[c];
        double* tail = out + R + kMomentRecord;
        tail[0] = (double)pl.holes; tail[1] = (double)pl.surplus; tail[2] = (double)pl.alive;
        tail[3] = (double)pl.remove; tail[4] = (double)pl.marks; tail[5] = (double)pl.skips;
        tail[6] = (double)w.status[c];
        tail[7] = 0;
    }
}

// ------------------------------------------------------- batch operators
template <typename V>
__global__ __launch_bounds__(kBlock) void k_op_accelerate(const CondParams* cond, int64_t n, const double* time,
                                                          const double* dt, double* pos, double* vel,
                                                          double* energy, double* gain) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const CondParams cp = *cond;
    Electron<V> e;
    e.x = pos[3 * i]; e.y = pos[3 * i + 1]; e.z = pos[3 * i + 2]; e.t = time[i];
    e.vx = (V)vel[3 * i]; e.vy = (V)vel[3 * i + 1]; e.vz = (V)vel[3 * i + 2];
    e.eps = energy_of(e.vx, e.vy, e.vz);
    double g = 0;
    flight<-1>(cp, e, dt[i], g);
    pos[3 * i] = e.x; pos[3 * i + 1] = e.y; pos[3 * i + 2] = e.z;
    vel[3 * i] = e.vx; vel[3 * i + 1] = e.vy; vel[3 * i + 2] = e.vz;
    energy[i] = e.eps;
    gain[i] = g;
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_op_collide(const CondParams* cond, DevProcesses pr, const V* tables,
                                                       int64_t n, double* vel, double* energy, const double* nue,
                                                       const double* variates, int32_t* pid, double* ejv,
                                                       double* eje, double* de) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const CondParams cp = *cond;
    const int64_t I = cp.I, P = pr.P;
    const V* cumul = tables + cp.tableOffset;
    const V* interp = cumul + I * P;
    const V* nutot = interp + I * P;
    Electron<V> e;
    e.x = e.y = e.z = e.t = 0;
    e.vx = (V)vel[3 * i]; e.vy = (V)vel[3 * i + 1]; e.vz = (V)vel[3 * i + 2];
    e.eps = (V)energy[i];
    e.nue = (V)nue[i];
    e.tfree = 0; e.stream = 0; e.event = 0;
    EventRng<V> rng(0, 0, 0, variates + (size_t)i * 11);
    CollOut<V> o;
    collide<V, true>(cp, pr, cumul, interp, nutot, e, rng, o);
    pid[i] = o.pid;
    vel[3 * i] = e.vx; vel[3 * i + 1] = e.vy; vel[3 * i + 2] = e.vz;
    energy[i] = e.eps;
    ejv[3 * i] = o.ejv[0]; ejv[3 * i + 1] = o.ejv[1]; ejv[3 * i + 2] = o.ejv[2];
    eje[i] = o.ejEps;
    de[i] = o.dE;
}

__global__ __launch_bounds__(kBlock) void k_op_philox(int64_t n, const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t c[4] = {ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]};
    philox10(c, key[2 * i], key[2 * i + 1]);
    for (int k = 0; k < 4; ++k) out[4 * i + k] = c[k];
}

// ------------------------------------------------------------- launchers
static inline unsigned grid_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <typename V>
hipError_t launch_init(const DevState<V>& s, const DevWork& w, int64_t totalMain, const int64_t* mainBase,
                       const double* thermalDev, const int32_t* slotCond, hipStream_t st) {
    if (totalMain == 0) return hipSuccess;
    timer_start(K_INIT, st);
    hipLaunchKernelGGL(k_init<V>, dim3(grid_of(totalMain)), dim3(kBlock), 0, st, s, w.conds, totalMain,
                       mainBase, thermalDev, slotCond);
    timer_stop(K_INIT, st);
    return hipGetLastError();
}

template <typename V, int FC>
static void launch_advance_case(const DevState<V>& s, const DevWork& w, const DevProcesses& pr, const V* tables,
                                size_t lds, hipStream_t st) {
    dim3 g(w.nBlocks), b(kBlock);
    if (pr.gasTempEffect == 0)
        hipLaunchKernelGGL((k_advance<V, FC, false>), g, b, lds, st, s, w, pr, tables);
    else
        hipLaunchKernelGGL((k_advance<V, FC, true>), g, b, lds, st, s, w, pr, tables);
}

template <typename V>
hipError_t launch_interval(const DevState<V>& s, const DevWork& w, const DevProcesses& pr, const V* tables,
                           uint32_t caseMask, hipStream_t st) {
    size_t lds = (size_t)3 * pr.P * 8 + (kBlock / kWave) * 8;
    dim3 g(w.nBlocks), b(kBlock);
    timer_start(K_ADVANCE, st);
    if (caseMask & (1u << FIELD_DC)) launch_advance_case<V, FIELD_DC>(s, w, pr, tables, lds, st);
    if (caseMask & (1u << FIELD_AC)) launch_advance_case<V, FIELD_AC>(s, w, pr, tables, lds, st);
    if (caseMask & (1u << FIELD_DC_B)) launch_advance_case<V, FIELD_DC_B>(s, w, pr, tables, lds, st);
    if (caseMask & (1u << FIELD_ECR)) launch_advance_case<V, FIELD_ECR>(s, w, pr, tables, lds, st);
    if (caseMask & (1u << FIELD_AC_B)) launch_advance_case<V, FIELD_AC_B>(s, w, pr, tables, lds, st);
    timer_stop(K_ADVANCE, st);
    timer_start(K_PLAN, st);
    hipLaunchKernelGGL(k_plan, dim3(w.nCond), dim3(1024), 0, st, w);
    timer_stop(K_PLAN, st);
    timer_start(K_MARKS, st);
    hipLaunchKernelGGL(k_count_marks<V>, g, b, 0, st, s, w);
    hipLaunchKernelGGL(k_plan_marks, dim3(w.nCond), dim3(1024), 0, st, w);
    timer_stop(K_MARKS, st);
    timer_start(K_APPLY, st);
    hipLaunchKernelGGL(k_apply_marks<V>, g, b, 0, st, s, w);
    timer_stop(K_APPLY, st);
    timer_start(K_PAIRS, st);
    hipLaunchKernelGGL(k_plan_pairs, dim3(w.nCond), dim3(1024), 0, st, w);
    timer_stop(K_PAIRS, st);
    timer_start(K_MOVE, st);
    hipLaunchKernelGGL(k_move<V>, g, b, 0, st, s, w);
    timer_stop(K_MOVE, st);
    timer_start(K_CLONE, st);
    hipLaunchKernelGGL(k_clone<V>, g, b, 0, st, s, w);
    timer_stop(K_CLONE, st);
    return hipGetLastError();
}

template <typename V>
hipError_t launch_sample(const DevState<V>& s, const DevWork& w, hipStream_t st) {
    timer_start(K_SAMPLE, st);
    hipLaunchKernelGGL(k_sample<V>, dim3(w.nBlocks), dim3(kBlock), 0, st, s, w);
    timer_stop(K_SAMPLE, st);
    return hipGetLastError();
}

template <typename V>
hipError_t launch_hist(const DevState<V>& s, const DevWork& w, const DevHist& h, hipStream_t st) {
    size_t lds = (size_t)h.nCells * 4;
    timer_start(K_HIST, st);
    hipLaunchKernelGGL(k_hist<V>, dim3(w.nBlocks), dim3(kBlock), lds, st, s, w, h);
    timer_stop(K_HIST, st);
    return hipGetLastError();
}

hipError_t launch_reduce(const DevWork& w, int P, hipStream_t st) {
    timer_start(K_REDUCE, st);
    hipLaunchKernelGGL(k_reduce, dim3(w.nCond), dim3(256), 0, st, w, P);
    timer_stop(K_REDUCE, st);
    return hipGetLastError();
}

template <typename V>
hipError_t launch_op_accelerate(const CondParams* cond, int64_t n, const double* time, const double* dt,
                                double* pos, double* vel, double* energy, double* gain, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_op_accelerate<V>, dim3(grid_of(n)), dim3(kBlock), 0, st, cond, n, time, dt, pos, vel,
                       energy, gain);
    return hipGetLastError();
}

template <typename V>
hipError_t launch_op_collide(const CondParams* cond, const DevProcesses& pr, const V* tables, int64_t n,
                             double* vel, double* energy, const double* nue, const double* variates,
                             int32_t* pid, double* ejv, double* eje, double* de, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_op_collide<V>, dim3(grid_of(n)), dim3(kBlock), 0, st, cond, pr, tables, n, vel, energy,
                       nue, variates, pid, ejv, eje, de);
    return hipGetLastError();
}

hipError_t launch_op_philox(int64_t n, const uint32_t* ctr, const uint32_t* key, uint32_t* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_op_philox, dim3(grid_of(n)), dim3(kBlock), 0, st, n, ctr, key, out);
    return hipGetLastError();
}

#define EKMC_INSTANTIATE(V)                                                                                   \
    template hipError_t launch_init<V>(const DevState<V>&, const DevWork&, int64_t, const int64_t*,           \
                                       const double*, const int32_t*, hipStream_t);                           \
    template hipError_t launch_interval<V>(const DevState<V>&, const DevWork&, const DevProcesses&, const V*, \
                                           uint32_t, hipStream_t);                                            \
    template hipError_t launch_sample<V>(const DevState<V>&, const DevWork&, hipStream_t);                    \
    template hipError_t launch_hist<V>(const DevState<V>&, const DevWork&, const DevHist&, hipStream_t);      \
    template hipError_t launch_op_accelerate<V>(const CondParams*, int64_t, const double*, const double*,     \
                                                double*, double*, double*, double*, hipStream_t);             \
    template hipError_t launch_op_collide<V>(const CondParams*, const DevProcesses&, const V*, int64_t,       \
                                             double*, double*, const double*, const double*, int32_t*,        \
                                             double*, double*, double*, hipStream_t);
EKMC_INSTANTIATE(double)
EKMC_INSTANTIATE(float)

}  // namespace ekmc
