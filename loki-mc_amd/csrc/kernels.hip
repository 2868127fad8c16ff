// kernels.hip -- gfx950 kernels of the electron Monte Carlo path.
//
// Work unit: a chunk of up to ~1k main electrons of one condition, owned by
// one wavefront, with its own newborn region and hole list.  One
// synchronization interval of every running condition is this sequence:
//   k_advance      persistent per field case: each wave takes chunks from a
//                  work queue and flies every electron of the chunk, and every
//                  electron ejected on the way, to t_sync.  Lanes take the
//                  next electron as soon as theirs is done (ballot-ranked
//                  queue), so a wave never waits for its slowest electron.
//                  Collisions are counted in LDS per wave and added to the
//                  condition's integer accumulators; attached electrons leave
//                  holes that the chunk's own newborns refill.
//   k_plan         per condition: scans of holes/surplus over its chunks, the
//                  population M, removals R or clones C.
//   k_marks / k_plan_marks / k_apply_marks
//                  exact uniform removal of R electrons: Bernoulli marks drawn
//                  by geometric skips, then the R marks with the smallest
//                  random keys (radix select) are removed.
//   k_plan_pairs / k_move / k_clone
//                  surplus newborns fill holes; clones of uniformly chosen
//                  electrons fill the holes that remain (rejection, as the
//                  reference does at BoltzmannMC.C:1368-1371).
//   k_sample       FP64 moments of the N electrons at t_sync, per chunk
//                  (BoltzmannMC.C:1410-1482).
//   k_finish       per condition: fixed-order sums of the chunk records, the
//                  integer accumulators, the plan; resets the queues.
// Every floating-point reduction runs in a fixed order (wave butterflies, then
// chunks in index order), everything else is integer, so runs are bitwise
// reproducible whatever the order in which waves take chunks.
#include "kernels.h"
#include "device_physics.h"

namespace ekmc {

const char* kKernelNames[K_COUNT] = {"init", "advance", "plan", "marks", "apply_marks",
                                     "plan_pairs", "move", "clone", "sample", "hist", "finish"};

static KernelTimer* g_timer = nullptr;
void set_kernel_timer(KernelTimer* t) { g_timer = t; }

// level 1 times the flight kernels only, level 2 every kernel
static inline bool timed(int k) { return g_timer && (g_timer->level >= 2 || (g_timer->level == 1 && k == K_ADVANCE)); }
static inline void timer_start(int k, hipStream_t st) {
    if (timed(k)) { (void)hipEventRecord(g_timer->start[k], st); }
}
static inline void timer_stop(int k, hipStream_t st) {
    if (timed(k)) { (void)hipEventRecord(g_timer->stop[k], st); g_timer->armed[k] = true; }
}

constexpr double kFix = 16777216.0;  // 2^24: fixed-point scale of energy sums [eV]
static inline unsigned grid_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
constexpr int kWavesPerBlock = kBlock / kWave;
// Null-collision events a lane may run in one phase-A step of the flight
// kernel: a null event costs a flight and a null test, not a whole step of
// the wave's bookkeeping.  Measured per interval on the 64-condition scan,
// mixed form: 1 / 2 / 3 / 4 events, 0.981 / 0.963 / 0.989 / 1.062 ms (longer
// runs leave more lanes idle in the loop than they save); FP64 form, whose
// second flight spills registers: 1 / 2 events, 1.596 / 1.663 ms.
template <typename V> constexpr int null_run() { return sizeof(V) == 8 ? 1 : 2; }

// LDS of one wave of the flight kernel: collision tallies [3P] (64-bit), the
// histogram grid of the chunk's condition [8] (FP64), the null-test envelope
// [2 kEnvBins] (V) and the EEDF counts of a histogram taken at load [histLds]
// (16-bit, two cells per word: a chunk holds at most 65535 main electrons)
// (32-bit).  The record ring is a separate static array.
template <typename V> __host__ __device__ constexpr size_t flight_wave_lds(int P, int histLds) {
    return (size_t)3 * P * 8 + 8 * 8 + (size_t)2 * kEnvBins * sizeof(V) + (size_t)((histLds + 1) / 2) * 4;
}

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
// number of set bits of `mask` below this lane
__device__ __forceinline__ int lane_rank(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
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
__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
    return v;
}
// exclusive prefix over the lanes of a wave
__device__ __forceinline__ int wave_exclusive_scan(int v) {
    int incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int t = __shfl_up(incl, o, kWave);
        if (lane_id() >= o) incl += t;
    }
    return incl - v;
}
// Wait until this wave's LDS operations are done (all lanes), so that other
// lanes of the same wave see them.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
}

// ------------------------------------------------------------ records
template <typename V> __device__ __forceinline__ void rec_to_electron(const Rec<V>& q, Electron<V>& e) {
    e.x = q.a.x; e.y = q.a.y; e.z = q.a.z; e.t = q.a.t;
    e.vx = q.a.vx; e.vy = q.a.vy; e.vz = q.a.vz;
    e.eps = energy_of(q.a.vx, q.a.vy, q.a.vz);
    e.tfree = q.a.tfree;
    e.nue = draw_part(q).nue; e.stream = draw_part(q).stream; e.event = draw_part(q).event;
}
template <typename V> __device__ __forceinline__ Rec<V> rec_get(const DevState<V>& s, int64_t i) {
    Rec<V> q;
    q.a = s.ra[i];
    if constexpr (split_record<V>()) q.b = s.rb[i];
    return q;
}
template <typename V> __device__ __forceinline__ void rec_put(const DevState<V>& s, int64_t i, const Rec<V>& q) {
    s.ra[i] = q.a;
    if constexpr (split_record<V>()) s.rb[i] = q.b;
}
template <typename V> __device__ __forceinline__ void load_rec(const DevState<V>& s, int64_t i, Electron<V>& e) {
    rec_to_electron(rec_get(s, i), e);
}
template <typename V> __device__ __forceinline__ void store_rec(const DevState<V>& s, int64_t i, const Electron<V>& e) {
    Rec<V> q = {};
    q.a.x = e.x; q.a.y = e.y; q.a.z = e.z; q.a.t = e.t;
    q.a.vx = e.vx; q.a.vy = e.vy; q.a.vz = e.vz; q.a.tfree = e.tfree;
    draw_part(q).nue = e.nue; draw_part(q).event = e.event; draw_part(q).stream = e.stream;
    rec_put(s, i, q);
}
template <typename V> __device__ __forceinline__ double rec_energy(const RecA<V>& q) {
    return (double)energy_of(q.vx, q.vy, q.vz);
}

// ------------------------------------------------------------ record ring
// The main records of a chunk stream through a per-wave ring in LDS of
// ring_segs<V>() segments of kRingSeg records, filled by direct-to-LDS loads
// issued up to ring_segs - 1 segments ahead of use: a lane taking the next
// main electron reads LDS, and the HBM latency of the record overlaps the
// flights of the wave.  Fill f of a chunk loads chunk-relative main records
// [f kRingSeg, (f + 1) kRingSeg) into segment f % ring_segs; piece k
// (16 bytes) of slot j of a segment is seg[k * kRingSeg + j].  Lane
// j < kRingSeg issues the loads of slot j (the LDS address a lane writes is
// the base plus 16 * lane), so every fill is rec_pieces vector memory
// instructions (lane 0 always has a record).  A wave's vector memory
// operations complete in issue order (vmcnt), so fill f has landed once at
// most rec_pieces * (fills issued after f) operations are outstanding: the
// first read of a segment waits for exactly that, not for the younger fills.
// A read completes inside its own instruction group, so a later fill of that
// segment cannot overtake it.
static_assert(kRingSeg == kWave / 2, "a segment's records once per half-wave");
template <typename V> __host__ __device__ constexpr int rec_pieces() { return (int)(sizeof(Rec<V>) / 16); }
template <typename V> __host__ __device__ constexpr int rec_pieces_a() { return (int)(sizeof(RecA<V>) / 16); }
static_assert(flight_ring_lds<float>() == (size_t)ring_segs<float>() * kWavesPerBlock * kRingSeg * sizeof(Rec<float>),
              "ring size");
typedef __attribute__((address_space(1))) void global_void;
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(3))) uint4 lds_uint4;  // LDS pointers stay LDS pointers

template <typename V> __device__ __forceinline__ lds_uint4* ring_segment(lds_uint4* ring, int record) {
    return ring + ((record / kRingSeg) % ring_segs<V>()) * (rec_pieces<V>() * kRingSeg);
}
// Loads of main records [loaded, loaded + kRingSeg) of a chunk of `count`
// into their segment.  PAIR: each direct-to-LDS load of the wave writes 1 KB,
// the lower half-wave piece k of the 32 records and the upper half-wave piece
// k + 1 (mixed flight kernel 0.900 -> 0.883 ms per interval on the
// 64-condition scan, FP64 1.477 -> 1.466 ms); otherwise the lower half-wave
// loads one piece per instruction.  A fill always holds record `loaded`, so
// it issues exactly ring_loads<V, PAIR>() loads.
template <typename V, bool PAIR> __host__ __device__ constexpr int ring_loads() {
    return PAIR ? rec_pieces<V>() / 2 : rec_pieces<V>();
}
template <typename V, bool PAIR>
__device__ __forceinline__ void ring_fill(const DevState<V>& s, int64_t first, int count, int loaded, lds_uint4* ring) {
    const int lane = lane_id();
    lds_uint4* seg = ring_segment<V>(ring, loaded);
    if constexpr (PAIR) {
        static_assert(rec_pieces<V>() % 2 == 0, "pieces are loaded in pairs");
        const int half = lane / kRingSeg;  // lanes 32-63 load the odd pieces
        const int idx = loaded + (lane % kRingSeg);
        if (idx < count) {
            const char* srcA = (const char*)(s.ra + first + idx);
            const char* srcB = (const char*)(s.rb + first + idx);
#pragma unroll
            for (int k = 0; k < rec_pieces<V>(); k += 2) {
                const int piece = k + half;
                const char* src = piece < rec_pieces_a<V>() ? srcA + 16 * piece : srcB + 16 * (piece - rec_pieces_a<V>());
                __builtin_amdgcn_global_load_lds((global_void*)src, (lds_void*)(seg + k * kRingSeg), 16, 0, 0);
            }
        }
    } else {
        const int idx = loaded + lane;
        if (lane < kRingSeg && idx < count) {
            const char* srcA = (const char*)(s.ra + first + idx);
            const char* srcB = (const char*)(s.rb + first + idx);
#pragma unroll
            for (int k = 0; k < rec_pieces<V>(); ++k) {
                const char* src = k < rec_pieces_a<V>() ? srcA + 16 * k : srcB + 16 * (k - rec_pieces_a<V>());
                __builtin_amdgcn_global_load_lds((global_void*)src, (lds_void*)(seg + k * kRingSeg), 16, 0, 0);
            }
        }
    }
}
// s_waitcnt with vmcnt(N) (expcnt and lgkmcnt not waited for)
template <int N> __device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}
// Wait until a fill followed by `later` younger fills has landed (each fill
// is ring_loads<V, PAIR>() loads; completion is in order).
template <typename V, bool PAIR> __device__ __forceinline__ void ring_wait(int later) {
    constexpr int p = ring_loads<V, PAIR>();
    static_assert(ring_segs<V>() <= 4, "ring_wait covers at most 3 younger fills");
    if (later <= 0) wait_vmcnt<0>();
    else if (later == 1) wait_vmcnt<p>();
    else if (later == 2) wait_vmcnt<2 * p>();
    else wait_vmcnt<3 * p>();
}
// The pieces of one ring slot at LDS byte address `a`, read and waited for
// in one group.
__device__ __forceinline__ void ring_pieces(uint32_t a, uint4 (&p)[4]) {
    asm volatile(
        "ds_read_b128 %0, %4\n\t"
        "ds_read_b128 %1, %4 offset:512\n\t"
        "ds_read_b128 %2, %4 offset:1024\n\t"
        "ds_read_b128 %3, %4 offset:1536\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(p[0]), "=&v"(p[1]), "=&v"(p[2]), "=&v"(p[3])
        : "v"(a)
        : "memory");
}
__device__ __forceinline__ void ring_pieces(uint32_t a, uint4 (&p)[6]) {
    asm volatile(
        "ds_read_b128 %0, %6\n\t"
        "ds_read_b128 %1, %6 offset:512\n\t"
        "ds_read_b128 %2, %6 offset:1024\n\t"
        "ds_read_b128 %3, %6 offset:1536\n\t"
        "ds_read_b128 %4, %6 offset:2048\n\t"
        "ds_read_b128 %5, %6 offset:2560\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(p[0]), "=&v"(p[1]), "=&v"(p[2]), "=&v"(p[3]), "=&v"(p[4]), "=&v"(p[5])
        : "v"(a)
        : "memory");
}
static_assert(kRingSeg * 16 == 512, "ring piece stride");
// LDS atomic add without return, issued as is: the compiler would otherwise
// order it after the outstanding ring loads (it cannot tell the histogram
// from the ring), which would expose their HBM latency.
__device__ __forceinline__ void lds_add_u32(uint32_t* p, uint32_t v) {
    asm volatile("ds_add_u32 %0, %1" : : "v"((uint32_t)(size_t)(lds_void*)p), "v"(v) : "memory");
}
// Slot j of a ring segment.
template <typename V> __device__ __forceinline__ void ring_read(const lds_uint4* seg, int j, Electron<V>& e) {
    uint4 p[rec_pieces<V>()];
    ring_pieces((uint32_t)(size_t)(seg + j), p);
    Rec<V> q;
    __builtin_memcpy(&q, p, sizeof(q));
    rec_to_electron(q, e);
}

// ------------------------------------------------------- block helpers
// Sum over a kBlock-thread block in a fixed order; result valid in thread 0.
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    int lane = lane_id(), wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i) r += red[i];
    __syncthreads();
    return r;
}
__device__ double block_max(double v, double* red) {
    v = wave_max(v);
    int lane = lane_id(), wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i) r = fmax(r, red[i]);
    __syncthreads();
    return r;
}

// Exclusive scan of `in[0,n)` (int32) into `out` (int64) by one block of
// `blockDim.x` threads (a multiple of the wave, at most 64 waves); returns the
// total in every thread.  Tiles of 8 consecutive values per thread: a serial
// scan in each thread, a shuffle scan of the thread totals in each wave, a
// scan of the wave totals by the first wave, three barriers per tile (a
// one-value-per-thread Hillis-Steele tile took twenty; a condition of 1e8
// electrons has about 98,000 work units).  Integer sums: any order gives the
// same offsets.  `lds` holds at least 2 kWave + 1 values.
__device__ int64_t block_exclusive_scan(const int32_t* in, int64_t* out, int64_t n, int64_t* lds) {
    constexpr int E = 8;
    const int lane = lane_id(), wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const int64_t tile = (int64_t)blockDim.x * E;
    int64_t running = 0;
    for (int64_t base = 0; base < n; base += tile) {
        const int64_t i0 = base + (int64_t)threadIdx.x * E;
        int64_t v[E], t = 0;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            v[k] = i0 + k < n ? in[i0 + k] : 0;
            t += v[k];
        }
        int64_t incl = t;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int64_t u = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += u;
        }
        if (lane == kWave - 1) lds[wid] = incl;
        __syncthreads();
        if (wid == 0) {
            const int64_t wt = lane < nw ? lds[lane] : 0;
            int64_t wi = wt;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int64_t u = __shfl_up(wi, o, kWave);
                if (lane >= o) wi += u;
            }
            if (lane < nw) lds[kWave + lane] = wi - wt;
            if (lane == kWave - 1) lds[2 * kWave] = wi;
        }
        __syncthreads();
        int64_t pre = running + lds[kWave + wid] + incl - t;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            if (i0 + k < n) out[i0 + k] = pre;
            pre += v[k];
        }
        running += lds[2 * kWave];
        __syncthreads();
    }
    return running;
}

// Index t of the range [off[t], off[t+1]) that contains i (largest t with
// off[t] <= i) over n exclusive offsets.
__device__ __forceinline__ int64_t owner_of(const int64_t* off, int64_t n, int64_t i) {
    int64_t lo = 0, hi = n - 1;
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
__global__ __launch_bounds__(kBlock) void k_init(DevState<V> s, const CondParams* conds, int nCond,
                                                 int64_t totalMain, const int64_t* mainBase,
                                                 const double* thermalDev) {
    int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= totalMain) return;
    int lo = 0, hi = nCond - 1;  // condition whose main range holds g
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (mainBase[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int c = lo;
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
    Electron<V> e;
    e.x = 0; e.y = 0; e.z = 0; e.t = 0;
    e.vx = (V)(n0 * dev); e.vy = (V)(n1 * dev); e.vz = (V)(n2 * dev);
    e.eps = 0;
    e.tfree = (V)kNonDef;
    e.nue = (V)cp.trial;
    e.stream = (uint64_t)i;
    e.event = 0;
    store_rec(s, slot, e);
    s.alive[slot] = 1;
}

// ----------------------------------------------------- newborn refill
// After a chunk's electrons are at t_sync (or after removals): surviving
// newborns of the chunk's region, in region order, fill the chunk's holes in
// hole-list order; the rest move to the front of the region.  Wave-level, in
// groups of 64 newborns: every lane loads its record before any lane stores
// (each store waits on the group's loads), sources are cleared before
// destinations are set, and destinations never lie beyond the group.
template <typename V>
__device__ void wave_refill(const DevState<V>& s, uint32_t* holeList, const Chunk& ch, int& holes, int nNew,
                            int& remaining) {
    const int lane = lane_id();
    int used = 0;
    for (int b = 0; b < nNew; b += kWave) {
        const int j = b + lane;
        const bool al = j < nNew && s.alive[ch.region + j] != 0;
        const uint64_t am = __ballot(al);
        const int r = used + lane_rank(am);
        int64_t dst = -1;
        if (al) {
            if (r < holes) dst = holeList[ch.first + r];
            else if (r - holes != j) dst = ch.region + (r - holes);
        }
        if (dst >= 0) {
            const Rec<V> q = rec_get(s, ch.region + j);
            rec_put(s, dst, q);
            s.alive[ch.region + j] = 0;
        }
        __builtin_amdgcn_wave_barrier();
        if (dst >= 0) s.alive[dst] = 1;
        used += __popcll(am);
    }
    const int filled = used < holes ? used : holes;
    for (int b = filled; b < holes; b += kWave) {  // remaining holes to the list front
        const int h = b + lane;
        uint32_t v = 0;
        if (h < holes) v = holeList[ch.first + h];
        __builtin_amdgcn_wave_barrier();
        if (h < holes) holeList[ch.first + h - filled] = v;
    }
    holes -= filled;
    remaining = used - filled;
}

// ------------------------------------------------------------- histograms
// One electron into the histograms of condition c (getTimeDependDistributions,
// BoltzmannMC.C:1551-1571; MathFunctions::histogramCount/histogram2DCount,
// MathFunctions.C:61-81,106-127): the EEDF into `lhist` (LDS, flushed later)
// or directly into HBM (phase row `row`, or -1), the EADF and EVDF into HBM
// when `cyl`.  The per-condition values are read once per chunk by the caller.
// Histogram modes of a flight launch, fixed per launch so that the paths a
// launch cannot take are not compiled into it: the EEDF in LDS and no
// cylindrically symmetric condition, the EEDF in LDS and EADF/EVDF in HBM,
// or everything in HBM.
enum HistMode { HIST_LDS = 0, HIST_LDS_CYL = 1, HIST_HBM = 2 };

template <typename V, int HIST>
__device__ __forceinline__ void hist_count(const DevHist& h, int c, const double* g, const Electron<V>& e,
                                           uint32_t* lhist, int row, bool cyl) {
    const double en = (double)e.eps;
    const int cell = (int)(en / g[0]);
    if (cell >= 0 && cell < h.nCells) {
        if (HIST != HIST_HBM) lds_add_u32(lhist + (cell >> 1), (cell & 1) ? 0x10000u : 1u);
        else {
            atomicAdd(&h.eedf[(size_t)c * h.nCells + cell], 1ull);
            if (row >= 0 && h.eedfPer) atomicAdd(&h.eedfPer[((size_t)c * h.nPhases + row) * h.nCells + cell], 1ull);
        }
    }
    if (HIST != HIST_LDS && cyl) {
        const double vx = e.vx, vy = e.vy, vz = e.vz;
        const double cosA = vz / sqrt(vx * vx + vy * vy + vz * vz);
        const int cy = (int)((cosA + 1.0) / g[1]);
        if (cell < h.nCells && cell >= 0 && cy < h.nCos && cy >= 0)
            atomicAdd(&h.eadf[((size_t)c * h.nCells + cell) * h.nCos + cy], 1ull);
        const int rx = (int)(sqrt(vx * vx + vy * vy) / g[2]);
        const int az = (int)((vz + g[4]) / g[3]);
        if (rx < h.nRad && rx >= 0 && az < h.nAx && az >= 0)
            atomicAdd(&h.evdf[((size_t)c * h.nRad + rx) * h.nAx + az], 1ull);
    }
}

// --------------------------------------------------------------- advance
// One chunk: every main electron and every newborn flies to t_sync (the
// passes of BoltzmannMC.C:630-681 for this set of electrons).
template <typename V, int FC, bool THERMAL, int HIST>
__device__ __forceinline__ void advance_chunk(const DevState<V>& s, const DevWork& w, const DevProcesses& pr,
                                              const V* tables, const DevHist& h, int cid, unsigned long long* cnt,
                                              lds_uint4* ring) {
    const int lane = lane_id();
    const int P = pr.P;
    long long* gainL = (long long*)(cnt + P);
    long long* lossL = gainL + P;
    const Chunk ch = w.chunks[cid];
    const int c = ch.cond;
    const CondParams cp = w.conds[c];
    const int64_t I = cp.I;
    const V* cumul = tables + cp.tableOffset;
    const V* interp = cumul + I * P;
    const V* nutot = interp + I * P;
    double* hg = (double*)(cnt + 3 * P);
    V* envHi = (V*)(hg + 8);
    V* envLo = envHi + kEnvBins;
    uint32_t* lhist = (uint32_t*)(envLo + kEnvBins);
    for (int i = lane; i < 3 * P; i += kWave) ((long long*)cnt)[i] = 0;
    for (int i = lane; i < 2 * kEnvBins; i += kWave) envHi[i] = nutot[I + i];
    // histograms of the sampled state (the state this interval starts from),
    // counted as main electrons are loaded (fill flag 2 of the condition);
    // the condition's grid is copied to LDS
    if (lane < 6) hg[lane] = h.grid[6 * (size_t)c + lane];
    const bool histOn = h.grid[6 * (size_t)c + 5] == 2.0;
    const bool histLds = HIST != HIST_HBM && histOn;
    const int histRow = histOn ? h.phaseRow[c] : -1;
    const bool histCyl = HIST != HIST_LDS && histOn && h.eadf && h.cyl[c] != 0;
    if (histLds)
        for (int i = lane; i < (h.nCells + 1) / 2; i += kWave) lhist[i] = 0;
    wave_lds_sync();

    // two record pieces per ring load, except with the cylindrical histograms
    // counted at load: there the pairs raised the flight kernel from 163 to
    // 169 VGPRs, past the 168 of three waves per SIMD
    constexpr bool kPairFill = HIST != HIST_LDS_CYL;
    const int count = ch.count;
    int nextMain = 0, nq = 0, nextNew = 0, holes = 0;
    int loaded = 0;  // main records whose ring loads are issued
    int landed = 0;  // ring fills known to have landed
    constexpr int kAhead = (ring_segs<V>() - 1) * kRingSeg;
    for (; loaded < count && loaded <= nextMain + kAhead; loaded += kRingSeg)
        ring_fill<V, kPairFill>(s, ch.first, count, loaded, ring);
    Electron<V> e;
    bool have = false, isNew = false, parked = false;
    int64_t slot = 0;
    uint32_t c0[4] = {0, 0, 0, 0};  // call 0 of a parked electron's event
    uint32_t iter = 0, nNull = 0, nReal = 0, nEsc = 0, status = 0;
    double fieldGain = 0, maxE = 0;
    while (true) {
        // lanes without an electron take the next main electron, then newborns
        // (every main slot holds an alive electron after population control)
        const uint64_t need = __ballot(!have);
        if (need) {
            const int nNeed = __popcll(need);
            const int availMain = (loaded < count ? loaded : count) - nextMain;
            const int takeMain = availMain < nNeed ? availMain : nNeed;
            const int availNew = nq - nextNew;
            const int takeNew = (nNeed - takeMain) < availNew ? (nNeed - takeMain) : availNew;
            if (takeMain > 0) {
                // main records [nextMain, nextMain + takeMain): the youngest
                // fill they need, and the fills issued after it
                const int need = (nextMain + takeMain - 1) / kRingSeg;
                if (need >= landed) {
                    ring_wait<V, kPairFill>(loaded / kRingSeg - 1 - need);
                    landed = need + 1;
                }
            }
            if (!have) {
                const int r = lane_rank(need);
                if (r < takeMain) {
                    slot = ch.first + nextMain + r;
                    isNew = false;
                    have = true;
                } else if (r - takeMain < takeNew) {
                    slot = ch.region + nextNew + (r - takeMain);
                    isNew = true;
                    have = true;
                }
                if (have) {
                    if (isNew) {
                        load_rec(s, slot, e);
                    } else {
                        const int idx = (int)(slot - ch.first);
                        ring_read<V>(ring_segment<V>(ring, idx), idx % kRingSeg, e);
                    }
                    iter = 0;
                    if (histOn && !isNew) hist_count<V, HIST>(h, c, hg, e, lhist, histRow, histCyl);
                }
            }
            nextMain += takeMain;
            nextNew += takeNew;
        }
        const uint64_t live = __ballot(have);
        if (live == 0) break;
        // refill the ring segments this step's lanes emptied, two or more at a
        // time, before the flights: every wait for a table load also waits for
        // the ring fills issued before it (a wave's vector memory operations
        // complete in order), so fewer, larger and earlier fills leave fewer
        // steps waiting on HBM latency (flight kernel per interval on the
        // 64-condition scan, mixed / FP64: 3 and 2 segments filled one at a
        // time at the end of a step 0.964 / 1.613 ms, 4 segments two at a time
        // 0.936 / 1.526 ms, the same filled before the flights 0.920 ms mixed)
        if (loaded + kRingSeg <= nextMain + kAhead)
            for (; loaded < count && loaded <= nextMain + kAhead; loaded += kRingSeg)
                ring_fill<V, kPairFill>(s, ch.first, count, loaded, ring);

        bool done = false, dead = false, spawn = false;
        CollOut<V> o;
        o.pid = -1;
        // phase A: flight, and the null test of the cold branch; a lane whose
        // event is a null collision flies on in the same step, up to null_run<V>()
        // events (the draws are keyed by the event, so the run length changes
        // no trajectory)
        if (have && !parked) {
            for (int run = 0; run < null_run<V>(); ++run) {
                double gain = 0;
                EventRng<V> rng(cp.seed, e.stream, e.event);
                const int st = step_a<V, FC, THERMAL>(cp, pr, nutot, envHi, envLo, e, rng, o, gain);
                fieldGain += gain;
                if (st == STEP_DONE) {
                    done = true;
                    break;
                }
                if (!isfinite(e.t) || !isfinite((double)e.eps)) {
                    status |= ST_NONFINITE;
                    e.t = cp.tsync;
                    done = true;
                    break;
                }
                if (o.escape) ++nEsc;
                if (st != STEP_NULL) {
                    parked = true;
                    c0[0] = rng.w0[0]; c0[1] = rng.w0[1]; c0[2] = rng.w0[2]; c0[3] = rng.w0[3];
                    break;
                }
                ++nNull;
                after_event(cp, e);
                if (++iter >= kMaxLoopPerSync) {
                    status |= ST_LOOP_CAP;
                    done = true;
                    break;
                }
            }
        }
        // phase B: real collisions, once enough lanes wait for one (or all do)
        const uint64_t pk = __ballot(parked);
        if (pk && (__popcll(pk) >= pr.parkLanes || pk == (live & ~__ballot(done)))) {
            if (parked) {
                EventRng<V> rng(cp.seed, e.stream, e.event, c0);
                step_b<V, THERMAL>(cp, pr, cumul, interp, e, rng, o);
                status |= o.flags;
                if (o.pid < 0) {
                    ++nNull;
                } else {
                    ++nReal;
                    atomicAdd(&cnt[o.pid], 1ull);
                    const long long fx = __double2ll_rn(o.dE * kFix);
                    if (o.dE >= 0) atomicAdd((unsigned long long*)&gainL[o.pid], (unsigned long long)fx);
                    else atomicAdd((unsigned long long*)&lossL[o.pid], (unsigned long long)fx);
                    const int type = pr.proc[o.pid].type;
                    if (type == 1) spawn = true;
                    else if (type == 2) { done = true; dead = true; }
                }
                after_event(cp, e);
                parked = false;
                if (++iter >= kMaxLoopPerSync) { status |= ST_LOOP_CAP; done = true; }
            }
        }
        // ejected electrons join the chunk's newborn queue, in lane order
        const uint64_t sp = __ballot(spawn);
        if (sp) {
            if (spawn) {
                const int idx = nq + lane_rank(sp);
                if (idx < ch.regionCap) {
                    Electron<V> n;
                    n.x = e.x; n.y = e.y; n.z = e.z; n.t = e.t;
                    n.vx = o.ejv[0]; n.vy = o.ejv[1]; n.vz = o.ejv[2]; n.eps = o.ejEps;
                    n.tfree = (V)kNonDef; n.nue = (V)cp.trial;
                    n.stream = child_stream(cp.seed, e.stream, e.event - 1);
                    n.event = 0;
                    store_rec(s, ch.region + idx, n);
                    s.alive[ch.region + idx] = 1;
                } else {
                    status |= ST_REGION_OVERFLOW;
                }
            }
            const int tot = nq + __popcll(sp);
            nq = tot < ch.regionCap ? tot : ch.regionCap;
        }
        // finished electrons: store, or leave a hole in the chunk's list
        const uint64_t dm = __ballot(done && dead && !isNew);
        if (done) {
            if (dead) {
                s.alive[slot] = 0;
                if (!isNew) w.holeList[ch.first + holes + lane_rank(dm)] = (uint32_t)slot;
            } else {
                store_rec(s, slot, e);
                maxE = fmax(maxE, (double)e.eps);
            }
            have = false;
        }
        holes += __popcll(dm);
    }

    int remaining = 0;
    wave_refill<V>(s, w.holeList, ch, holes, nq, remaining);

    const double fg = wave_sum(fieldGain);
    const double mx = wave_max(maxE);
    const uint32_t nn = wave_sum_u(nNull), nr = wave_sum_u(nReal), ne = wave_sum_u(nEsc);
    const uint32_t stw = wave_or(status);
    if (lane == 0) {
        w.holesC[cid] = holes;
        w.surplusC[cid] = remaining;
        double* t = w.tailC + (size_t)cid * CT_SIZE;
        t[CT_FIELD] = fg;
        t[CT_NULL] = (double)nn;
        t[CT_REAL] = (double)nr;
        t[CT_EVENTS] = (double)nn + (double)nr;
        t[CT_ESCAPE] = (double)ne;
        t[CT_MAXE] = mx;
        t[CT_GROWTH] = 0;
        t[CT_PAD] = 0;
        if (stw) atomicOr(&w.status[c], stw);
    }
    wave_lds_sync();
    for (int p = lane; p < P; p += kWave) {
        const unsigned long long n = cnt[p];
        if (n) {
            atomicAdd(&w.accCount[(size_t)c * P + p], n);
            const long long gq = gainL[p], lq = lossL[p];
            if (gq) atomicAdd((unsigned long long*)&w.accGain[(size_t)c * P + p], (unsigned long long)gq);
            if (lq) atomicAdd((unsigned long long*)&w.accLoss[(size_t)c * P + p], (unsigned long long)lq);
        }
    }
    if (histLds) {
        const int row = h.phaseRow[c];
        for (int i = lane; i < h.nCells; i += kWave) {
            const uint32_t n = (lhist[i >> 1] >> ((i & 1) * 16)) & 0xffffu;
            if (!n) continue;
            atomicAdd(&h.eedf[(size_t)c * h.nCells + i], (unsigned long long)n);
            if (row >= 0 && h.eedfPer)
                atomicAdd(&h.eedfPer[((size_t)c * h.nPhases + row) * h.nCells + i], (unsigned long long)n);
        }
    }
    wave_lds_sync();
}

// Waves per SIMD the flight kernel is compiled for.  FP64: 256 registers per
// lane, two waves; left to itself the compiler takes 290 (256 VGPRs and 34
// AGPRs), one wave per SIMD, and the few VGPRs spilled at two waves cost less
// than the latency one wave cannot hide (measured 2.27 -> 1.60 ms per
// interval on the 64-condition scan in FP64).  Mixed form: the compiler's
// choice (150-166 VGPRs, three waves, for the FP32 flight formulas); asking
// for three waves explicitly gave the same occupancy and a slower schedule
// (0.962 -> 1.000 ms), asking for four (with a two-segment record ring to
// fit the LDS) spilled 33-41 VGPRs and ran at 1.062 ms.
template <typename V, int FC> constexpr int flight_waves_per_eu() { return sizeof(V) == 8 ? 2 : 1; }

// Persistent flight kernel of one field case: every wave takes chunks from
// the list through an atomic counter until the list is exhausted.
template <typename V, int FC, bool THERMAL, int HIST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(flight_waves_per_eu<V, FC>()))) void k_advance(DevState<V> s, DevWork w, DevProcesses pr, const V* tables,
                                                    DevHist h, const int32_t* list, int32_t nList,
                                                    uint32_t* queue) {
    extern __shared__ unsigned char smem[];
    constexpr int kRingWave = ring_segs<V>() * rec_pieces<V>() * kRingSeg;  // 16-byte pieces per wave
    __shared__ uint4 rings[kWavesPerBlock * kRingWave];
    const int wid = threadIdx.x / kWave;
    // the process records in LDS: a collision reads its process's constants
    // at LDS latency, outside the in-order queue of the ring loads
    uint4* procL = (uint4*)smem;
    for (int i = threadIdx.x; i < pr.P * (int)(sizeof(ProcRec) / 16); i += kBlock)
        procL[i] = ((const uint4*)pr.proc)[i];
    __syncthreads();
    DevProcesses prl = pr;
    prl.proc = (const ProcRec*)smem;
    unsigned long long* cnt =
        (unsigned long long*)(smem + (size_t)pr.P * sizeof(ProcRec) + (size_t)wid * flight_wave_lds<V>(pr.P, h.histLds));
    lds_uint4* ring = (lds_uint4*)(rings + wid * kRingWave);
    while (true) {
        uint32_t item = 0;
        if (lane_id() == 0) item = atomicAdd(queue, 1u);
        item = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl((int)item, 0, kWave));
        if (item >= (uint32_t)nList) break;
        advance_chunk<V, FC, THERMAL, HIST>(s, w, prl, tables, h, list[item], cnt, ring);
    }
}

// ------------------------------------------------------ population control
__global__ __launch_bounds__(1024) void k_plan(DevWork w) {
    __shared__ int64_t lds[1024];
    const int c = blockIdx.x;
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    const int64_t H = block_exclusive_scan(w.holesC + cp.chunkBegin, w.holeOff + cp.chunkBegin, cp.nChunks, lds);
    const int64_t S = block_exclusive_scan(w.surplusC + cp.chunkBegin, w.surOff + cp.chunkBegin, cp.nChunks, lds);
    if (threadIdx.x == 0) {
        PopPlan pl = {};
        pl.holes = H;
        pl.surplus = S;
        pl.alive = cp.nMain - H + S;
        pl.remove = S > H ? S - H : 0;
        const double R = (double)pl.remove;
        pl.markProb = pl.remove > 0 ? fmin(1.0, (R + 8.0 * sqrt(R) + 20.0) / (double)pl.alive) : 0.0;
        w.plans[c] = pl;
    }
}

// Bernoulli(p) marks over a chunk's candidates [0, L): main slots
// [0, count) (holes are not alive and never count) and the surplus
// newborns [count, L).  Lane l draws geometric skips over its segment
// [l L / 64, (l+1) L / 64) with Philox keyed by (chunk, lane, interval).
struct MarkGen {
    uint64_t seed, key;
    uint32_t sync, call;
    double logq;
    bool all;
    int64_t pos, hi;
    uint32_t w[4];
    __device__ MarkGen(const CondParams& cp, const Chunk& ch, int lane, double p, int64_t L) {
        seed = cp.seed;
        key = ((uint64_t)(uint32_t)lane << 32) | (uint32_t)ch.local;
        sync = (uint32_t)cp.syncIndex;
        call = 0;
        all = !(p < 1.0);
        logq = all ? -1.0 : log1p(-p);
        pos = (int64_t)lane * L / kWave - 1;
        hi = (int64_t)(lane + 1) * L / kWave;
    }
    __device__ bool next() {
        if (all) return ++pos < hi;
        const bool h = (call & 1) != 0;
        if (!h) philox_call(seed, key, sync, P_POPCTL, call >> 1, w);
        ++call;
        const uint32_t hiw = h ? w[2] : w[0], low = h ? w[3] : w[1];
        const uint64_t x = ((((uint64_t)hiw) << 32) | low) >> 11;
        const double u = ((double)x + 1.0) * (1.0 / 9007199254740992.0);  // (0,1]
        const double skip = floor(log(u) / logq);
        if (!(skip < (double)(hi - pos - 1))) { pos = hi; return false; }
        pos += 1 + (int64_t)skip;
        return true;
    }
};

__device__ __forceinline__ bool mark_alive(const uint8_t* alive, const Chunk& ch, int64_t pos) {
    return pos >= ch.count || alive[ch.first + pos] != 0;
}

__global__ __launch_bounds__(kBlock) void k_marks(const uint8_t* alive, DevWork w) {
    const int lane = lane_id();
    const int cid = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (cid >= w.nChunks) return;
    const Chunk ch = w.chunks[cid];
    const CondParams& cp = w.conds[ch.cond];
    if (!cp.active) return;
    const PopPlan pl = w.plans[ch.cond];
    if (pl.remove == 0) return;
    MarkGen g(cp, ch, lane, pl.markProb, (int64_t)ch.count + w.surplusC[cid]);
    uint32_t m = 0;
    while (g.next()) m += mark_alive(alive, ch, g.pos) ? 1u : 0u;
    m = wave_sum_u(m);
    if (lane == 0) w.markC[cid] = (int32_t)m;
}

// Key of the mark with condition-wide index j (uniform 32-bit word).
__device__ __forceinline__ uint32_t mark_key(const CondParams& cp, int64_t j) {
    uint32_t w[4];
    philox_call(cp.seed, (uint64_t)j, (uint32_t)cp.syncIndex, P_POPCTL,
                0x8000u | ((uint32_t)(cp.syncIndex >> 32) & 0x7fffu), w);
    return w[0];
}

// Scan of marks, then the removal threshold: of the K marks, the R with the
// smallest (key, index) go, which is a uniform R-subset.  The R-th smallest
// key comes from a 4-pass, 8-bit radix select; ties on that key are ordered
// by index.
constexpr int kTieCap = 1024;
__global__ __launch_bounds__(1024) void k_plan_marks(DevWork w) {
    __shared__ int64_t lds[1024];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t ties[kTieCap];
    __shared__ uint32_t sh[4];  // prefix, mask, need, tie count
    const int c = blockIdx.x;
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    PopPlan pl = w.plans[c];
    if (pl.remove == 0) return;
    const int64_t K = block_exclusive_scan(w.markC + cp.chunkBegin, w.markOff + cp.chunkBegin, cp.nChunks, lds);
    const int64_t R = pl.remove;
    if (K < R || K > 0xffffffffLL) {
        if (threadIdx.x == 0) {
            pl.marks = K;
            pl.status |= ST_POPCTL;
            w.plans[c] = pl;
            atomicOr(&w.status[c], (uint32_t)ST_POPCTL);
        }
        return;
    }
    if (threadIdx.x == 0) { sh[0] = 0; sh[1] = 0; sh[2] = (uint32_t)R; sh[3] = 0; }
    __syncthreads();
    uint32_t* keys = w.markKeys + cp.base;  // K <= the condition's pool slots
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 24 - 8 * pass;
        for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        const uint32_t prefix = sh[0], mask = sh[1];
        for (int64_t j = threadIdx.x; j < K; j += blockDim.x) {
            uint32_t k;
            if (pass == 0) { k = mark_key(cp, j); keys[j] = k; }
            else k = keys[j];
            if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x < kWave) {
            // the bucket holding the need-th smallest key: lane l scans
            // buckets [4l, 4l + 4) after an exclusive scan of the lane sums
            // (bucket 255 takes the rest, as a serial scan would)
            const int l = threadIdx.x;
            const uint32_t need = sh[2];
            uint32_t b[4], sum = 0;
            for (int q = 0; q < 4; ++q) { b[q] = hist[4 * l + q]; sum += b[q]; }
            uint32_t incl = sum;
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o, kWave);
                if (l >= o) incl += t;
            }
            uint32_t cum = incl - sum;
            int found = -1;
            uint32_t before = 0;
            for (int q = 0; q < 4; ++q) {
                const int bkt = 4 * l + q;
                if (found < 0 && (cum + b[q] >= need || bkt == 255)) { found = bkt; before = cum; }
                cum += b[q];
            }
            // the first lane with a bucket is the answer
            const uint64_t has = __ballot(found >= 0);
            const int first = __ffsll((long long)has) - 1;
            if (l == first) {
                sh[0] = prefix | ((uint32_t)found << shift);
                sh[1] = mask | (255u << shift);
                sh[2] = need - before;
            }
        }
        __syncthreads();
    }
    const uint32_t T = sh[0], need = sh[2];  // R-th smallest key; its rank among keys == T
    for (int64_t j = threadIdx.x; j < K; j += blockDim.x) {
        if (keys[j] == T) {
            const uint32_t slot = atomicAdd(&sh[3], 1u);
            if (slot < kTieCap) ties[slot] = (uint32_t)j;
        }
    }
    __syncthreads();
    const uint32_t nt = sh[3];
    if (nt > kTieCap || need < 1 || need > nt) {
        if (threadIdx.x == 0) {
            pl.marks = K;
            pl.status |= ST_POPCTL;
            w.plans[c] = pl;
            atomicOr(&w.status[c], (uint32_t)ST_POPCTL);
        }
        return;
    }
    // bitonic sort of the tied indices (padded to a power of two)
    int n2 = 1;
    while (n2 < (int)nt) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x)
        if (i >= (int)nt) ties[i] = 0xffffffffu;
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                int l = i ^ j;
                if (l > i) {
                    bool up = (i & k) == 0;
                    uint32_t a = ties[i], bb = ties[l];
                    if ((a > bb) == up) { ties[i] = bb; ties[l] = a; }
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        pl.marks = K;
        pl.thrKey = T;
        pl.thrIndex = ties[need - 1];
        w.plans[c] = pl;
    }
}

// Marks in chunk order, then lane order, then position order; a mark at or
// below the threshold is removed.  Removed main electrons become
// holes of the chunk, then the chunk refills them from its newborns.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_apply_marks(DevState<V> s, DevWork w) {
    const int lane = lane_id();
    const int cid = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (cid >= w.nChunks) return;
    const Chunk ch = w.chunks[cid];
    const CondParams& cp = w.conds[ch.cond];
    if (!cp.active) return;
    const PopPlan pl = w.plans[ch.cond];
    if (pl.remove == 0 || (pl.status & ST_POPCTL)) return;
    const int sur = w.surplusC[cid];
    const int64_t L = (int64_t)ch.count + sur;
    int mine = 0;
    {
        MarkGen g(cp, ch, lane, pl.markProb, L);
        while (g.next()) mine += mark_alive(s.alive, ch, g.pos) ? 1 : 0;
    }
    int64_t gi = w.markOff[cid] + wave_exclusive_scan(mine);
    int holes = w.holesC[cid];
    double growth = 0;
    MarkGen g(cp, ch, lane, pl.markProb, L);
    bool more = g.next();
    while (__ballot(more)) {
        bool kill = false, killMain = false;
        int64_t slot = 0;
        if (more) {
            if (mark_alive(s.alive, ch, g.pos)) {
                const uint32_t k = w.markKeys[cp.base + gi];
                if (k < pl.thrKey || (k == pl.thrKey && gi <= pl.thrIndex)) {
                    kill = true;
                    killMain = g.pos < ch.count;
                    slot = killMain ? ch.first + g.pos : ch.region + (g.pos - ch.count);
                }
                ++gi;
            }
        }
        const uint64_t km = __ballot(killMain);
        if (kill) {
            growth -= rec_energy(s.ra[slot]);
            s.alive[slot] = 0;
            if (killMain) w.holeList[ch.first + holes + lane_rank(km)] = (uint32_t)slot;
        }
        holes += __popcll(km);
        if (more) more = g.next();
    }
    int remaining = 0;
    wave_refill<V>(s, w.holeList, ch, holes, sur, remaining);
    const double gsum = wave_sum(growth);
    if (lane == 0) {
        w.holesC[cid] = holes;
        w.surplusC[cid] = remaining;
        w.tailC[(size_t)cid * CT_SIZE + CT_GROWTH] += gsum;
    }
}

__global__ __launch_bounds__(1024) void k_plan_pairs(DevWork w) {
    __shared__ int64_t lds[1024];
    const int c = blockIdx.x;
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    PopPlan pl = w.plans[c];
    if (pl.holes == 0 && pl.surplus == 0) return;
    const int64_t H = block_exclusive_scan(w.holesC + cp.chunkBegin, w.holeOff + cp.chunkBegin, cp.nChunks, lds);
    const int64_t S = block_exclusive_scan(w.surplusC + cp.chunkBegin, w.surOff + cp.chunkBegin, cp.nChunks, lds);
    if (threadIdx.x == 0) {
        pl.holes = H;
        pl.surplus = S;
        w.plans[c] = pl;
    }
}

// Pair i (i < min(H, S)): hole i of the condition <- surplus newborn i.  The
// waves of a condition's chunks stride over the pairs.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_move(DevState<V> s, DevWork w) {
    const int lane = lane_id();
    const int cid = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (cid >= w.nChunks) return;
    const Chunk me = w.chunks[cid];
    const CondParams& cp = w.conds[me.cond];
    if (!cp.active) return;
    const PopPlan pl = w.plans[me.cond];
    const int64_t n = pl.holes < pl.surplus ? pl.holes : pl.surplus;
    const int64_t* hoff = w.holeOff + cp.chunkBegin;
    const int64_t* soff = w.surOff + cp.chunkBegin;
    const int64_t stride = (int64_t)cp.nChunks * kWave;
    for (int64_t i = (int64_t)me.local * kWave + lane; i < n; i += stride) {
        const int64_t th = owner_of(hoff, cp.nChunks, i);
        const Chunk& hc = w.chunks[cp.chunkBegin + th];
        const int64_t dst = w.holeList[hc.first + (i - hoff[th])];
        const int64_t ts = owner_of(soff, cp.nChunks, i);
        const Chunk& sc = w.chunks[cp.chunkBegin + ts];
        const int64_t src = sc.region + (i - soff[ts]);
        rec_put(s, dst, rec_get(s, src));
        s.alive[src] = 0;
        s.alive[dst] = 1;
    }
}

// Holes S..H-1 get clones of uniformly drawn alive main electrons.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_clone(DevState<V> s, DevWork w) {
    const int lane = lane_id();
    const int cid = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (cid >= w.nChunks) return;
    const Chunk me = w.chunks[cid];
    const CondParams& cp = w.conds[me.cond];
    if (!cp.active) return;
    const PopPlan pl = w.plans[me.cond];
    if (pl.holes <= pl.surplus) return;
    const int64_t* hoff = w.holeOff + cp.chunkBegin;
    const int64_t stride = (int64_t)cp.nChunks * kWave;
    double growth = 0;
    uint32_t status = 0;
    for (int64_t i = pl.surplus + (int64_t)me.local * kWave + lane; i < pl.holes; i += stride) {
        const int64_t th = owner_of(hoff, cp.nChunks, i);
        const Chunk& hc = w.chunks[cp.chunkBegin + th];
        const int64_t dst = w.holeList[hc.first + (i - hoff[th])];
        // uniform alive source by rejection over the main slots (BoltzmannMC.C:1368-1371);
        // draw k of hole i: counter {i, k >> 16, interval, clone | (k & 0xffff)},
        // so no two of the 2^20 draws share a counter (i < 2^32)
        int64_t src = -1;
        for (uint32_t k = 0; k < (1u << 20); ++k) {
            double u = u53(cp.seed, (uint64_t)(uint32_t)i | ((uint64_t)(k >> 16) << 32), (uint32_t)cp.syncIndex,
                           P_CLONE, k & 0xffffu);
            int64_t cand = (int64_t)fmin(floor(u * (double)cp.nMain), (double)(cp.nMain - 1));
            if (s.alive[cp.base + cand] == 1) { src = cp.base + cand; break; }
        }
        if (src < 0) { status |= ST_POPCTL; continue; }
        Rec<V> q = rec_get(s, src);
        growth += rec_energy(q.a);
        uint32_t wds[4];
        philox_call(cp.seed ^ 0x5851F42D4C957F2Dull, draw_part(q).stream, (uint32_t)dst, P_CLONE,
                    (uint32_t)(cp.syncIndex & 0xffffu), wds);
        draw_part(q).stream = ((uint64_t)wds[1] << 32) | wds[0];
        draw_part(q).event = 0;
        rec_put(s, dst, q);
        s.alive[dst] = 2;  // alive, but not a clone source in this pass (BoltzmannMC.C:1374)
    }
    const uint32_t st = wave_or(status);
    const double g = wave_sum(growth);
    if (lane == 0) {
        if (st) atomicOr(&w.status[me.cond], st);
        w.tailC[(size_t)cid * CT_SIZE + CT_GROWTH] += g;
    }
}

// ----------------------------------------------------------------- sample
// Ensemble sums of one sampling instant (calculateMeanDataForSwarmParams,
// BoltzmannMC.C:1410-1482) in FP64, per chunk: eps, r, v, r r^T, r v^T,
// count, max eps.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_sample(DevState<V> s, DevWork w) {
    const int lane = lane_id();
    const int cid = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (cid >= w.nChunks) return;
    const Chunk ch = w.chunks[cid];
    if (!w.conds[ch.cond].active) return;
    double m[kMomentRecord];
    for (int k = 0; k < kMomentRecord; ++k) m[k] = 0;
    auto add = [&](const RecA<V>& q) {
        const double r[3] = {q.x, q.y, q.z};
        const double v[3] = {(double)q.vx, (double)q.vy, (double)q.vz};
        const double e = rec_energy(q);
        m[0] += e;
        for (int a = 0; a < 3; ++a) { m[1 + a] += r[a]; m[4 + a] += v[a]; }
        for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb) { m[7 + 3 * a + bb] += r[a] * r[bb]; m[16 + 3 * a + bb] += r[a] * v[bb]; }
        m[25] += 1;
        m[26] = fmax(m[26], e);
    };
    // two records per lane in flight; each lane still adds its records in
    // slot order.  The clone marks (alive flag 2) of this interval's
    // population control end here, read with the records: the flight kernel
    // then stores no flag for an electron that survives its interval (a
    // one-byte store per electron that shared HBM sectors with its
    // neighbours: flight kernel 0.917 -> 0.894 ms per interval on the
    // 64-condition scan, this kernel 1 us slower)
    int64_t k = lane;
    for (; k + kWave < ch.count; k += 2 * kWave) {
        const RecA<V> q0 = s.ra[ch.first + k], q1 = s.ra[ch.first + k + kWave];
        const uint8_t a0 = s.alive[ch.first + k], a1 = s.alive[ch.first + k + kWave];
        add(q0);
        add(q1);
        if (a0 == 2) s.alive[ch.first + k] = 1;
        if (a1 == 2) s.alive[ch.first + k + kWave] = 1;
    }
    if (k < ch.count) {
        add(s.ra[ch.first + k]);
        if (s.alive[ch.first + k] == 2) s.alive[ch.first + k] = 1;
    }
    for (int k = 0; k < kMomentRecord - 1; ++k) m[k] = wave_sum(m[k]);
    m[26] = wave_max(m[26]);
    if (lane == 0)
        for (int k = 0; k < kMomentRecord; ++k) w.momC[(size_t)cid * kMomentRecord + k] = m[k];
}

// Histograms of one sampling instant (getTimeDependDistributions,
// BoltzmannMC.C:1551-1571; MathFunctions::histogramCount/histogram2DCount,
// MathFunctions.C:61-81,106-127), added to integer accumulators.  One block
// per range of main slots of one condition; the EEDF is counted in LDS when
// it fits, else directly in HBM.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_hist(DevState<V> s, const SlotRange* ranges, DevHist h, int ldsCells) {
    extern __shared__ unsigned char smem[];
    const SlotRange rg = ranges[blockIdx.x];
    const int c = rg.cond;
    const double* g = h.grid + 6 * (size_t)c;
    if (g[5] != 1.0) return;  // flag 2: counted by the flight kernel as it loads the state
    const bool useLds = ldsCells >= h.nCells;
    uint32_t* lhist = (uint32_t*)smem;
    if (useLds) {
        for (int i = threadIdx.x; i < h.nCells; i += kBlock) lhist[i] = 0;
        __syncthreads();
    }
    const double eStep = g[0], cStep = g[1], rStep = g[2], aStep = g[3], maxSpeed = g[4];
    const bool cyl = h.cyl[c] != 0 && h.eadf;
    const int row = h.phaseRow[c];
    for (int64_t k = threadIdx.x; k < rg.count; k += kBlock) {
        const RecA<V> q = s.ra[rg.first + k];
        const double e = rec_energy(q);
        const int cell = (int)(e / eStep);
        if (cell >= 0 && cell < h.nCells) {
            if (useLds) atomicAdd(&lhist[cell], 1u);
            else {
                atomicAdd(&h.eedf[(size_t)c * h.nCells + cell], 1ull);
                if (row >= 0 && h.eedfPer) atomicAdd(&h.eedfPer[((size_t)c * h.nPhases + row) * h.nCells + cell], 1ull);
            }
        }
        if (cyl) {
            const double vx = q.vx, vy = q.vy, vz = q.vz;
            const double cosA = vz / sqrt(vx * vx + vy * vy + vz * vz);
            const int cy = (int)((cosA + 1.0) / cStep);
            if (cell < h.nCells && cell >= 0 && cy < h.nCos && cy >= 0)
                atomicAdd(&h.eadf[((size_t)c * h.nCells + cell) * h.nCos + cy], 1ull);
            const int rx = (int)(sqrt(vx * vx + vy * vy) / rStep);
            const int az = (int)((vz + maxSpeed) / aStep);
            if (rx < h.nRad && rx >= 0 && az < h.nAx && az >= 0)
                atomicAdd(&h.evdf[((size_t)c * h.nRad + rx) * h.nAx + az], 1ull);
        }
        if (h.evdf3) {
            // cells of width 2 v_max / n3 from -v_max on each axis; floor, so
            // that a velocity below -v_max is not counted in the first cell
            const double inv = h.n3 / (2.0 * maxSpeed);
            const double fx = floor(((double)q.vx + maxSpeed) * inv), fy = floor(((double)q.vy + maxSpeed) * inv),
                         fz = floor(((double)q.vz + maxSpeed) * inv);
            if (fx >= 0 && fx < h.n3 && fy >= 0 && fy < h.n3 && fz >= 0 && fz < h.n3)
                atomicAdd(&h.evdf3[(((size_t)c * h.n3 + (size_t)fx) * h.n3 + (size_t)fy) * h.n3 + (size_t)fz], 1ull);
        }
    }
    if (!useLds) return;
    __syncthreads();
    for (int i = threadIdx.x; i < h.nCells; i += kBlock) {
        const uint32_t n = lhist[i];
        if (!n) continue;
        atomicAdd(&h.eedf[(size_t)c * h.nCells + i], (unsigned long long)n);
        if (row >= 0 && h.eedfPer)
            atomicAdd(&h.eedfPer[((size_t)c * h.nPhases + row) * h.nCells + i], (unsigned long long)n);
    }
}

// ----------------------------------------------------------------- finish
// summary[c] = [ event record (3P + 8) | moment sums (kMomentRecord) | plan (8) ]
// Chunk records are summed in a fixed order: thread t adds chunks t, t+256,
// ... in order, the wave adds its lanes by butterflies, and thread f adds the
// four waves in order.  The integer accumulators are converted and cleared
// for the next interval.  Block 0 resets the flight-kernel work queues.
constexpr int kFinishFields = CT_SIZE + kMomentRecord;
// Fixed-order sums (maxima for the two maximum fields) over the records
// [first, first + n) of a chunk-tail array and a moment array, by one
// 256-thread block: each thread takes records threadIdx.x, + 256, ..., the
// waves reduce by shuffles and thread f < kFinishFields adds the four wave
// results of field f in wave order; the result is valid in those threads.
__device__ double finish_reduce(const double* tsrc, const double* msrc, size_t first, int n,
                                double (*red)[256 / kWave]) {
    double acc[kFinishFields];
#pragma unroll
    for (int f = 0; f < kFinishFields; ++f) acc[f] = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const size_t j = first + k;
        const double* t = tsrc + j * CT_SIZE;
        const double* m = msrc + j * kMomentRecord;
#pragma unroll
        for (int f = 0; f < CT_SIZE; ++f) acc[f] = f == CT_MAXE ? fmax(acc[f], t[f]) : acc[f] + t[f];
#pragma unroll
        for (int f = 0; f < kMomentRecord; ++f)
            acc[CT_SIZE + f] = f == kMomentRecord - 1 ? fmax(acc[CT_SIZE + f], m[f]) : acc[CT_SIZE + f] + m[f];
    }
    const int lane = lane_id(), wid = threadIdx.x / kWave;
#pragma unroll
    for (int f = 0; f < kFinishFields; ++f) {
        const bool isMax = f == CT_MAXE || f == kFinishFields - 1;
        const double v = isMax ? wave_max(acc[f]) : wave_sum(acc[f]);
        if (lane == 0) red[f][wid] = v;
    }
    __syncthreads();
    double tot = 0;
    if (threadIdx.x < kFinishFields) {
        const int f = threadIdx.x;
        const bool isMax = f == CT_MAXE || f == kFinishFields - 1;
        tot = red[f][0];
        for (int i = 1; i < 256 / kWave; ++i) tot = isMax ? fmax(tot, red[f][i]) : tot + red[f][i];
    }
    return tot;
}

// First pass for conditions with many work units: group g's sums, in the
// record layout of the chunk tails and moments.
__global__ __launch_bounds__(256) void k_finish_group(DevWork w) {
    __shared__ double red[kFinishFields][256 / kWave];
    const int g = blockIdx.x;
    if (!w.conds[w.groupCond[g]].active) return;
    const double tot = finish_reduce(w.tailC, w.momC, (size_t)w.groupFirst[g], w.groupCount[g], red);
    if (threadIdx.x < kFinishFields) {
        const int f = threadIdx.x;
        if (f < CT_SIZE) w.tailG[(size_t)g * CT_SIZE + f] = tot;
        else w.momG[(size_t)g * kMomentRecord + (f - CT_SIZE)] = tot;
    }
}

__global__ __launch_bounds__(256) void k_finish(DevWork w, int P) {
    __shared__ double red[kFinishFields][256 / kWave];
    const int c = blockIdx.x;
    if (c == 0 && threadIdx.x < 8) w.queue[threadIdx.x] = 0;
    const CondParams cp = w.conds[c];
    if (!cp.active) return;
    double* out = w.summary + (size_t)c * w.summarySize;
    const int R = w.recSize;
    // the work units themselves, or the sums of their reduction groups
    const double tot = w.nGroups ? finish_reduce(w.tailG, w.momG, (size_t)w.condGroup[c], w.condGroups[c], red)
                                 : finish_reduce(w.tailC, w.momC, (size_t)cp.chunkBegin, cp.nChunks, red);
    if (threadIdx.x < kFinishFields) {
        const int f = threadIdx.x;
        if (f < CT_SIZE) out[3 * P + f] = tot;
        else out[R + (f - CT_SIZE)] = tot;
    }
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        const size_t i = (size_t)c * P + p;
        out[p] = (double)w.accCount[i];
        out[P + p] = (double)w.accGain[i] / kFix;
        out[2 * P + p] = (double)w.accLoss[i] / kFix;
        w.accCount[i] = 0;
        w.accGain[i] = 0;
        w.accLoss[i] = 0;
    }
    if (threadIdx.x == 0) {
        const PopPlan pl = w.plans[c];
        double* tail = out + R + kMomentRecord;
        tail[0] = (double)pl.holes; tail[1] = (double)pl.surplus; tail[2] = (double)pl.alive;
        tail[3] = (double)pl.remove; tail[4] = (double)pl.marks; tail[5] = (double)pl.thrIndex;
        tail[6] = (double)w.status[c];
        tail[7] = 0;
    }
}

// ----------------------------------------------------------------- tables
// Linear interpolation as gsl_interp_linear evaluates it: the interval is the
// last one with x[j] <= v (the last point belongs to the last interval), and
// y = y_j + (v - x_j)/dx*(y_{j+1} - y_j).
__device__ __forceinline__ double table_at(const double* x, const double* y, int n, double v) {
#pragma clang fp contract(off)
    int lo = 0, hi = n - 2;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (x[mid] <= v) lo = mid; else hi = mid - 1;
    }
    const double dx = x[lo + 1] - x[lo];
    return dx > 0.0 ? y[lo] + (v - x[lo]) / dx * (y[lo + 1] - y[lo]) : 0.0;
}

// Interpolated cross section x relative density of process p at `energy`
// (interpolateCrossSections, BoltzmannMC.C:571-609), FP64.
__device__ __forceinline__ double table_value(const DevTableSrc& a, int p, double energy) {
#pragma clang fp contract(off)
    const int b = a.srcBegin[p], n = a.srcEnd[p] - b;
    double value = 0;
    if (a.isSuper[p]) {
        // Klein-Rosseland: (g_i/g_f)(1 + E_th/u) sigma_if(u + E_th)
        if (energy > a.minE[p] && energy <= a.maxE[p]) {
            const double thr = a.thrPartner[p];
            value = a.statW[p] * (1.0 + thr / energy) * table_at(a.energy + b, a.sigma + b, n, energy + thr) *
                    a.relDens[p];
        }
    } else if (energy >= a.minE[p] && energy <= a.maxE[p]) {
        value = table_at(a.energy + b, a.sigma + b, n, energy) * a.relDens[p];
    }
    return value;
}

// Rows of a table block: every (row, process) value by its own thread into
// LDS, then each row's running sum in process order by one thread -- the
// arithmetic and order of the host FP64 build (BoltzmannMC.C:571-614).
// A block holds table_rows_per_block(P) rows.
inline __host__ __device__ int table_rows_per_block(int P) { return P >= kBlock ? 1 : kBlock / P; }
template <typename V>
__global__ __launch_bounds__(kBlock) void k_table_rows(DevTableSrc a, double N, double step, V* tables,
                                                       double* nutot64) {
#pragma clang fp contract(off)
    extern __shared__ double vals[];  // [rows * P]
    const int P = a.P, I = a.I;
    const int rows = table_rows_per_block(P);
    const int i0 = blockIdx.x * rows;
    V* cumul = tables;
    V* interp = tables + (size_t)I * P;
    V* nutot = interp + (size_t)I * P;
    for (int q = threadIdx.x; q < rows * P; q += kBlock) {
        const int i = i0 + q / P;
        if (i >= I) break;
        const double v = table_value(a, q % P, i * step);
        vals[q] = v;
        interp[(size_t)i0 * P + q] = (V)v;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < rows; r += kBlock) {
        const int i = i0 + r;
        if (i >= I) break;
        double acc = 0;
        for (int p = 0; p < P; ++p) {
            acc += vals[r * P + p];
            cumul[(size_t)i * P + p] = (V)acc;
        }
        acc *= N * sqrt(i * step * 2.0 * kCharge / kMass);
        nutot[i] = (V)acc;
        nutot64[i] = acc;
    }
}

// Running maximum of the total collision frequency (BoltzmannMC.C:612-613):
// one block; each thread takes a contiguous segment.
__global__ __launch_bounds__(1024) void k_table_max(const double* nutot64, double* numax64, int I) {
    __shared__ double lds[1024];
    const int per = (I + 1023) / 1024;
    const int b = threadIdx.x * per, e = min(I, b + per);
    double m = 0;
    for (int i = b; i < e; ++i) m = fmax(m, nutot64[i]);
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // inclusive max-scan of the segment maxima
        const double v = threadIdx.x >= o ? lds[threadIdx.x - o] : 0.0;
        __syncthreads();
        lds[threadIdx.x] = fmax(lds[threadIdx.x], v);
        __syncthreads();
    }
    double run = threadIdx.x > 0 ? lds[threadIdx.x - 1] : 0.0;
    for (int i = b; i < e; ++i) {
        run = fmax(nutot64[i], run);
        numax64[i] = run;
    }
}

// Null-test envelope: for bin b (rows b<<s .. (b+1)<<s, the last row being
// the partner of the bin's last interpolation), the largest and smallest
// total frequency, widened by a few ulp of V so that every interpolated value
// w1*a + w2*b of the bin's rows lies inside.
template <typename V> struct EnvMargin;
template <> struct EnvMargin<float> { static constexpr double up = 1.0 + 0x1p-20, down = 1.0 - 0x1p-20; };
template <> struct EnvMargin<double> { static constexpr double up = 1.0 + 0x1p-48, down = 1.0 - 0x1p-48; };

template <typename V>
__global__ __launch_bounds__(kEnvBins) void k_table_env(V* tables, int I, int P, int shift) {
    const V* nutot = tables + 2 * (size_t)I * P;
    V* envHi = tables + 2 * (size_t)I * P + I;
    V* envLo = envHi + kEnvBins;
    const int b = threadIdx.x;
    const int k0 = b << shift;
    if (k0 > I - 1) { envHi[b] = (V)0; envLo[b] = (V)0; return; }
    const int k1 = min(((b + 1) << shift), I - 1);
    V hi = nutot[k0], lo = nutot[k0];
    for (int k = k0 + 1; k <= k1; ++k) { hi = fmax(hi, nutot[k]); lo = fmin(lo, nutot[k]); }
    envHi[b] = (V)((double)hi * EnvMargin<V>::up);
    envLo[b] = (V)((double)lo * EnvMargin<V>::down);
}

template <typename V>
hipError_t launch_tables(const DevTableSrc& src, double N, double step, int envShift, V* tables, double* nutot64,
                         double* numax64, hipStream_t st) {
    const int rows = table_rows_per_block(src.P);
    hipLaunchKernelGGL(k_table_rows<V>, dim3((src.I + rows - 1) / rows), dim3(kBlock),
                       (size_t)rows * src.P * sizeof(double), st, src, N, step, tables, nutot64);
    hipLaunchKernelGGL(k_table_env<V>, dim3(1), dim3(kEnvBins), 0, st, tables, src.I, src.P, envShift);
    hipLaunchKernelGGL(k_table_max, dim3(1), dim3(1024), 0, st, nutot64, numax64, src.I);
    return hipGetLastError();
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
static inline unsigned chunk_grid(int nChunks) { return (unsigned)((nChunks + kWavesPerBlock - 1) / kWavesPerBlock); }
// Dynamic LDS of a flight-kernel block: the process records [P], then the
// per-wave areas.
template <typename V> static inline size_t flight_lds(int P, int histLds) {
    return (size_t)P * sizeof(ProcRec) + (size_t)kWavesPerBlock * flight_wave_lds<V>(P, histLds);
}

template <typename V>
hipError_t launch_init(const DevState<V>& s, const DevWork& w, int64_t totalMain, const int64_t* mainBase,
                       const double* thermalDev, hipStream_t st) {
    if (totalMain == 0) return hipSuccess;
    timer_start(K_INIT, st);
    hipLaunchKernelGGL(k_init<V>, dim3(grid_of(totalMain)), dim3(kBlock), 0, st, s, w.conds, w.nCond, totalMain,
                       mainBase, thermalDev);
    timer_stop(K_INIT, st);
    return hipGetLastError();
}

// Histogram mode of the flight launches of an engine: the EEDF in LDS when
// the engine counts it there, EADF/EVDF code only when a condition is
// cylindrically symmetric.
static inline int hist_mode(const DevHist& h) {
    if (h.histLds < h.nCells) return HIST_HBM;
    return h.eadf ? HIST_LDS_CYL : HIST_LDS;
}

template <typename V, int FC, bool TH>
static const void* flight_kernel(int hist) {
    if (hist == HIST_LDS) return (const void*)k_advance<V, FC, TH, HIST_LDS>;
    if (hist == HIST_LDS_CYL) return (const void*)k_advance<V, FC, TH, HIST_LDS_CYL>;
    return (const void*)k_advance<V, FC, TH, HIST_HBM>;
}

template <typename V, int FC, bool TH, int HIST>
static void launch_advance_hist(const DevState<V>& s, const DevWork& w, const DevProcesses& pr, const V* tables,
                                const DevHist& h, const FlightLists& fl, hipStream_t st) {
    dim3 g(fl.grid[FC]), b(kBlock);
    const size_t lds = flight_lds<V>(pr.P, h.histLds);
    hipLaunchKernelGGL((k_advance<V, FC, TH, HIST>), g, b, lds, st, s, w, pr, tables, h, fl.list[FC], fl.count[FC],
                       w.queue + FC);
}

template <typename V, int FC, bool TH>
static void launch_advance_th(const DevState<V>& s, const DevWork& w, const DevProcesses& pr, const V* tables,
                              const DevHist& h, const FlightLists& fl, hipStream_t st) {
    switch (hist_mode(h)) {
        case HIST_LDS: launch_advance_hist<V, FC, TH, HIST_LDS>(s, w, pr, tables, h, fl, st); break;
        case HIST_LDS_CYL: launch_advance_hist<V, FC, TH, HIST_LDS_CYL>(s, w, pr, tables, h, fl, st); break;
        default: launch_advance_hist<V, FC, TH, HIST_HBM>(s, w, pr, tables, h, fl, st); break;
    }
}

template <typename V, int FC>
static void launch_advance_case(const DevState<V>& s, const DevWork& w, const DevProcesses& pr, const V* tables,
                                const DevHist& h, const FlightLists& fl, hipStream_t st) {
    if (fl.count[FC] == 0) return;
    if (pr.gasTempEffect == 0) launch_advance_th<V, FC, false>(s, w, pr, tables, h, fl, st);
    else launch_advance_th<V, FC, true>(s, w, pr, tables, h, fl, st);
}

template <typename V>
int flight_blocks_per_cu(int fc, bool thermal, int P, const DevHist& h) {
    const void* k = nullptr;
    const int hm = hist_mode(h);
    switch (fc) {
        case FIELD_DC: k = thermal ? flight_kernel<V, FIELD_DC, true>(hm) : flight_kernel<V, FIELD_DC, false>(hm); break;
        case FIELD_AC: k = thermal ? flight_kernel<V, FIELD_AC, true>(hm) : flight_kernel<V, FIELD_AC, false>(hm); break;
        case FIELD_DC_B:
            k = thermal ? flight_kernel<V, FIELD_DC_B, true>(hm) : flight_kernel<V, FIELD_DC_B, false>(hm);
            break;
        case FIELD_ECR:
            k = thermal ? flight_kernel<V, FIELD_ECR, true>(hm) : flight_kernel<V, FIELD_ECR, false>(hm);
            break;
        case FLIGHT_DC_ANY:
            k = thermal ? flight_kernel<V, FLIGHT_DC_ANY, true>(hm) : flight_kernel<V, FLIGHT_DC_ANY, false>(hm);
            break;
        default: k = thermal ? flight_kernel<V, FIELD_AC_B, true>(hm) : flight_kernel<V, FIELD_AC_B, false>(hm); break;
    }
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kBlock, flight_lds<V>(P, h.histLds)) != hipSuccess || n < 1)
        n = 1;
    return n;
}

template <typename V>
hipError_t launch_flights(const DevState<V>& s, const DevWork& w, const DevProcesses& pr, const V* tables,
                          const DevHist& h, const FlightLists& fl, hipStream_t st) {
    timer_start(K_ADVANCE, st);
    launch_advance_case<V, FIELD_DC>(s, w, pr, tables, h, fl, st);
    launch_advance_case<V, FIELD_AC>(s, w, pr, tables, h, fl, st);
    launch_advance_case<V, FIELD_DC_B>(s, w, pr, tables, h, fl, st);
    launch_advance_case<V, FIELD_ECR>(s, w, pr, tables, h, fl, st);
    launch_advance_case<V, FIELD_AC_B>(s, w, pr, tables, h, fl, st);
    launch_advance_case<V, FLIGHT_DC_ANY>(s, w, pr, tables, h, fl, st);
    timer_stop(K_ADVANCE, st);
    return hipGetLastError();
}

template <typename V>
hipError_t launch_popctl(const DevState<V>& s, const DevWork& w, hipStream_t st) {
    const dim3 g(chunk_grid(w.nChunks)), b(kBlock);
    timer_start(K_PLAN, st);
    hipLaunchKernelGGL(k_plan, dim3(w.nCond), dim3(1024), 0, st, w);
    timer_stop(K_PLAN, st);
    timer_start(K_MARKS, st);
    hipLaunchKernelGGL(k_marks, g, b, 0, st, s.alive, w);
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
    hipLaunchKernelGGL(k_sample<V>, dim3(chunk_grid(w.nChunks)), dim3(kBlock), 0, st, s, w);
    timer_stop(K_SAMPLE, st);
    return hipGetLastError();
}

template <typename V>
hipError_t launch_hist(const DevState<V>& s, const DevHist& h, hipStream_t st) {
    if (h.nRanges == 0) return hipSuccess;
    const int ldsCells = h.nCells <= 8192 ? h.nCells : 0;
    const size_t lds = (size_t)ldsCells * 4;
    timer_start(K_HIST, st);
    hipLaunchKernelGGL(k_hist<V>, dim3(h.nRanges), dim3(kBlock), lds, st, s, h.ranges, h, ldsCells);
    timer_stop(K_HIST, st);
    return hipGetLastError();
}

hipError_t launch_finish(const DevWork& w, int P, hipStream_t st) {
    timer_start(K_REDUCE, st);
    if (w.nGroups) hipLaunchKernelGGL(k_finish_group, dim3(w.nGroups), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_finish, dim3(w.nCond), dim3(256), 0, st, w, P);
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
                                       const double*, hipStream_t);                                           \
    template hipError_t launch_flights<V>(const DevState<V>&, const DevWork&, const DevProcesses&, const V*,  \
                                          const DevHist&, const FlightLists&, hipStream_t);                   \
    template hipError_t launch_popctl<V>(const DevState<V>&, const DevWork&, hipStream_t);                    \
    template hipError_t launch_sample<V>(const DevState<V>&, const DevWork&, hipStream_t);                    \
    template hipError_t launch_hist<V>(const DevState<V>&, const DevHist&, hipStream_t);                     \
    template int flight_blocks_per_cu<V>(int, bool, int, const DevHist&);                                                \
    template hipError_t launch_tables<V>(const DevTableSrc&, double, double, int, V*, double*, double*,        \
                                         hipStream_t);                                                       \
    template hipError_t launch_op_accelerate<V>(const CondParams*, int64_t, const double*, const double*,     \
                                                double*, double*, double*, double*, hipStream_t);             \
    template hipError_t launch_op_collide<V>(const CondParams*, const DevProcesses&, const V*, int64_t,       \
                                             double*, double*, const double*, const double*, int32_t*,        \
                                             double*, double*, double*, hipStream_t);
EKMC_INSTANTIATE(double)
EKMC_INSTANTIATE(float)

}  // namespace ekmc
