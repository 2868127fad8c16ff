// device_physics.h -- per-electron device arithmetic of the Monte Carlo path:
// Philox variates, analytic free flights, null-collision channel choice and
// the three collision operators.  One template serves both arithmetic forms:
//   V = double : plain FP64, the drop-in form; every step follows the
//                reference's operation order (BoltzmannMC.C:804-1280) so that
//                channel indices agree bit for bit with a CPU FP64 build.
//   V = float  : mixed form; velocities, energies, flight and scattering in
//                FP32, tables in FP32, while the clock, positions and the
//                energy left after a threshold is subtracted stay in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include "common.h"

namespace ekmc {

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon et al., SC'11): ten rounds, Weyl key bumps between.
// Each round's two 32x32-bit products are formed as full 64-bit products.
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

__device__ __forceinline__ void philox_call(uint64_t seed, uint64_t stream, uint32_t event,
                                            uint32_t purpose, uint32_t call, uint32_t w[4]) {
    w[0] = (uint32_t)stream; w[1] = (uint32_t)(stream >> 32); w[2] = event;
    w[3] = (purpose << 16) | call;
    philox10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Stream of an electron ejected by `parent` at its event `ev`.
__device__ __forceinline__ uint64_t child_stream(uint64_t seed, uint64_t parent, uint32_t ev) {
    uint32_t w[4];
    philox_call(seed, parent, ev, P_CHILD, 0, w);
    return ((uint64_t)w[1] << 32) | w[0];
}

// Slot intervals (ekmc.h): 0 (0,1); 1-5,7,10 (0,1]; 6,8,9 [0,1].
__device__ __forceinline__ constexpr int slot_kind(int s) {
    return s == 0 ? 0 : ((s == 6 || s == 8 || s == 9) ? 2 : 1);
}

// Slot -> Philox call of the event: call 0 carries the free time and the
// channel-choice variate, so a null event costs one call; a cold collision
// draws cos(chi) and the azimuth from call 1; thermal and ionization slots
// come from later calls.
//   FP64 (one slot = 53 bits of two words, two slots per call):
//     call 0 {0,5}  1 {6,7}  2 {1,2}  3 {3,4}  4 {8,9}  5 {10}
//   FP32: the same calls; a slot is the top 24 bits of the first of its two
//     words, i.e. the 24 leading bits of the FP64 variate.  With the same
//     seed the two forms then draw the same variates, and their trajectories
//     stay together until a rounding difference flips a decision.  (Packing
//     four FP32 slots into call 0 saved the second call of a cold collision,
//     about 1 % of the flight kernel, but left the two forms uncorrelated.)
__device__ __forceinline__ constexpr int call64(int s) {
    return (s == 0 || s == 5) ? 0 : (s == 6 || s == 7) ? 1 : (s <= 2) ? 2 : (s <= 4) ? 3 : (s <= 9) ? 4 : 5;
}
__device__ __forceinline__ constexpr int half64(int s) {
    return (s == 5 || s == 7 || s == 2 || s == 4 || s == 9) ? 1 : 0;
}

// Variates of one event: call 0 cached in w0, one other call in w1.  With a
// tape (batch operators) the slot values are read from it instead.
template <typename V> struct EventRng {
    uint64_t seed, stream;
    uint32_t event;
    bool have0;
    int other;
    uint32_t w0[4], w1[4];
    const double* tape;
    __device__ EventRng(uint64_t sd, uint64_t st, uint32_t ev, const double* tp = nullptr)
        : seed(sd), stream(st), event(ev), have0(false), other(-1), tape(tp) {}
    // with call 0 already drawn
    __device__ EventRng(uint64_t sd, uint64_t st, uint32_t ev, const uint32_t c0[4])
        : seed(sd), stream(st), event(ev), have0(true), other(-1), tape(nullptr) {
        w0[0] = c0[0]; w0[1] = c0[1]; w0[2] = c0[2]; w0[3] = c0[3];
    }
    __device__ __forceinline__ void prime0() {
        if (!have0) { philox_call(seed, stream, event, P_EVENT, 0u, w0); have0 = true; }
    }
    __device__ __forceinline__ const uint32_t* words(int call) {
        if (call == 0) { prime0(); return w0; }
        if (other != call) { philox_call(seed, stream, event, P_EVENT, (uint32_t)call, w1); other = call; }
        return w1;
    }
    __device__ __forceinline__ V u(int s);
};

template <> __device__ __forceinline__ double EventRng<double>::u(int s) {
    if (tape) return tape[s];
    const uint32_t* w = words(call64(s));
    const int h = half64(s);
    const uint32_t a = h ? w[2] : w[0], b = h ? w[3] : w[1];
    const uint64_t x = ((((uint64_t)a) << 32) | b) >> 11;
    const double k = 1.0 / 9007199254740992.0;
    const int kind = slot_kind(s);
    return kind == 0 ? ((double)x + 0.5) * k : (kind == 1 ? ((double)x + 1.0) * k : (double)x * k);
}

template <> __device__ __forceinline__ float EventRng<float>::u(int s) {
    if (tape) return (float)tape[s];
    const uint32_t* w = words(call64(s));
    const uint32_t v = half64(s) ? w[2] : w[0];
    const float x = (float)(v >> 8);
    const float k = 1.0f / 16777216.0f;
    const int kind = slot_kind(s);
    return kind == 0 ? (x + 0.5f) * k : (kind == 1 ? (x + 1.0f) * k : x * k);
}

// Free time -ln(U)/nu_trial (BoltzmannMC.C:652-653): FP64 in the reference
// order; FP32 with the hardware logarithm and the reciprocal frequency.
__device__ __forceinline__ double free_time(double u, const CondParams& c) { return -log(u) / c.trial; }
__device__ __forceinline__ float free_time(float u, const CondParams& c) {
    return -__logf(u) * (float)c.trialInv;
}

// sin and cos of 2 pi u (azimuths): FP64 as the reference computes them;
// FP32 with the hardware sine/cosine, whose argument is in turns.
__device__ __forceinline__ void turn_sincos(double u, double& s, double& co) {
    const double a = (2.0 * kPi) * u;
    s = sin(a); co = cos(a);
}
__device__ __forceinline__ void turn_sincos(float u, float& s, float& co) {
    s = __builtin_amdgcn_sinf(u); co = __builtin_amdgcn_cosf(u);
}
__device__ __forceinline__ double log_of(double x) { return log(x); }
__device__ __forceinline__ float log_of(float x) { return __logf(x); }

// ---------------------------------------------------------------- electron
template <typename V> struct Electron {
    double x, y, z, t;
    V vx, vy, vz, eps, tfree, nue;
    uint64_t stream;
    uint32_t event;
};

template <typename V> __device__ __forceinline__ V sqr_norm(V a, V b, V c) {
#pragma clang fp contract(off)
    return a * a + b * b + c * c;
}

// electron energy [eV] of a velocity, 0.5*m_e*|v|^2/e in the reference order
__device__ __forceinline__ double energy_of(double vx, double vy, double vz) {
#pragma clang fp contract(off)
    return 0.5 * kMass * sqr_norm(vx, vy, vz) / kCharge;
}
__device__ __forceinline__ float energy_of(float vx, float vy, float vz) {
#pragma clang fp contract(off)
    return (float)(0.5 * kMass / kCharge) * (vx * vx + vy * vy + vz * vz);
}
// FP64 energy of an FP32 velocity (threshold subtraction in the mixed form)
__device__ __forceinline__ double energy_of_d(float vx, float vy, float vz) {
    double a = vx, b = vy, c = vz;
    return 0.5 * kMass * (a * a + b * b + c * c) / kCharge;
}
__device__ __forceinline__ double energy_of_d(double vx, double vy, double vz) { return energy_of(vx, vy, vz); }

// sin(x) - x and cos(x) - 1 without cancellation, for the FP32 flights
__device__ __forceinline__ float sin_minus_x(float x) {
    if (fabsf(x) < 0.5f) {
        float x2 = x * x;
        return -x * x2 * (1.0f / 6.0f) * (1.0f - x2 * (1.0f / 20.0f) * (1.0f - x2 * (1.0f / 42.0f) * (1.0f - x2 * (1.0f / 72.0f))));
    }
    return sinf(x) - x;
}
__device__ __forceinline__ float cos_minus_1(float x) {
    float s = sinf(0.5f * x);
    return -2.0f * s * s;
}

// --------------------------------------------------------------- flights
// accelerateElectron (BoltzmannMC.C:804-905), FP64 form in the reference order.
template <int FC>
__device__ inline void flight(const CondParams& c, Electron<double>& e, double dt, double& gain) {
    double prev = e.eps;
    const double q = kCharge, me = kMass;
    double v0 = e.vx, v1 = e.vy, v2 = e.vz;
    const int fc = FC >= 0 ? FC : c.fieldCase;
    switch (fc) {
        case FIELD_DC: {
            e.x += v0 * dt + c.accel[0] * (0.5 * dt * dt);
            e.y += v1 * dt + c.accel[1] * (0.5 * dt * dt);
            e.z += v2 * dt + c.accel[2] * (0.5 * dt * dt);
            e.vx = v0 + c.accel[0] * dt;
            e.vy = v1 + c.accel[1] * dt;
            e.vz = v2 + c.accel[2] * dt;
        } break;
        case FIELD_AC: {
            double w = c.omega, phi = w * e.t;
            double sinPhi = sin(phi), cosPhi = cos(phi);
            double wDt = w * dt, phase = wDt + phi;
            double sinPhase = sin(phase), cosPhase = cos(phase);
            double e_me_w = q / (me * w);
            double aux1 = e_me_w / w * (cosPhase + wDt * sinPhi - cosPhi);
            double aux2 = e_me_w * (sinPhi - sinPhase);
            e.x += v0 * dt + c.E[0] * aux1;
            e.y += v1 * dt;
            e.z += v2 * dt + c.E[2] * aux1;
            e.vx = v0 + c.E[0] * aux2;
            e.vz = v2 + c.E[2] * aux2;
        } break;
        case FIELD_DC_B: {
            double W = c.cyclotron, WDt = W * dt;
            double sW = sin(WDt), cW = cos(WDt);
            double sW_W = sW / W, aux2 = (cW - 1.0) / W;
            double vEx = q * c.E[0] / (me * W), az = q * c.E[2] / me;
            e.x += v0 * sW_W + (v1 + vEx) * aux2;
            e.y += -v0 * aux2 + v1 * sW_W + vEx * (sW_W - dt);
            e.z += v2 * dt - 0.5 * az * dt * dt;
            e.vx = v0 * cW - v1 * sW - vEx * sW;
            e.vy = v0 * sW + v1 * cW + vEx * (cW - 1.0);
            e.vz = v2 - az * dt;
        } break;
        case FIELD_ECR: {
            double W = c.cyclotron, phi = W * e.t;
            double sinPhi = sin(phi), cosPhi = cos(phi);
            double WDt = W * dt, sW = sin(WDt), cW = cos(WDt);
            double phase = WDt + phi, sinPhase = sin(phase), cosPhase = cos(phase);
            double cosOpp = cos(phi - WDt);
            double e_me_W = q / (me * W);
            double vEx = e_me_W * c.E[0], vEz = e_me_W * c.E[2];
            double sW_W = sW / W, cWm1_W = (cW - 1.0) / W, cosPhase_W = cosPhase / W;
            e.x += v0 * sW_W + v1 * cWm1_W - 0.25 * vEx * (cosPhase_W - cosOpp / W + 2.0 * dt * sinPhase);
            e.y += -v0 * cWm1_W + v1 * sW_W + 0.5 * vEx * (dt * cosPhase_W - 2.0 * cWm1_W * sinPhi - cosPhi * sW_W);
            e.z += v2 * dt + vEz * (cosPhase_W - cosPhi / W + dt * sinPhi);
            e.vx = v0 * cW - v1 * sW - 0.5 * vEx * (WDt * cosPhase + cosPhi * sW);
            e.vy = v0 * sW + v1 * cW - 0.5 * vEx * (WDt * cW * sinPhi + (WDt * cosPhi - sinPhi) * sW);
            e.vz = v2 + vEz * (sinPhi - sinPhase);
        } break;
        default: {
            double w = c.omega, W = c.cyclotron, phi = w * e.t;
            double sinPhi = sin(phi), cosPhi = cos(phi);
            double wDt = w * dt, swD = sin(wDt), cwD = cos(wDt);
            double phase = wDt + phi, sinPhase = sin(phase), cosPhase = cos(phase);
            double WDt = W * dt, sW = sin(WDt), cW = cos(WDt);
            double w2 = w * w, W2 = W * W, d = w2 - W2;
            double vEx = q / (me * W) * c.E[0], vEz = q / (me * w) * c.E[2];
            double vEx_d = vEx / d, WvEx_d = W * vEx_d;
            double sW_W = sW / W, cWm1_W = (cW - 1.0) / W, wSinPhi = w * sinPhi;
            e.x += v0 * sW_W + v1 * cWm1_W + WvEx_d * (cosPhi * (cwD - cW) + (wSinPhi * sW_W - sinPhi * swD));
            e.y += -v0 * cWm1_W + v1 * sW_W - vEx_d / w * ((W2 + w2 * cW - w2) * sinPhi + W2 * (w * cosPhi * sW_W - sinPhase));
            e.z += v2 * dt + vEz * ((cosPhase - cosPhi) / w + dt * sinPhi);
            e.vx = v0 * cW - v1 * sW + WvEx_d * w * (cW * sinPhi - sinPhase + cosPhi / w * W * sW);
            e.vy = v0 * sW + v1 * cW + vEx_d * W2 * (cosPhase - cosPhi * cW + wSinPhi * sW_W);
            e.vz = v2 + vEz * (sinPhi - sinPhase);
        } break;
    }
    e.eps = energy_of(e.vx, e.vy, e.vz);
    gain = e.eps - prev;
}

// Mixed form: FP32 velocities and displacements, FP64 clock and positions.
// The DC, AC and DC+B cases use cancellation-free rewrites of the same
// closed forms with per-condition FP32 constants; the two AC+B cases run the
// FP64 formulas and round.
template <int FC>
__device__ inline void flight(const CondParams& c, Electron<float>& e, double dt, double& gain) {
    float prev = e.eps;
    float v0 = e.vx, v1 = e.vy, v2 = e.vz;
    float dtf = (float)dt;
    const int fc = FC >= 0 ? FC : c.fieldCase;
    switch (fc) {
        case FIELD_DC: {
            float hdt2 = 0.5f * dtf * dtf;
            e.x += (double)(v0 * dtf + c.fA[0] * hdt2);
            e.y += (double)(v1 * dtf + c.fA[1] * hdt2);
            e.z += (double)(v2 * dtf + c.fA[2] * hdt2);
            e.vx = v0 + c.fA[0] * dtf;
            e.vy = v1 + c.fA[1] * dtf;
            e.vz = v2 + c.fA[2] * dtf;
        } break;
        case FIELD_AC: {
            // phase in turns from the FP64 clock (fraction of f t), then FP32
            double ft = c.freq * e.t;
            ft -= floor(ft);
            const float phT = (float)ft, dT = (float)(c.freq * dt);
            const float d = (float)(c.omega * dt);
            const float sinPhi = __builtin_amdgcn_sinf(phT), cosPhi = __builtin_amdgcn_cosf(phT);
            const float sh = sinf(0.5f * d);
            const float dS = 2.0f * __builtin_amdgcn_cosf(phT + 0.5f * dT) * sh;  // sin(phi+d) - sin(phi)
            const float dmsin = -sin_minus_x(d);                                   // d - sin(d)
            const float C1pd = sinPhi * dmsin - 2.0f * cosPhi * sh * sh;           // cos(phi+d)-cos(phi)+d sin(phi)
            const float aux1 = c.fEmw * c.fInvOmega * C1pd;
            const float aux2 = -c.fEmw * dS;
            e.x += (double)(v0 * dtf + c.fE0 * aux1);
            e.y += (double)(v1 * dtf);
            e.z += (double)(v2 * dtf + c.fE2 * aux1);
            e.vx = v0 + c.fE0 * aux2;
            e.vz = v2 + c.fE2 * aux2;
        } break;
        case FIELD_DC_B: {
            // rotation by x = W dt from sin and cos of x/2:
            // sin x = 2 s c, cos x - 1 = -2 s^2, sin x - x by series for small x
            const float x = (float)(c.cyclotron * dt);
            float sh, ch;
            sincosf(0.5f * x, &sh, &ch);
            const float sW = 2.0f * sh * ch, cm1 = -2.0f * sh * sh;
            const float smx = fabsf(x) < 0.5f ? sin_minus_x(x) : sW - x;
            const float vEx = c.fvEx, iW = c.fInvW;
            e.x += (double)((v0 * sW + (v1 + vEx) * cm1) * iW);
            e.y += (double)((-v0 * cm1 + v1 * sW + vEx * smx) * iW);
            e.z += (double)(v2 * dtf - 0.5f * c.fAz * dtf * dtf);
            e.vx = v0 + v0 * cm1 - (v1 + vEx) * sW;
            e.vy = v0 * sW + v1 + (v1 + vEx) * cm1;
            e.vz = v2 - c.fAz * dtf;
        } break;
        default: {
            Electron<double> d;
            d.x = e.x; d.y = e.y; d.z = e.z; d.t = e.t;
            d.vx = v0; d.vy = v1; d.vz = v2; d.eps = prev;
            double g;
            flight<FC>(c, d, dt, g);
            e.x = d.x; e.y = d.y; e.z = d.z;
            e.vx = (float)d.vx; e.vy = (float)d.vy; e.vz = (float)d.vz;
        } break;
    }
    e.eps = energy_of(e.vx, e.vy, e.vz);
    gain = (double)e.eps - (double)prev;
}

// --------------------------------------------------------------- samplers
// AngularScatteringFunctions.h:27-74; `slot` is the variate slot.
template <typename V>
__device__ __forceinline__ V cos_chi(const DevProcesses& pr, int p, V energy, V energyAfter, EventRng<V>& rng,
                                    int slot) {
    switch (pr.proc[p].scat) {
        case 1: return (V)1;  // forward
        case 2: {             // bornDipole
            V loss = (V)pr.proc[p].energyLoss;
            V s = sqrt(energyAfter) + sqrt(energy);
            V ratio = loss / (s * s);
            V r2 = ratio * ratio;
            return (V)1 + (V)2 * r2 / ((V)1 - r2) * ((V)1 - pow(r2, -rng.u(slot)));
        }
        case 3:  // surendra
            return ((V)2 + energy - (V)2 * pow((V)1 + energy, rng.u(slot))) / energy;
        case 4: {  // coulombScreen
            V eps = pr.proc[p].params[0] == 0 ? energy : energyAfter;
            V sp = (V)pr.proc[p].params[1] / eps;
            V R = rng.u(slot);
            return (sp + (V)1 - ((V)2 * sp + (V)1) * R) / (sp + (V)1 - R);
        }
        default: return (V)1 - (V)2 * rng.u(slot);  // isotropic
    }
}

template <typename V> struct Sph { V norm, sinT, cosT, sinP, cosP; };

// MathFunctions::cart2sph (MathFunctions.C:143-163)
template <typename V> __device__ inline bool cart2sph(V x, V y, V z, Sph<V>& s) {
    V x2My2 = x * x + y * y;
    V nxy = sqrt(x2My2);
    s.norm = sqrt(x2My2 + z * z);
    if (s.norm == (V)0) return false;
    s.sinT = nxy / s.norm;
    s.cosT = z / s.norm;
    if (nxy != (V)0) { s.sinP = y / nxy; s.cosP = x / nxy; }
    else { s.sinP = (V)1; s.cosP = (V)0; }
    return true;
}

// MathFunctions::eulerTransformation (MathFunctions.C:129-141)
template <typename V>
__device__ __forceinline__ void euler(V sinChi, V cosChi, V sinEta, V cosEta, const Sph<V>& s, V out[3]) {
    V sCsE = sinChi * sinEta;
    V aux = sinChi * cosEta * s.cosT + cosChi * s.sinT;
    out[0] = -sCsE * s.sinP + aux * s.cosP;
    out[1] = sCsE * s.cosP + aux * s.sinP;
    out[2] = -sinChi * cosEta * s.sinT + cosChi * s.cosT;
}

template <typename V> struct CollOut {
    int pid;
    uint32_t flags;  // ST_ZERO_VELOCITY
    V ejv[3];
    V ejEps;
    double dE;
    bool escape;
};

// Interpolated table value of two rows (no contraction: bit-exact FP64 form)
template <typename V> __device__ __forceinline__ V lerp2(V w1, V a, V w2, V b) {
#pragma clang fp contract(off)
    return w1 * a + w2 * b;
}

// Lower-bound search of BoltzmannMC.C:1061-1088 (and the per-gas variant,
// :986-1010) followed by the zero-rate walk (:1091-1093).
template <typename V>
__device__ inline int channel_search(const V* c1, const V* c2, const V* i1, const V* i2, V w1, V w2,
                                     int L, int R, V target, V prevSum, V ref, V speed, bool perGas) {
#pragma clang fp contract(off)
    int chosen = -1;
    bool broke = false;
    auto probe = [&](int t) -> V {
        return perGas ? prevSum + (w1 * c1[t] + w2 * c2[t] - ref) * speed : w1 * c1[t] + w2 * c2[t];
    };
    // The probes of the sequential loop (t = (L + R) / 2, then the left or
    // right half), two levels per round trip: the middles of both halves are
    // loaded with the current middle, so a search makes half the dependent
    // loads.  Every speculative index lies in [L, R].  (Three levels would
    // raise the flight kernel above the registers of three waves per SIMD.)
    while (L != R) {
        const int t0 = (L + R) / 2;
        const int a1 = (L + t0) / 2, b1 = (t0 + 1 + R) / 2;
        const V v0 = probe(t0), va1 = probe(a1), vb1 = probe(b1);
        if (target < v0) R = t0;
        else if (target > v0) L = t0 + 1;
        else { chosen = t0; broke = true; break; }
        if (L == R) break;
        const bool left0 = R == t0;  // range [L, t0], middle a1; else [t0 + 1, R], middle b1
        const int t1 = left0 ? a1 : b1;
        const V v1 = left0 ? va1 : vb1;
        if (target < v1) R = t1;
        else if (target > v1) L = t1 + 1;
        else { chosen = t1; broke = true; break; }
    }
    if (!broke) chosen = L;
    while (chosen > 0 && w1 * i1[chosen] == (V)0 && w2 * i2[chosen] == (V)0) --chosen;
    return chosen;
}

// performCollision (BoltzmannMC.C:907-1113) with conservativeCollision
// (:1115-1193), ionizationCollision (:1195-1272) and attachmentCollision
// (:1274-1280), in parts: the cold null test, the cold and thermal channel
// choices, and the collision operator of the chosen process.  `cumul`,
// `interp` [I*P] and `nutot` [I] are this condition's tables.

// Thermal target motion for this electron (BoltzmannMC.C:916)
template <typename V>
__device__ __forceinline__ bool is_thermal(const CondParams& c, const DevProcesses& pr, const Electron<V>& e) {
    const int gte = pr.gasTempEffect;
    return gte == 1 || (gte == 2 && (double)e.eps < 20.0 * c.gasEnergy);
}

// Table rows and weights of the cold branch (BoltzmannMC.C:1035-1047)
template <typename V> struct ColdRow {
    int k1, k2;
    V w1, w2;
};
template <typename V> __device__ __forceinline__ ColdRow<V> cold_row(const CondParams& c, const Electron<V>& e) {
    ColdRow<V> r;
    const V over = e.eps / (V)c.step;
    r.k1 = (int)fmin(over, (V)(c.I - 1));
    r.k2 = (int)fmin((V)(r.k1 + 1), (V)(c.I - 1));
    r.w1 = (V)r.k2 - over;
    if (r.w1 < (V)0) r.w1 = (V)0;
    r.w2 = (V)1 - r.w1;
    return r;
}

// Null test of the cold branch (BoltzmannMC.C:1049-1053): true for a real
// collision.  Also flags a flight whose trial frequency was below the total.
template <typename V>
__device__ __forceinline__ bool cold_null_test(const CondParams& c, const V* nutot, const Electron<V>& e,
                                               EventRng<V>& rng, CollOut<V>& o) {
    const ColdRow<V> r = cold_row(c, e);
    const V randF = e.nue * rng.u(5);
    const V tot = lerp2(r.w1, nutot[r.k1], r.w2, nutot[r.k2]);
    o.escape = e.nue < tot;
    return !(randF > tot);
}

// The same decision with the per-bin envelope envHi/envLo of the total
// frequency (bounds of every interpolated value a bin's rows can give,
// widened by a few ulp): only a variate between the bounds reads the table.
// The result, and the escape flag, equal those of cold_null_test.
template <typename V>
__device__ __forceinline__ bool cold_null_test_env(const CondParams& c, const V* nutot, const V* envHi,
                                                   const V* envLo, const Electron<V>& e, EventRng<V>& rng,
                                                   CollOut<V>& o) {
    const ColdRow<V> r = cold_row(c, e);
    const V randF = e.nue * rng.u(5);
    const int bin = r.k1 >> c.envShift;
    const V hi = envHi[bin], lo = envLo[bin];
    const bool needTot = (randF > lo && !(randF > hi)) || e.nue < hi;
    V tot = hi;
    if (needTot) tot = lerp2(r.w1, nutot[r.k1], r.w2, nutot[r.k2]);
    o.escape = needTot ? e.nue < tot : false;
    if (randF > hi) return false;
    if (!(randF > lo)) return true;
    return !(randF > tot);
}

// Cold channel choice of a real collision (BoltzmannMC.C:1055-1097)
template <typename V>
__device__ __forceinline__ int cold_choose(const CondParams& c, const DevProcesses& pr, const V* cumul,
                                           const V* interp, const Electron<V>& e, EventRng<V>& rng) {
    const int P = pr.P;
    const ColdRow<V> r = cold_row(c, e);
    const V randF = e.nue * rng.u(5);
    const V* c1 = cumul + (size_t)r.k1 * P;
    const V* c2 = cumul + (size_t)r.k2 * P;
    const V* i1 = interp + (size_t)r.k1 * P;
    const V* i2 = interp + (size_t)r.k2 * P;
    const V speed = sqrt(sqr_norm(e.vx, e.vy, e.vz));
    const V randCS = randF / (V)c.N / speed;
    return channel_search(c1, c2, i1, i2, r.w1, r.w2, 0, P - 1, randCS, (V)0, (V)0, speed, false);
}

// Thermal channel choice (BoltzmannMC.C:916-1031): target velocity Vt from
// three normals; -1 for a null collision.
template <typename V>
__device__ __forceinline__ int thermal_choose(const CondParams& c, const DevProcesses& pr, const V* cumul,
                                              const V* interp, const Electron<V>& e, EventRng<V>& rng, V Vt[3]) {
    const int P = pr.P, I = c.I;
    const V step = (V)c.step;
    // MathFunctions::unitNormalRand3 (MathFunctions.C:54-59)
    const V r5 = rng.u(5);
    V r1 = rng.u(1), r2 = rng.u(2), r3 = rng.u(3), r4 = rng.u(4);
    V s2, c2r, s4, c4r;
    turn_sincos(r2, s2, c2r);
    turn_sincos(r4, s4, c4r);
    V aux1 = sqrt((V)-2 * log_of(r1));
    V vers[3] = {aux1 * c2r, aux1 * s2, sqrt((V)-2 * log_of(r3)) * c4r};
    V randCS = e.nue * r5 / (V)c.N;
    V prevSum = (V)0;
    for (int g = 0; g < pr.G; ++g) {
        if (pr.gasFrac[g] == 0) continue;
        int L = pr.gasFirst[g], R = pr.gasLast[g];
        V sd = (V)pr.proc[L].thermalStd;
        for (int k = 0; k < 3; ++k) Vt[k] = vers[k] * sd;
        V dx = e.vx - Vt[0], dy = e.vy - Vt[1], dz = e.vz - Vt[2];
        V rel = sqrt(sqr_norm(dx, dy, dz));
        V over = (V)0.5 * (V)pr.proc[L].rMass * rel * rel / (V)kCharge / step;
        int k1 = (int)fmin(over, (V)(I - 1));
        int k2 = (int)fmin((V)(k1 + 1), (V)(I - 1));
        // the weights snap to {0, 1} (BoltzmannMC.C:959-967): row k1 has
        // weight 1 unless k2 - over < 0.  A snapped pair picks one row, and
        // 1 x + 0 x = x for the finite table values, so that row passed twice
        // with weights (1, 0) gives the same numbers from one load per entry
        const int kr = (V)k2 - over < (V)0 ? k2 : k1;
        const V* cr = cumul + (size_t)kr * P;
        const V* ir = interp + (size_t)kr * P;
        V ref = (V)0;
        if (L > 0) ref = lerp2((V)1, cr[L - 1], (V)0, cr[L - 1]);
        V lim = prevSum + (lerp2((V)1, cr[R], (V)0, cr[R]) - ref) * rel;
        if (randCS > lim) { prevSum = lim; continue; }
        return channel_search(cr, cr, ir, ir, (V)1, (V)0, L, R, randCS, prevSum, ref, rel, true);
    }
    return -1;
}

// The collision operator of process `chosen` (BoltzmannMC.C:1101-1280).
template <typename V, bool THERMAL>
__device__ __forceinline__ void apply_collision(const DevProcesses& pr, Electron<V>& e, EventRng<V>& rng,
                                                CollOut<V>& o, int chosen, const V Vt[3], bool thermal) {
    o.pid = chosen;
    const ProcRec& pc = pr.proc[chosen];
    const int type = pc.type;
    const V inc = e.eps;
    if (type == 2) {
        o.dE = -(double)inc;  // attachmentCollision: the electron is lost
        return;
    }
    if (THERMAL && thermal && type == 0) {
        // conservativeCollision with the target's thermal velocity
        double M = pc.tMass, mu = pc.rMass, loss = pc.energyLoss;
        Sph<V> s;
        if (!cart2sph(e.vx - Vt[0], e.vy - Vt[1], e.vz - Vt[2], s)) { o.flags |= ST_ZERO_VELOCITY; o.pid = -1; return; }
        double relE = 0.5 * mu * (double)s.norm * (double)s.norm / kCharge;
        double relAfter = relE - loss;
        if (relAfter <= 0) { o.pid = -1; return; }
        V cc = cos_chi<V>(pr, chosen, (V)relE, (V)relAfter, rng, 6);
        V sc = sqrt((V)1 - cc * cc);
        V se, ce;
        turn_sincos(rng.u(7), se, ce);
        V u[3];
        euler(sc, cc, se, ce, s, u);
        double rel2 = (double)s.norm * (double)s.norm - 2.0 / mu * loss * kCharge;
        if (!(rel2 > 0)) { o.pid = -1; return; }  // rounding at the threshold edge
        V sp = (V)sqrt(rel2);
        V fM = (V)(M / (kMass + M)), fe = (V)(kMass / (kMass + M));
        e.vx = fM * (sp * u[0]) + (fe * e.vx + fM * Vt[0]);
        e.vy = fM * (sp * u[1]) + (fe * e.vy + fM * Vt[1]);
        e.vz = fM * (sp * u[2]) + (fe * e.vz + fM * Vt[2]);
        e.eps = energy_of(e.vx, e.vy, e.vz);
        o.dE = (double)e.eps - (double)inc;
        return;
    }
    // conservativeCollision (cold) and ionizationCollision share the
    // scattering of the incident electron: one sampler, azimuth and rotation
    const bool ion = type == 1;
    const double loss = pc.energyLoss;  // energy loss, or the ionization threshold
    if (ion && (double)inc < loss) { o.pid = -1; return; }
    Sph<V> s;
    if (!cart2sph(e.vx, e.vy, e.vz, s)) { o.flags |= ST_ZERO_VELOCITY; o.pid = -1; return; }
    const double after = (double)inc - loss;  // energy after the loss; net energy of an ionization
    if (!ion && after <= 0) { o.pid = -1; return; }
    const V net = (V)after;
    V ej = (V)0, scE = net;
    if (ion) {
        if (pr.ionizOp == 2) {
            V w = (V)pc.wParam;
            ej = w * tan(rng.u(8) * atan(net / ((V)2 * w)));
        } else if (pr.ionizOp == 3) {
            ej = rng.u(8) * net;
        } else {
            ej = (V)pr.sharingFactor * net;
        }
        scE = net - ej;
    }
    const bool boeuf = ion && pc.scat == 5;
    V cc;
    if (boeuf) cc = sqrt(scE / net);
    else cc = cos_chi<V>(pr, chosen, inc, scE, rng, 6);
    const V sc = sqrt((V)1 - cc * cc);
    V se, ce;
    turn_sincos(rng.u(7), se, ce);
    V u[3];
    euler(sc, cc, se, ce, s, u);
    V sp;
    if (ion) {
        sp = sqrt((V)2 * scE * (V)kCharge / (V)kMass);
    } else {
        const double M = pc.tMass, mu = pc.rMass;
        double v2after = (double)s.norm * (double)s.norm - 2.0 / kMass * loss * kCharge;
        if (!(v2after > 0)) { o.pid = -1; return; }  // rounding at the threshold edge
        sp = sqrt((V)v2after * ((V)1 - (V)(2.0 * mu / (kMass + M)) * ((V)1 - cc)));
    }
    e.vx = sp * u[0]; e.vy = sp * u[1]; e.vz = sp * u[2];
    if (!ion) {
        e.eps = energy_of(e.vx, e.vy, e.vz);
        o.dE = (double)e.eps - (double)inc;
        return;
    }
    // the ejected electron (ionizationCollision, BoltzmannMC.C:1228-1267)
    V cCe, sCe, sEe, cEe;
    if (boeuf) {
        cCe = sqrt(ej / net); sCe = sqrt((V)1 - cCe * cCe);
        sEe = -se; cEe = -ce;
    } else {
        cCe = cos_chi<V>(pr, chosen, inc, ej, rng, 9);
        sCe = sqrt((V)1 - cCe * cCe);
        turn_sincos(rng.u(10), sEe, cEe);
    }
    euler(sCe, cCe, sEe, cEe, s, u);
    V spE = sqrt((V)2 * ej * (V)kCharge / (V)kMass);
    o.ejv[0] = spE * u[0]; o.ejv[1] = spE * u[1]; o.ejv[2] = spE * u[2];
    o.ejEps = ej;
    e.eps = scE;
    o.dE = -loss;
}

template <typename V> __device__ __forceinline__ void clear_out(CollOut<V>& o) {
    o.pid = -1; o.flags = 0; o.ejEps = (V)0; o.dE = 0; o.escape = false;
    o.ejv[0] = o.ejv[1] = o.ejv[2] = (V)0;
}

// performCollision: null test, channel choice and collision operator.
template <typename V, bool THERMAL>
__device__ inline void collide(const CondParams& c, const DevProcesses& pr, const V* cumul, const V* interp,
                               const V* nutot, Electron<V>& e, EventRng<V>& rng, CollOut<V>& o) {
    clear_out(o);
    V Vt[3] = {(V)0, (V)0, (V)0};
    const bool thermal = THERMAL && is_thermal(c, pr, e);
    int chosen;
    if (THERMAL && thermal) {
        chosen = thermal_choose(c, pr, cumul, interp, e, rng, Vt);
        if (chosen == -1) return;
    } else {
        if (!cold_null_test(c, nutot, e, rng, o)) return;
        chosen = cold_choose(c, pr, cumul, interp, e, rng);
    }
    apply_collision<V, THERMAL>(pr, e, rng, o, chosen, Vt, thermal);
}

// ------------------------------------------------------ two-phase step
// One step of an electron in the passes of BoltzmannMC.C:637-680, split so
// that a wave can run the real collision operator for many electrons at
// once.  Phase A: draw a free time if none is pending (the first variate of
// call 0 of the electron's current event), fly to t_sync or to the collision
// and, in the cold branch, run the null test.  Phase B: the channel choice
// and collision operator of a real (or thermal) collision.
enum StepResult { STEP_DONE = 1, STEP_NULL = 2, STEP_REAL = 3 };

template <typename V, int FC, bool THERMAL>
__device__ __forceinline__ int step_a(const CondParams& c, const DevProcesses& pr, const V* nutot, const V* envHi,
                                      const V* envLo, Electron<V>& e, EventRng<V>& rng, CollOut<V>& o,
                                      double& gain) {
    if (e.t == c.tsync) return STEP_DONE;
    const bool pending = e.tfree < (V)0;  // pending free time undefined (NON_DEF is negative)
    if (pending || e.t + (double)e.tfree <= c.tsync) rng.prime0();
    if (pending) {
        e.tfree = free_time(rng.u(0), c);
        e.nue = (V)c.trial;
    }
    // one flight for both outcomes (a wave holds both kinds of lanes)
    const bool partial = e.t + (double)e.tfree > c.tsync;
    const double dt = partial ? c.tsync - e.t : (double)e.tfree;
    if (FC == FLIGHT_DC_ANY) {
        if (c.fieldCase == FIELD_DC) flight<FIELD_DC>(c, e, dt, gain);
        else flight<FIELD_DC_B>(c, e, dt, gain);
    } else {
        flight<FC>(c, e, dt, gain);
    }
    if (partial) {
        e.t = c.tsync;
        e.tfree = (V)((double)e.tfree - dt);
        return STEP_DONE;
    }
    e.t += (double)e.tfree;
    if (THERMAL && is_thermal(c, pr, e)) {
        o.escape = false;  // only the cold null test compares the trial frequency with the total
        return STEP_REAL;
    }
    return cold_null_test_env(c, nutot, envHi, envLo, e, rng, o) ? STEP_REAL : STEP_NULL;
}

template <typename V, bool THERMAL>
__device__ __forceinline__ void step_b(const CondParams& c, const DevProcesses& pr, const V* cumul, const V* interp,
                                       Electron<V>& e, EventRng<V>& rng, CollOut<V>& o) {
    clear_out(o);
    V Vt[3] = {(V)0, (V)0, (V)0};
    const bool thermal = THERMAL && is_thermal(c, pr, e);
    int chosen;
    if (THERMAL && thermal) {
        chosen = thermal_choose(c, pr, cumul, interp, e, rng, Vt);
        if (chosen == -1) return;
    } else {
        chosen = cold_choose(c, pr, cumul, interp, e, rng);
    }
    apply_collision<V, THERMAL>(pr, e, rng, o, chosen, Vt, thermal);
}

// After a collision event: the next event's free time is drawn when the next
// step starts, except for a collision exactly at t_sync, whose next free
// time is drawn now with this interval's trial frequency (BoltzmannMC.C:672).
template <typename V> __device__ __forceinline__ void after_event(const CondParams& c, Electron<V>& e) {
    e.event += 1;
    e.tfree = (V)kNonDef;
    if (e.t == c.tsync) {
        EventRng<V> r(c.seed, e.stream, e.event);
        e.tfree = free_time(r.u(0), c);
        e.nue = (V)c.trial;
    }
}

}  // namespace ekmc
