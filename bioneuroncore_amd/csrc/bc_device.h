// bc_device.h — device-side biophysics of the per-tick step, operating on register-resident
// segment state. Every expression keeps the reference's float/double type and evaluation
// order (the kernels are built with -ffp-contract=off) and all exp calls go through
// bc::exp, which returns the bits of the reference's glibc ::exp. The functions cite the
// reference routine they follow.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bc_libm.h"

#define BC_DEV __device__ static inline __attribute__((always_inline))

namespace bc {

constexpr double kFaraday = 96485.0;
constexpr float kGas = 8.314f;
constexpr float kTemp = 310.15f;
constexpr float kPi = 3.14159265f;
constexpr float kEvr = 0.2f / (1.0f - 0.2f);
constexpr float kAtoPa = 1e12f;
constexpr float kPsToS = 1e-12f;
constexpr float kMmToUm = 1e3f;
constexpr float kCbarNa = 55.964f, kCbarK = 40.514f, kCbarCa = 0.12775f, kCbarCl = 36.20f;
constexpr float kSigmaE = 3.612f;
constexpr float kFrtMv = 0.03742f;
constexpr float kChargeToPa = 96.485f;

__device__ __constant__ static const float kZ[7] = {1.0f, 1.0f, 2.0f, -1.0f, 2.0f, -1.0f, 1.0f};
__device__ __constant__ static const float kDiff[7] = {1.33f, 1.96f, 0.22f, 2.03f, 0.10f, 0.0f, 9.31f};
__device__ __constant__ static const float kBath[7] = {140.0f, 3.0f, 1.2f, 130.0f, 1.0f, 0.0f, 4.0e-5f};

enum : uint8_t {
    M_PUMP = 7, M_AMPA = 8, M_NMDA = 9, M_GABAA = 11, M_SK = 103, M_BK = 104, M_K2P = 107,
    M_KV7 = 112, M_KVDR = 123, M_HCN1 = 129, M_NALCN = 133, M_CAV21 = 140, M_NAV16 = 141,
    M_NAV12 = 142, M_KV31 = 150, M_KV11 = 156, M_CX36 = 114, M_CX43 = 198,
    M_GLAST = 147, M_GLYT1 = 148, M_GAT3 = 149, M_EAAT3 = 183, M_GLYT2 = 184, M_GAT1 = 185,
    M_HCN2 = 130, M_PUMP_A3 = 197, M_PUMP_GLIA = 110, M_PMCA = 98, M_NCX = 99, M_KCC2 = 100,
    M_NKCC1 = 122, M_CA_BUFFER = 102, M_CAV13 = 1, M_CLC2 = 2, M_KV42 = 85, M_KV21 = 196,
    M_NAP = 105, M_NAP_TAD = 195, M_KIR = 106, M_NAV12DEV = 193,
};

// The core channel set: pump, ionotropic receptors, the K+/Na+/Ca2+/HCN/NALCN channels of the
// tree and chain workloads. Every other covered ID is an extended type.
__host__ __device__ static inline bool modifier_core(uint8_t t) {
    switch (t) {
        case M_PUMP: case M_AMPA: case M_NMDA: case M_GABAA: case M_SK: case M_BK: case M_K2P:
        case M_KV7: case M_KVDR: case M_HCN1: case M_NALCN: case M_CAV21: case M_NAV16:
        case M_NAV12: case M_KV31: case M_KV11: case 4:
            return true;
        default:
            return false;
    }
}

// Ionotropic types (kernel_dispatch.h:283-289) and the channel types myelin blocks
// (kernel_dispatch.h:299-309) among the covered IDs; myelin blocks no pump or transporter.
__host__ __device__ static inline bool ionotropic_type(uint8_t t) {
    return t <= 2 || (t >= 8 && t <= 16) || t == 97 || t == 134 || t == 194;
}
__host__ __device__ static inline bool myelin_blocks(uint8_t t) {
    return ionotropic_type(t) || t == M_K2P || t == M_NALCN || t == M_KV31 || t == M_KV11 || t == M_KV7 ||
           t == M_KVDR || t == M_CAV21 || t == M_HCN1 || t == M_SK || t == M_BK || t == M_KV42 ||
           t == M_NAP || t == M_NAP_TAD || t == M_HCN2 || t == 131 || t == 132 || t == 154 || t == 157 ||
           t == 205 || t == 206 || t == 207 || t == 202 || t == 203 || t == 204 || t == 168 || t == 200 ||
           t == 201 || t == 191 || t == 159 || t == 96 || t == 84 || t == 152 || t == 151 || t == 153 ||
           t == 158 || t == 121 || t == 209;
}

// G-protein-coupled receptor types: they read state[active_vars + 1].
__host__ __device__ static inline bool gprotein_type(uint8_t t) {
    return (t >= 17 && t <= 62) || t == 70 || t == 74;
}

// Modifiers that read the segment's extracellular transmitters in step 1 (the ionotropic
// receptors) or also take them up (the transporters).
__host__ __device__ static inline bool modifier_reads_nt(uint8_t t) {
    return t == M_AMPA || t == M_NMDA || t == M_GABAA || (t >= M_GLAST && t <= M_GAT3) ||
           (t >= M_EAAT3 && t <= M_GAT1) || (t >= 10 && t <= 64) || (t >= 67 && t <= 68) ||
           (t >= 70 && t <= 76) || t == 96 || t == 97 || t == 134 || t == 194 || t == 3;
}

__host__ __device__ static inline bool modifier_supported(uint8_t t) {
    switch (t) {
        case M_PUMP: case M_AMPA: case M_NMDA: case M_GABAA: case M_SK: case M_BK: case M_K2P:
        case M_KV7: case M_KVDR: case M_HCN1: case M_NALCN: case M_CAV21: case M_NAV16:
        case 4:   // myelin: no current, no state (hh_biological.cpp:450-453); blocks others
        case M_NAV12: case M_KV31: case M_KV11: case M_CX36: case M_CX43: case M_GLAST:
        case M_GLYT1: case M_GAT3: case M_EAAT3: case M_GLYT2: case M_GAT1: case M_HCN2:
        case M_PUMP_A3: case M_PUMP_GLIA: case M_PMCA: case M_NCX: case M_KCC2: case M_NKCC1:
        case M_CA_BUFFER: case M_CAV13: case M_CLC2: case M_KV42: case M_KV21: case M_NAP:
        case M_NAP_TAD: case M_KIR: case M_NAV12DEV: case 10: case 12: case 13: case 14: case 97:
        case 194: case 15: case 134: case 63: case 64: case 67: case 68: case 70: case 71: case 72:
        case 73: case 74: case 131: case 132: case 169: case 170: case 171: case 177: case 199:
        case 154: case 157: case 205: case 206: case 207: case 202: case 203: case 204: case 160:
        case 161: case 166: case 167: case 208: case 178: case 176: case 168: case 200: case 201:
        case 191: case 210: case 162: case 211: case 175: case 182: case 172: case 173: case 174:
        case 179: case 163: case 164: case 180: case 159: case 101: case 96: case 84: case 144:
        case 145: case 146: case 152: case 186: case 190: case 86: case 126: case 127: case 128:
        case 136: case 3: case 5: case 77: case 78: case 79: case 188: case 16: case 87: case 88:
        case 89: case 90: case 91: case 92: case 192: case 6: case 94: case 95: case 108: case 111:
        case 115: case 116: case 117: case 120: case 82: case 83: case 113: case 137: case 75: case 76:
        // IDs without code in the reference (reserved, or plugin stubs with no plugin loaded):
        // no current and no state change there (kernel_dispatch.h:565, dispatch_generated.cpp:
        // 26-31), which is the device switch's default
        case 0: case 80: case 81: case 93: case 124: case 125: case 138: case 139: case 143:
            return true;
        default:
            return (t >= 17 && t <= 62) || t >= 212;   // the G-protein-coupled receptor families
    }
}

// Segment state held in registers while its modifiers run (step 1) or while the receptors of
// its incoming synapses run (step 3).
struct SegRegs {
    double V;
    double in[7];
    double ex[7];
    float atp;
    float area;    // membrane area: sphere 4*pi*r^2 or cylinder 2*pi*r*L (float)
    double vol;    // volume (double, run_tick_common.h:260-264)
    double cpm;    // F * vol / 1000, the charge-per-mole factor every current divides by
    float pkc, pka, camp;   // first PKC state[0] (live), first cAMP/PKA state[1] and state[0] (constants)
    float camkii;           // state[0] of the last CaMKII modifier, read by the connexins
    float camkii01;         // its state[0] + state[1], read by Kv4.2
    float cabp4;            // state[0] of the first CaBP4 modifier, set before each Cav1.4 runs
    float dag;              // state[1] of the first PLC/IP3/DAG modifier, set before eCB synthesis runs
    float ip3;              // its state[0], set before IP3R runs
    float gq;               // the Gq receptors' state[0] summed, set before PLC runs
    float er;               // ER_Ca2 (segments with SERCA, IP3R or RyR)
    float temp, txn;        // temperature, transcription signal (segments with 16, 77-79 or 88)
    float xa, xb;           // the last mTORC1 / MAPK / NF-kB states protein synthesis, cofilin
                            // and MMP-9 read, set before they run
    float len;              // segment length, set before a physical sensor runs
    uint32_t sid;           // Segment.id of the record, set before thermal noise runs
    float nt[15];  // neurotransmitter (used by receptors)
    // exp(-zFV/RT) of the GHK helpers for the two V values the channels pass (V rounded to
    // float, and V as double): V is fixed while one segment's modifiers or receptors run, so
    // every GHK call reuses these instead of recomputing the same exp.
    double vkf, vkd;        // keys
    double emf, edf, emd, edd;
    double vkmg, emg;       // exp(-0.062 V) of the NMDA Mg2+ block, keyed on V
};

BC_DEV double vn_mono(double V) { return kFaraday * V * 0.001 / (kGas * kTemp); }
BC_DEV double vn_di(double V) { return 2.0 * kFaraday * V * 0.001 / (kGas * kTemp); }
BC_DEV void seg_consts_init(SegRegs& s) { s.cpm = kFaraday * s.vol / 1000.0; }
BC_DEV void ghk_cache_init(SegRegs& s) {
    double vf = (float)s.V;
    s.vkf = vf;
    s.vkd = s.V;
    s.emf = bc::exp(-vn_mono(vf));
    s.edf = bc::exp(-vn_di(vf));
    s.emd = bc::exp(-vn_mono(s.V));
    s.edd = bc::exp(-vn_di(s.V));
    s.vkmg = __builtin_nan("");
}
BC_DEV void mg_cache_init(SegRegs& s) {
    s.vkmg = s.V;
    s.emg = bc::exp(-0.062f * s.V);
}
// kSegV: the caller passes V == s.V (the double key) and both caches are initialised, so the
// cached value is the exp itself, NaN V included (the cache was computed by the same expression).
template <bool kSegV = false>
BC_DEV double exp_mono(const SegRegs& s, double V, double vn) {
    if (kSegV) return s.emd;
    if (V == s.vkf) return s.emf;
    if (V == s.vkd) return s.emd;
    return bc::exp(-vn);
}
template <bool kSegV = false>
BC_DEV double exp_di(const SegRegs& s, double V, double vn) {
    if (kSegV) return s.edd;
    if (V == s.vkf) return s.edf;
    if (V == s.vkd) return s.edd;
    return bc::exp(-vn);
}

// One modifier slot: states 0..6 live in SoA memory; state[7] (effective area) is a register.
struct ModSlot {
    double* st;
    size_t stride;
    double dens;
    float a7;
    uint8_t av;
    __device__ double get(int k) const { return st[(size_t)k * stride]; }
    __device__ void set(int k, double v) const { st[(size_t)k * stride] = v; }
};

template <class M>
BC_DEV float eff_area(const M& m, const SegRegs& s) { return (m.a7 > 0.0f) ? m.a7 : s.area; }
template <int N>
BC_DEV void floor_n(double* a) {
#pragma unroll
    for (int i = 0; i < N; i++)
        if (a[i] < 1e-9) a[i] = 1e-9;
}
BC_DEV void floor7(double* a) {
#pragma unroll
    for (int i = 0; i < 7; i++)
        if (a[i] < 1e-9) a[i] = 1e-9;
}

// ---- GHK (run_tick_common.h:274-312, modifier_helpers.h:48-75) ----
template <bool kSegV = false>
BC_DEV double ghk_mono(const SegRegs& s, double V, double ci, double co) {
    double vn = vn_mono(V);
    if (fabs(vn) < 1e-6) return kFaraday * (ci - co);
    double k = kFaraday / (kGas * kTemp);
    if (vn > 20.0) return k * kFaraday * V * 0.001 * ci;
    if (vn < -20.0) return k * kFaraday * V * 0.001 * co;
    double e = exp_mono<kSegV>(s, V, vn);
    return k * kFaraday * V * 0.001 * (ci - co * e) / (1.0 - e);
}
template <bool kSegV = false>
BC_DEV double ghk_di(const SegRegs& s, double V, double ci, double co) {
    double vn = vn_di(V);
    if (fabs(vn) < 1e-6) return 2.0 * kFaraday * (ci - co);
    double k = 4.0 * kFaraday / (kGas * kTemp);
    if (vn > 20.0) return k * kFaraday * V * 0.001 * ci;
    if (vn < -20.0) return k * kFaraday * V * 0.001 * co;
    double e = exp_di<kSegV>(s, V, vn);
    return k * kFaraday * V * 0.001 * (ci - co * e) / (1.0 - e);
}
BC_DEV double ghk_di_nocache(double V, double ci, double co) {
    double vn = vn_di(V);
    if (fabs(vn) < 1e-6) return 2.0 * kFaraday * (ci - co);
    double k = 4.0 * kFaraday / (kGas * kTemp);
    if (vn > 20.0) return k * kFaraday * V * 0.001 * ci;
    if (vn < -20.0) return k * kFaraday * V * 0.001 * co;
    double e = bc::exp(-vn);
    return k * kFaraday * V * 0.001 * (ci - co * e) / (1.0 - e);
}
template <bool kSegV = false>
BC_DEV double ghk_an(const SegRegs& s, double V, double ci, double co) {
    double vn = vn_mono(V);
    if (fabs(vn) < 1e-6) return kFaraday * (co - ci);
    double k = kFaraday / (kGas * kTemp);
    if (vn > 20.0) return k * kFaraday * V * 0.001 * co;
    if (vn < -20.0) return k * kFaraday * V * 0.001 * ci;
    double e = exp_mono<kSegV>(s, V, vn);
    return k * kFaraday * V * 0.001 * (co - ci * e) / (1.0 - e);
}
__host__ __device__ constexpr double p_single(float gamma_ps, float z, float cbar) {
    // evaluated at compile time for constant gamma: clang folds each IEEE operation exactly
    return (double)(gamma_ps * kPsToS) * (double)(kGas * kTemp) / ((double)(z * z) * kFaraday * kFaraday * cbar);
}
struct P3 { double na, k, ca; };
__host__ __device__ constexpr P3 p_partial3(float gamma_ps, float fna, float fk, float fca) {
    double w = fna * kCbarNa + fk * kCbarK + fca * 4.0 * kCbarCa;
    return P3{p_single((float)(gamma_ps * fna * kCbarNa / w), 1.0f, kCbarNa),
              p_single((float)(gamma_ps * fk * kCbarK / w), 1.0f, kCbarK),
              p_single((float)(gamma_ps * fca * 4.0 * kCbarCa / w), 2.0f, kCbarCa)};
}
__host__ __device__ constexpr P3 p_nalcn3(float gamma_ps) {
    const double fna = 1.3, fk = 1.2, fca = 0.5;
    double w = fna * kCbarNa + fk * kCbarK + fca * 4.0 * kCbarCa;
    return P3{p_single((float)(gamma_ps * fna * kCbarNa / w), 1.0f, kCbarNa),
              p_single((float)(gamma_ps * fk * kCbarK / w), 1.0f, kCbarK),
              p_single((float)(gamma_ps * fca * 4.0 * kCbarCa / w), 2.0f, kCbarCa)};
}

// ---- gates (modifier_helpers.h:132-271, 909-1013) ----
BC_DEV float vtrap(float x, float y) {
    if (fabsf(x / y) < 1e-6f) return y * (1.0f - x / (2.0f * y));
    return x / (1.0f - bc::exp(-x / y));
}
struct NavK { float am_off, am_k, ah_off, bh_off, h_off; };
BC_DEV float nav_am(const NavK& g, float V) { return 0.182f * vtrap(V + g.am_off, g.am_k); }
BC_DEV float nav_bm(const NavK& g, float V) { return 0.124f * vtrap(-(V + g.am_off), g.am_k); }
BC_DEV float nav_ah(const NavK& g, float V) { return 0.024f * vtrap(V + g.ah_off, 5.0f); }
BC_DEV float nav_bh(const NavK& g, float V) { return 0.0091f * vtrap(-(V + g.bh_off), 5.0f); }

BC_DEV float cav21_minf(float V) { return 1.0f / (1.0f + bc::exp(-(V - -15.0f) / (6.0f + 1e-6f))); }
BC_DEV float cav21_hinf(float V, float ca_um) {
    float h = 1.0f / (1.0f + bc::exp((V - -50.0f) / 8.0f));
    h *= 1.0f / (1.0f + ca_um / 3.0f);
    return h;
}
BC_DEV float kdr_tau(float V, float vh) {
    float tau = 14.6f * bc::exp(-(V - vh) / 30.1f);
    if (tau < 0.3f) tau = 0.3f;
    if (tau > 200.0f) tau = 200.0f;
    return tau;
}
// x^4 for a float x: x*x is exact in double, so one rounding remains; this is the correctly
// rounded value that glibc's pow(x, 4.0) returns except in rare last-bit cases, and every use
// rounds the result to float right after (probability of a differing float ~1e-12).
BC_DEV double pow4(float x) {
    double a = (double)x * (double)x;
    return a * a;
}
BC_DEV float sk_pinf(float ca) {
    if (ca < 1e-9f) ca = 1e-9f;
    float r = 0.0004f / ca;
    return 1.0f / (1.0f + pow4(r));
}
BC_DEV float bk_po(float V, float ca) {
    float vn = kFaraday * V * 1e-3f / (kGas * kTemp);
    float J = 0.104f * bc::exp(0.55f * vn);
    float L = 3.7e-8f * bc::exp(0.30f * vn);
    if (ca < 0.0f) ca = 0.0f;
    float K = ca / 0.044f;
    float o = 1.0f + K * 23.0f + J * 32.0f + J * K * 23.0f * 32.0f * 3.7f;
    float c = 1.0f + K + J + J * K * 3.7f;
    o *= o; o *= o;
    c *= c; c *= c;
    float num = L * o;
    return num / (num + c);
}
BC_DEV float pump_vnorm(float V, float na, float ko, float atp_level, float kmna = 10.0f, float kmk = 1.5f) {
    float na3 = na * na * na, k2 = ko * ko;
    float kna3 = kmna * kmna * kmna, kk2 = kmk * kmk;
    float fna = na3 / (na3 + kna3), fk = k2 / (k2 + kk2);
    float atp = atp_level;
    if (atp < 0.0f) atp = 0.0f;
    float fatp = atp / (atp + 0.5f);
    float eta = 0.8f * kFaraday * (V * 1e-3f) / (kGas * kTemp);
    float a2 = 60.0f * bc::exp(0.5f * eta);
    float b2 = 6.0f * bc::exp(-(1.0f - 0.5f) * eta);
    float a1 = 200.0f * fna * fatp;
    float a3 = 200.0f * fk;
    float a4 = 60.0f;
    float denom = a3 / a1 + (b2 + a3) / a2 + 1.0f + a3 / a4;
    float turnover = a3 / denom;
    float k3 = 200.0f;
    float vmax = k3 / (k3 / 200.0f + 1.0f + k3 / 60.0f);
    return turnover / vmax;
}

// ---- channel currents: return pA (reference sign convention) ----
// Na/K-ATPase (7), its alpha3 isoform (197: K_m Na 25 mM, ouabain block by state[5] nM, IC50
// 20 nM) and the glial alpha2 pump (110: K_m K 1 mM) (hh_biological.cpp:57-141, 3414-3445).
enum : int { PUMP_A1 = 0, PUMP_A3 = 1, PUMP_GLIA = 2 };
template <int kKind = PUMP_A1>
BC_DEV double ch_pump(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    float vn = pump_vnorm(s.V, s.in[0], s.ex[1], s.atp, kKind == PUMP_A3 ? 25.0f : 10.0f,
                          kKind == PUMP_GLIA ? 1.0f : 1.5f);
    if (kKind == PUMP_A3) {
        double s5 = m.get(5);
        double oua = (s5 > 0.0f) ? s5 : 0.0f;
        vn /= (1.0f + oua / 20.0f);
    }
    double N = m.dens * area;
    float ip = N * vn * 75.0f * 1.602176634e-19f * kAtoPa;
    float vol = s.vol;
    double cpm = kFaraday * vol / 1000.0;
    float tr = ip * dt / cpm;
    float na_av = (s.in[0] - 1e-9f) / 3.0f;
    float k_av = (s.ex[1] - 1e-9f) * kEvr / 2.0f;
    if (na_av < 0.0f) na_av = 0.0f;
    if (k_av < 0.0f) k_av = 0.0f;
    float mx = (na_av < k_av) ? na_av : k_av;
    if (tr > mx) tr = mx;
    s.in[0] -= 3.0f * tr;
    s.ex[0] += 3.0f * tr / kEvr;
    s.in[1] += 2.0f * tr;
    s.ex[1] -= 2.0f * tr / kEvr;
    floor7(s.in);
    floor7(s.ex);
    s.atp -= ip * dt * 1e-8f;
    if (s.atp < 0.0f) s.atp = 0.0f;
    return -ip;
}

BC_DEV double ch_nav(const ModSlot& m, SegRegs& s, float dt, const NavK& g, float gamma, float ttx_ic50,
                     float pkx) {   // hh_biological.cpp:193-277
    float area = eff_area(m, s);
    if (s.V < -200.0f) s.V = -200.0f;
    if (s.V > 200.0f) s.V = 200.0f;
    float V = s.V;
    float am = nav_am(g, V), bm = nav_bm(g, V);
    float minf = am / (am + bm);
    float mtau = 1.0f / (am + bm);
    float hinf = 1.0f / (1.0f + bc::exp((V + g.h_off) / 6.2f));
    float htau = 1.0f / (nav_ah(g, V) + nav_bh(g, V));
    const double s0 = m.get(0), s1 = m.get(1), s5 = m.get(5);   // all loads before the stores
    double mv = minf + (s0 - minf) * bc::exp(-dt / mtau);
    double hv = hinf + (s1 - hinf) * bc::exp(-dt / htau);
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m.set(0, mv);
    m.set(1, hv);
    float op = mv * mv * mv * hv;
    double ttx = (s5 > 0.0f) ? s5 : 0.0f;
    op /= (1.0f + ttx / ttx_ic50);
    op *= (1.0f - pkx * (s.pkc + s.pka) * 0.5f);
    if (op < 0.0f) op = 0.0f;
    double P = p_single(gamma, 1.0f, kCbarNa);
    double flux = ghk_mono(s, V, s.in[0], s.ex[0]);
    double ia = P * flux * (m.dens * area) * op;
    double ina = ia * kAtoPa;
    double cpv = s.cpm;
    s.in[0] -= (ina * dt) / cpv;
    s.ex[0] += (ina * dt) / (cpv * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -ina;
}

BC_DEV double ch_cav21(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:364-388
    float V = s.V;
    float mi = cav21_minf(V);
    double s0 = m.get(0), s1 = m.get(1);
    s0 += (mi - s0) * dt / 1.0f;
    float ca = s.in[2] * kMmToUm;
    float hi = cav21_hinf(V, ca);
    s1 += (hi - s1) * dt / 300.0f;
    if (s0 < 0.0f) s0 = 0.0f;
    if (s1 < 0.0f) s1 = 0.0f;
    if (s0 > 1.0f) s0 = 1.0f;
    if (s1 > 1.0f) s1 = 1.0f;
    m.set(0, s0);
    m.set(1, s1);
    return 0.0f;
}

BC_DEV double k_out(SegRegs& s, double ik, float dt) {
    double cpv = s.cpm;
    s.in[1] -= (ik * dt) / cpv;
    s.ex[1] += (ik * dt) / (cpv * kEvr);
    floor7(s.ex);
    return -ik;
}

BC_DEV double ch_kv31(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:1043-1100
    float V = s.V;
    float area = eff_area(m, s);
    double s0 = m.get(0), s1 = m.get(1), s2 = m.get(2), s5 = m.get(5);
    float mi = 1.0f / (1.0f + bc::exp(-(V - 10.0f) / (10.0f + 1e-6f)));
    float tau_m = 1.0f;
    s0 += (mi - s0) * dt / tau_m;
    if (s0 < 0.0f) s0 = 0.0f;
    if (s0 > 1.0f) s0 = 1.0f;
    float hi = 1.0f / (1.0f + bc::exp((V + 30.0f) / 6.0f));
    float tau_h = 2000.0f;
    s1 += (hi - s1) * dt / tau_h;
    if (s1 < 0.0f) s1 = 0.0f;
    if (s1 > 1.0f) s1 = 1.0f;
    float ki = s.in[1], ko = s.ex[1];
    double P = p_single(27.0f, 1.0f, kCbarK);
    double op = s0 * s0 * s0 * s1;
    double tea = (s5 > 0.0f) ? s5 : 0.0f;
    op /= (1.0f + tea / 0.2f);
    float kp = 5.6e-6f, kd = 1e-4f;
    s2 += (kp * s.pkc * (1.0f - s2) - kd * s2) * dt;
    if (s2 < 0.0f) s2 = 0.0f;
    if (s2 > 1.0f) s2 = 1.0f;
    op *= (1.0f - 0.27f * s2);
    m.set(0, s0);
    m.set(1, s1);
    m.set(2, s2);
    double flux = ghk_mono(s, V, ki, ko);
    double ia = P * flux * (m.dens * area) * op;
    return k_out(s, ia * kAtoPa, dt);
}

BC_DEV double ch_kv11(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:1189-1231
    float V = s.V;
    float area = eff_area(m, s);
    double s0 = m.get(0), s1 = m.get(1), s5 = m.get(5), s6 = m.get(6);
    float vh = -30.0f, kk = 6.0f, gamma = 10.0f;
    float mi = 1.0f / (1.0f + bc::exp(-(V - vh) / (kk + 1e-6f)));
    float tau_m = kdr_tau(V, vh);
    s0 += (mi - s0) * dt / tau_m;
    if (s0 < 0.0f) s0 = 0.0f;
    if (s0 > 1.0f) s0 = 1.0f;
    float hi = 1.0f / (1.0f + bc::exp((V + 40.0f) / 6.0f));
    float tau_h = 3000.0f;
    s1 += (hi - s1) * dt / tau_h;
    if (s1 < 0.0f) s1 = 0.0f;
    if (s1 > 1.0f) s1 = 1.0f;
    m.set(0, s0);
    m.set(1, s1);
    double op = s0 * s0 * s0 * s0 * s1;
    double dtx = (s5 > 0.0f) ? s5 : 0.0f;
    double tea = (s6 > 0.0f) ? s6 : 0.0f;
    op /= (1.0f + dtx / 2.0f) * (1.0f + tea / 170.0f);
    float ki = s.in[1], ko = s.ex[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    double ia = P * flux * (m.dens * area) * op;
    return k_out(s, ia * kAtoPa, dt);
}

BC_DEV double ch_kvdr(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:2763-2798
    float V = s.V;
    float area = eff_area(m, s);
    double mv = m.get(0), s1 = m.get(1), s2 = m.get(2), s4 = m.get(4);
    double vh = (s1 != 0.0f) ? s1 : -30.0f;
    double kk = (s2 > 0.0f) ? s2 : 6.0f;
    double tau = (s4 > 0.0f) ? s4 : kdr_tau(V, vh);
    float mi = 1.0f / (1.0f + bc::exp(-(V - vh) / (kk + 1e-6f)));
    mv += (mi - mv) * dt / (tau + 1e-6f);
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m.set(0, mv);
    float ki = s.in[1], ko = s.ex[1];
    double P = p_single(12.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    double ia = P * flux * (m.dens * area) * mv;
    double ik = ia * kAtoPa;
    float cpmv = s.cpm;
    s.in[1] -= (ik * dt) / cpmv;
    s.ex[1] += (ik * dt) / (cpmv * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -ik;
}

BC_DEV double ch_sk(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:2859-2888
    float V = s.V;
    float area = eff_area(m, s);
    double op = m.get(0), s3 = m.get(3), s4 = m.get(4), s5 = m.get(5);
    double gamma = (s3 > 0.0f) ? s3 : 10.0f;
    double t_open = (s4 > 0.0f) ? s4 : 5.0f;
    double t_close = (s5 > 0.0f) ? s5 : 50.0f;
    float ca = s.in[2];
    float pi = sk_pinf(ca);
    float tau = (pi > op) ? t_open : t_close;
    op += (pi - op) * dt / tau;
    if (op < 0.0f) op = 0.0f;
    if (op > 1.0f) op = 1.0f;
    m.set(0, op);
    float ki = s.in[1], ko = s.ex[1];
    double P = (s3 > 0.0f) ? p_single(gamma, 1.0f, kCbarK) : p_single(10.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    double ia = P * flux * (m.dens * area) * op;
    float ik = (float)(ia * kAtoPa);
    double cpv = s.cpm;
    s.in[1] -= (ik * dt) / cpv;
    s.ex[1] += (ik * dt) / (cpv * kEvr);
    floor7(s.ex);
    return -ik;
}

BC_DEV double ch_bk(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:2922-2949
    float V = s.V;
    float area = eff_area(m, s);
    double op = m.get(0), s1 = m.get(1), s2 = m.get(2);
    double gamma = (s1 > 0.0f) ? s1 : 250.0f;
    double tau = (s2 > 0.0f) ? s2 : 1.0f;
    float ca = s.in[2];
    float pi = bk_po(V, ca);
    op += (pi - op) * dt / tau;
    if (op < 0.0f) op = 0.0f;
    if (op > 1.0f) op = 1.0f;
    m.set(0, op);
    float ki = s.in[1], ko = s.ex[1];
    double P = (s1 > 0.0f) ? p_single(gamma, 1.0f, kCbarK) : p_single(250.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    double ia = P * flux * (m.dens * area) * op;
    float ik = (float)(ia * kAtoPa);
    double cpv = s.cpm;
    s.in[1] -= (ik * dt) / cpv;
    s.ex[1] += (ik * dt) / (cpv * kEvr);
    floor7(s.ex);
    return -ik;
}

BC_DEV double ch_k2p(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:3063-3085
    float V = s.V;
    float area = eff_area(m, s);
    double s0 = m.get(0), s1 = m.get(1);
    double gamma = (s0 > 0.0f) ? s0 : 100.0f;
    double po = (s1 > 0.0f) ? s1 : 0.012f;
    float ki = s.in[1], ko = s.ex[1];
    double P = (s0 > 0.0f) ? p_single(gamma, 1.0f, kCbarK) : p_single(100.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    double ia = P * flux * (m.dens * area) * po;
    return k_out(s, ia * kAtoPa, dt);
}

BC_DEV double ch_kv7(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:3158-3220
    float V = s.V;
    float area = eff_area(m, s);
    double mv = m.get(0), s1 = m.get(1), s2 = m.get(2), s4 = m.get(4), s5 = m.get(5), s6 = m.get(6);
    double vh = (s1 != 0.0f) ? s1 : -38.0f;
    double ks = (s2 > 0.0f) ? s2 : 10.0f;
    double t_act = (s4 > 0.0f) ? s4 : 100.0f;
    double t_deact = (s5 > 0.0f) ? s5 : 60.0f;
    double pip2 = (s6 >= 0.0f) ? s6 : 1.0f;
    float mi = 1.0f / (1.0f + bc::exp(-(V - vh) / ks));
    float tau_m = t_deact + (t_act - t_deact) * mi;
    if (tau_m < 1.0f) tau_m = 1.0f;
    mv += (mi - mv) * dt / tau_m;
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m.set(0, mv);
    if (pip2 < 0.01f) return 0.0f;
    float ki = s.in[1], ko = s.ex[1];
    double P = p_single(7.0f, 1.0f, kCbarK);
    float sup = 1.0f - 0.8f * s.pkc;
    if (sup < 0.0f) sup = 0.0f;
    double flux = ghk_mono(s, V, ki, ko);
    double ia = P * flux * (m.dens * area) * (mv * pip2 * sup);
    return k_out(s, ia * kAtoPa, dt);
}

// HCN1 (129: V1/2 -69 mV, tau 84 ms, 0.84 pS) and HCN2 (kHcn2, 130: -95 mV, 557 ms, 1.54 pS)
// (modifier_helpers.h:1442-1484, hh_biological.cpp:3598-3599)
template <int kHcn = 1>   // 1-4: HCN1-HCN4 (hh_biological.cpp:3598-3601, modifier_helpers.h:1064-1079)
BC_DEV double ch_hcn1(const ModSlot& m, SegRegs& s, float dt) {
    constexpr float kVh = kHcn == 1 ? -69.0f : kHcn == 2 ? -95.0f : kHcn == 3 ? -77.0f : -100.0f;
    constexpr float kTopen = kHcn == 1 ? 84.0f : kHcn == 2 ? 557.0f : kHcn == 3 ? 1408.0f : 3586.0f;
    constexpr float kGamma = kHcn == 1 ? 0.84f : kHcn == 2 ? 1.54f : kHcn == 3 ? 0.54f : 0.51f;
    double V = s.V;
    float area = eff_area(m, s);
    double s1 = m.get(1), s2 = m.get(2), s3 = m.get(3), s4 = m.get(4), s5 = m.get(5);
    float vh = (s1 != 0) ? s1 : kVh;
    float k = (s2 > 0) ? s2 : 6.0f;
    float gamma = (s3 > 0) ? s3 : kGamma;
    float t_open = (s4 > 0) ? s4 : kTopen;
    float t_close = (s5 > 0) ? s5 : 300.0f;
    float camp = s.camp;
    float ve = vh + 15.0f * camp / (camp + 0.5f);
    float mv = m.get(0);
    float vf = V;
    float mi = 1.0f / (1.0f + bc::exp((vf - ve) / k));
    float tau = (V < ve) ? (t_open * bc::exp((V - ve) / 15.0f)) : t_close;
    mv += (mi - mv) * dt / tau;
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m.set(0, mv);
    const P3 pp = (s3 > 0) ? p_partial3(gamma, 0.2f, 0.8f, 0.0f) : p_partial3(kGamma, 0.2f, 0.8f, 0.0f);
    double pna = pp.na, pk = pp.k;
    double N = m.dens * area;
    double fna = ghk_mono(s, V, s.in[0], s.ex[0]);
    double fk = ghk_mono(s, V, s.in[1], s.ex[1]);
    double in = pna * fna * N * mv * kAtoPa;
    double ik = pk * fk * N * mv * kAtoPa;
    double cpm = s.cpm;
    s.in[0] -= (in * dt) / cpm;
    s.in[1] -= (ik * dt) / cpm;
    s.ex[0] += (in * dt) / (cpm * kEvr);
    s.ex[1] += (ik * dt) / (cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -(in + ik);
}

BC_DEV double ch_nalcn(const ModSlot& m, SegRegs& s, float dt) {   // hh_biological.cpp:3609-3641
    float V = s.V;
    float area = eff_area(m, s);
    double s0 = m.get(0);
    double gamma = (s0 > 0.0f) ? s0 : 27.0f;
    const P3 pp = (s0 > 0.0f) ? p_nalcn3(gamma) : p_nalcn3(27.0f);
    double pna = pp.na, pk = pp.k, pca = pp.ca;
    double N = m.dens * area;
    float ina = (float)(pna * ghk_mono(s, V, s.in[0], s.ex[0]) * N * kAtoPa);
    float ik = (float)(pk * ghk_mono(s, V, s.in[1], s.ex[1]) * N * kAtoPa);
    float ica = (float)(pca * ghk_di(s, V, s.in[2], s.ex[2]) * N * kAtoPa);
    double cpm = s.cpm;
    s.in[0] -= (ina * dt) / cpm;
    s.in[1] -= (ik * dt) / cpm;
    s.in[2] -= (ica * dt) / (2.0f * cpm);
    s.ex[0] += (ina * dt) / (cpm * kEvr);
    s.ex[1] += (ik * dt) / (cpm * kEvr);
    s.ex[2] += (ica * dt) / (2.0f * cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -(ina - ik + ica);
}

// Ionotropic receptor (modifier_helpers.h:1213-1276).
struct IonoK { float gamma, ec50, t_rise, t_decay; int nt; float pna, pk, pca; bool anion; float t_desens; };

// Transmitter k of the segment (k a constant after inlining: the array stays in registers).
BC_DEV float nt_pick(const SegRegs& s, int k) {
    switch (k) {
        case 0: return s.nt[0];
        case 1: return s.nt[1];
        case 2: return s.nt[2];
        case 3: return s.nt[3];
        case 4: return s.nt[4];
        case 5: return s.nt[5];
        case 6: return s.nt[6];
        case 7: return s.nt[7];
        case 8: return s.nt[8];
        case 9: return s.nt[9];
        case 11: return s.nt[11];
        default: return s.nt[10];
    }
}
// state[k] of a modifier for k <= 7: state[7] is the effective area the caller set (a register
// here); larger k, past the record's state array, are refused at load.
BC_DEV double state_any(const ModSlot& m, uint32_t k) { return k < 7 ? m.get(k) : (double)m.a7; }

// scale: the reference's mod_scale factor of the current (modifier_helpers.h:1216), 1 except
// for the MSN receptors 75/76.
template <bool kSegV = false>
BC_DEV double iono(const ModSlot& m, SegRegs& s, float dt, const IonoK& p, float scale = 1.0f) {
    float area = eff_area(m, s);
    double V = s.V;
    double cpm = s.cpm;
    float gamma = p.gamma;
    if (m.av > 0) {
        double g = state_any(m, m.av);
        if (g > 0.0f) gamma = g;
    }
    float nt = nt_pick(s, p.nt);
    float r_inf = nt / (nt + p.ec50);
    double s0 = m.get(0);
    float r = s0;
    float tau = (r_inf > r) ? p.t_rise : p.t_decay;
    s0 += (r_inf - r) * dt / tau;
    if (s0 < 0.0f) s0 = 0.0f;
    if (s0 > 1.0f) s0 = 1.0f;
    m.set(0, s0);
    float op = s0;
    if (p.t_desens > 0.0f) {
        float kd = p.ec50;
        float d = m.get(1);
        float d_inf = nt / (nt + kd);
        d += (d_inf - d) * dt / p.t_desens;
        if (d < 0.0f) d = 0.0f;
        if (d > 1.0f) d = 1.0f;
        m.set(1, d);
        op *= (1.0f - d);
    }
    double N = m.dens * area;
    if (p.anion) {
        double pcl = (gamma == p.gamma) ? p_single(p.gamma, 1.0f, kCbarCl) : p_single(gamma, 1.0f, kCbarCl);
        double fcl = ghk_an<kSegV>(s, V, s.in[3], s.ex[3]);
        double icl = pcl * fcl * N * op * scale * kAtoPa;
        s.in[3] += (icl * dt) / cpm;
        floor7(s.in);
        return -icl;
    }
    const P3 pp = (gamma == p.gamma) ? p_partial3(p.gamma, p.pna, p.pk, p.pca)
                                     : p_partial3(gamma, p.pna, p.pk, p.pca);
    double pna = pp.na, pk = pp.k, pca = pp.ca;
    double ina = pna * ghk_mono<kSegV>(s, V, s.in[0], s.ex[0]) * N * op * scale * kAtoPa;
    double ik = pk * ghk_mono<kSegV>(s, V, s.in[1], s.ex[1]) * N * op * scale * kAtoPa;
    double itot = ina + ik;
    s.in[0] -= (ina * dt) / cpm;
    s.in[1] -= (ik * dt) / cpm;
    if (p.pca > 0.0f) {
        double ica = pca * ghk_di<kSegV>(s, V, s.in[2], s.ex[2]) * N * op * scale * kAtoPa;
        itot += ica;
        s.in[2] -= (ica * dt) / (2.0f * cpm);
    }
    floor7(s.in);
    return -itot;
}

// NMDA (9) and the GluN2B NMDA (k2b, 194: rise 10 ms, decay 150 ms).
template <bool kSegV = false, bool k2b = false>   // kSegV: also requires mg_cache_init
BC_DEV double rec_nmda(const ModSlot& m, SegRegs& s, float dt) {   // receptors_biological.cpp:48-62
    const IonoK k = {50.0f, 1.0f, k2b ? 10.0f : 5.0f, k2b ? 150.0f : 50.0f, 0, 0.08f, 0.08f, 0.84f, false, 0.0f};
    double I = iono<kSegV>(m, s, dt, k);
    double emg = (kSegV || s.V == s.vkmg) ? s.emg : bc::exp(-0.062f * s.V);
    float mg = 1.0f / (1.0f + (1.0f / 3.57f) * emg);
    double r = I * mg;
    // nmda_cdi_coagonist (modifier_helpers.h:1281-1291)
    double ca_um = s.in[2] * kMmToUm;
    double cdi_inf = 1.0 - 0.8f * ca_um / (ca_um + 1.0f);
    double s2 = m.get(2);
    float tau = (cdi_inf < s2) ? 500.0f : 9000.0f;
    s2 += (cdi_inf - s2) * dt / tau;
    if (s2 < 0.0f) s2 = 0.0f;
    if (s2 > 1.0f) s2 = 1.0f;
    m.set(2, s2);
    float coag = 10.0f / (10.0f + 1.0f);
    return r * (s2 * coag);
}

// Receptor-slot dispatch when every receptor in the network is AMPA, NMDA or GABA-A; V is the
// segment's V and both exp caches are initialised. Floors ions 0-3 only: the caller floors
// ions 4-6 (untouched here) once after its walk, which floor7 per call amounts to. One body with the three parameter sets
// selected per lane (the same operations as iono<true> with the constants of
// dispatch(): AMPA and NMDA lanes of a wavefront run it together.
// Receptor slot whose states 0-2 were loaded ahead (the post-side walk issues every load of a
// chain step at once); a state beyond 2 (a gamma override index) is read from memory.
struct RegSlot {
    double st[3];        // states 0-2, loaded with the chain step; the rest are read on use
    const double* mem;   // state 0 of this receptor, stride `stride`
    size_t stride;
    double dens;
    float a7;
    uint8_t av;
    __device__ double get(int k) const {
        if (k == 0) return st[0];
        if (k == 1) return st[1];
        if (k == 2) return st[2];
        return mem[(size_t)k * stride];   // an active_vars parameter state
    }
    __device__ void set(int k, double v) {
        if (k == 0) st[0] = v;
        else if (k == 1) st[1] = v;
        else st[2] = v;   // the receptors below set states 0, 1 (AMPA, GABA-A) and 2 (NMDA)
    }
};

template <class Slot>
BC_DEV double dispatch_receptor(uint8_t type, Slot& m, SegRegs& s, float dt) {
    const bool nmda = type == M_NMDA, gaba = type == M_GABAA;
    const float gamma0 = nmda ? 50.0f : (gaba ? 28.0f : 8.0f);
    const float ec50 = gaba || nmda ? 1.0f : 0.5f;
    const float t_rise = nmda ? 5.0f : (gaba ? 1.0f : 0.5f);
    const float t_decay = nmda ? 50.0f : (gaba ? 5.0f : 3.0f);
    const float t_desens = nmda ? 0.0f : (gaba ? 2000.0f : 3.0f);
    float area = eff_area(m, s);
    double V = s.V;
    double cpm = s.cpm;
    float gamma = gamma0;
    if (m.av > 0 && m.av < 7) {
        double g = m.get(m.av);
        if (g > 0.0f) gamma = g;
    }
    float nt = gaba ? s.nt[1] : s.nt[0];
    float r_inf = nt / (nt + ec50);
    double s0 = m.get(0);
    float r = s0;
    float tau = (r_inf > r) ? t_rise : t_decay;
    s0 += (r_inf - r) * dt / tau;
    if (s0 < 0.0f) s0 = 0.0f;
    if (s0 > 1.0f) s0 = 1.0f;
    m.set(0, s0);
    float op = s0;
    if (t_desens > 0.0f) {
        float d = m.get(1);
        float d_inf = nt / (nt + ec50);
        d += (d_inf - d) * dt / t_desens;
        if (d < 0.0f) d = 0.0f;
        if (d > 1.0f) d = 1.0f;
        m.set(1, d);
        op *= (1.0f - d);
    }
    double N = m.dens * area;
    const float scale = 1.0f;
    if (gaba) {
        constexpr double kPcl = p_single(28.0f, 1.0f, kCbarCl);
        double pcl = (gamma == 28.0f) ? kPcl : p_single(gamma, 1.0f, kCbarCl);
        double fcl = ghk_an<true>(s, V, s.in[3], s.ex[3]);
        double icl = pcl * fcl * N * op * scale * kAtoPa;
        s.in[3] += (icl * dt) / cpm;
        floor_n<4>(s.in);   // ions 4-6: floored by the caller once after its walk
        return -icl;
    }
    constexpr P3 kPA = p_partial3(8.0f, 0.5f, 0.5f, 0.0f);
    constexpr P3 kPN = p_partial3(50.0f, 0.08f, 0.08f, 0.84f);
    const float fna = nmda ? 0.08f : 0.5f, fca = nmda ? 0.84f : 0.0f;
    P3 pp;
    if (gamma == gamma0) pp = nmda ? kPN : kPA;
    else pp = p_partial3(gamma, fna, fna, fca);
    double ina = pp.na * ghk_mono<true>(s, V, s.in[0], s.ex[0]) * N * op * scale * kAtoPa;
    double ik = pp.k * ghk_mono<true>(s, V, s.in[1], s.ex[1]) * N * op * scale * kAtoPa;
    double itot = ina + ik;
    s.in[0] -= (ina * dt) / cpm;
    s.in[1] -= (ik * dt) / cpm;
    if (nmda) {   // pca > 0
        double ica = pp.ca * ghk_di<true>(s, V, s.in[2], s.ex[2]) * N * op * scale * kAtoPa;
        itot += ica;
        s.in[2] -= (ica * dt) / (2.0f * cpm);
    }
    floor_n<4>(s.in);
    double I = -itot;
    if (!nmda) return I;
    // Mg2+ block and nmda_cdi_coagonist, as rec_nmda
    float mg = 1.0f / (1.0f + (1.0f / 3.57f) * s.emg);
    double rr = I * mg;
    double ca_um = s.in[2] * kMmToUm;
    double cdi_inf = 1.0 - 0.8f * ca_um / (ca_um + 1.0f);
    double s2 = m.get(2);
    float tau2 = (cdi_inf < s2) ? 500.0f : 9000.0f;
    s2 += (cdi_inf - s2) * dt / tau2;
    if (s2 < 0.0f) s2 = 0.0f;
    if (s2 > 1.0f) s2 = 1.0f;
    m.set(2, s2);
    float coag = 10.0f / (10.0f + 1.0f);
    return rr * (s2 * coag);
}

// Connexins Cx36 (114) and Cx43 (198) (hh_biological.cpp:3266-3300, 3310-3341; the two bodies
// are the same): connexon turnover over hours, CaMKII/PKA phosphorylation over seconds and a
// pH-gated open probability; no current. Gap junctions read state[0] and state[4].
BC_DEV double ch_connexin(const ModSlot& m, SegRegs& s, float dt) {
    double synth = (m.get(2) > 0.0f) ? m.get(2) : 0.001f;
    double degr = (m.get(3) > 0.0f) ? m.get(3) : 0.0005f;
    double pH = (m.get(5) > 4.0f) ? m.get(5) : 7.2f;
    double n = m.get(0);
    n += (synth - degr * n) * dt / 3600000.0f;
    if (n < 0.0f) n = 0.0f;
    if (n > 1.0f) n = 1.0f;
    m.set(0, n);
    double pka = m.get(6);
    float p_target = s.camkii * 0.6f + pka * 0.3f;
    double ph = m.get(1);
    ph += (p_target - ph) * dt / 5000.0f;
    if (ph < 0.0f) ph = 0.0f;
    if (ph > 1.0f) ph = 1.0f;
    m.set(1, ph);
    float ph_f = (pH < 6.8f) ? (pH - 6.0f) / 0.8f : 1.0f;
    if (ph_f < 0.0f) ph_f = 0.0f;
    double p_cam = 1.0f - ph * 0.5f;
    double p_pka = 1.0f + ph * 0.3f;
    double po = (p_cam * 0.5f + p_pka * 0.5f) * ph_f;
    if (po < 0.05f) po = 0.05f;
    if (po > 2.0f) po = 2.0f;
    m.set(4, po);
    return 0.0;
}

// Cav1.3 (1, hh_biological.cpp:142-185): L-type m^2 h, fixed time constants, Ca-dependent h;
// the extracellular Ca falls with the intracellular, as the reference writes it.
BC_DEV double ch_cav13(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2), s3 = m.get(3), s4 = m.get(4);
    double vh = (s3 != 0.0f) ? -fabs(s3) : -35.0f;
    double k = (s4 > 0.0f) ? s4 : 7.0f;
    double gamma = (s2 > 0.0f) ? s2 : 25.0f;
    float mi = 1.0f / (1.0f + bc::exp(-(V - vh) / (k + 1e-6f)));
    double g0 = m.get(0), g1 = m.get(1);
    g0 += (mi - g0) * dt / 2.0f;
    float ca = s.in[2] * kMmToUm;
    float hi = 1.0f / (1.0f + bc::exp((V + 52.0f) / 8.0f));
    hi *= 1.0f / (1.0f + ca / 3.0f);
    g1 += (hi - g1) * dt / 100.0f;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g1 < 0.0f) g1 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    if (g1 > 1.0f) g1 = 1.0f;
    m.set(0, g0);
    m.set(1, g1);
    float ci = s.in[2], co = s.ex[2];
    double P = p_single(gamma, 2.0f, kCbarCa);
    double op = g0 * g0 * g1;
    double flux = ghk_di(s, V, ci, co);
    double ica = P * flux * (m.dens * area) * op * kAtoPa;
    double cpm = s.cpm;
    s.in[2] -= (ica * dt) / (2.0f * cpm);
    s.ex[2] -= (ica * dt) / (2.0f * cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -ica;
}

// ClC-2 leak (2, hh_biological.cpp:390-414): state[0] forced to 1 below 0.5, GHK anion current;
// the charge-per-mole factor is a float in this one.
BC_DEV double ch_cl(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double g0 = m.get(0);
    if (g0 < 0.5f) { g0 = 1.0f; m.set(0, g0); }
    double s2 = m.get(2);
    double gamma = (s2 > 0.0f) ? s2 : 2.5f;
    float ci = s.in[3], co = s.ex[3];
    double P = p_single(gamma, 1.0f, kCbarCl);
    double flux = ghk_an(s, V, ci, co);
    double icl = P * flux * (m.dens * area) * g0 * kAtoPa;
    float vol = s.vol;
    float cpm = kFaraday * vol / 1000.0;
    s.in[3] += (icl * dt) / cpm;
    s.ex[3] -= (icl * dt) / (cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -icl;
}

// Kv4.2 A-type (85, hh_biological.cpp:789-836, modifier_helpers.h:463-472): m^3 h with V-dependent
// recovery; CaMKII (state[0] + state[1] of the last CaMKII modifier) phosphorylates state[3].
BC_DEV double ch_kv42(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2), s4 = m.get(4), s5 = m.get(5);
    double vh = (s4 != 0.0f) ? -fabs(s4) : -10.0f;
    double k = (s5 > 0.0f) ? s5 : 13.5f;
    double gamma = (s2 > 0.0f) ? s2 : 5.0f;
    float mi = 1.0f / (1.0f + bc::exp(-(V - vh) / (k + 1e-6f)));
    double g0 = m.get(0), g1 = m.get(1), g3 = m.get(3);
    g0 += (mi - g0) * dt / 1.5f;
    float hi = 1.0f / (1.0f + bc::exp((V + 65.0f) / 8.0f));
    float th = 15.0f + (170.0f - 15.0f) / (1.0f + bc::exp((V - -60.0f) / 8.0f));
    g1 += (hi - g1) * dt / th;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g1 < 0.0f) g1 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    if (g1 > 1.0f) g1 = 1.0f;
    g3 += (s.camkii01 * 0.5f - g3) * dt / 5000.0f;
    if (g3 < 0.0f) g3 = 0.0f;
    if (g3 > 1.0f) g3 = 1.0f;
    m.set(0, g0);
    m.set(1, g1);
    m.set(3, g3);
    double geff = gamma * (1.0f - g3 * 0.7f);
    float ki = s.in[1], ko = s.ex[1];
    double P = p_single(geff, 1.0f, kCbarK);
    double op = g0 * g0 * g0 * g1;
    double flux = ghk_mono(s, V, ki, ko);
    return k_out(s, P * flux * (m.dens * area) * op * kAtoPa, dt);
}

// Kv2.1 delayed rectifier (196, hh_biological.cpp:1392-1428, modifier_helpers.h:584-620): m^4 h
// with a U-shaped inactivation, TEA block by state[5].
BC_DEV double ch_kv21(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2), s5 = m.get(5);
    double gamma = (s2 > 0.0f) ? s2 : 8.0f;
    double g0 = m.get(0), g1 = m.get(1);
    float mi = 1.0f / (1.0f + bc::exp(-(V - 0.0f) / 9.4f));
    float tm = 15.0f + (73.0f - 15.0f) / (1.0f + bc::exp((V - 0.0f) / 10.0f));
    g0 += (mi - g0) * dt / tm;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    float left = 1.0f / (1.0f + bc::exp(-(V - -26.5f) / 5.2f));
    float right = 1.0f / (1.0f + bc::exp((V - 15.0f) / 10.0f));
    float hi = 1.0f - (1.0f - 0.2f) * left * right;
    float th = 4000.0f + (300.0f - 4000.0f) / (1.0f + bc::exp((V - -80.0f) / 15.0f));
    g1 += (hi - g1) * dt / th;
    if (g1 < 0.0f) g1 = 0.0f;
    if (g1 > 1.0f) g1 = 1.0f;
    m.set(0, g0);
    m.set(1, g1);
    double op = g0 * g0 * g0 * g0 * g1;
    double tea = (s5 > 0.0f) ? s5 : 0.0f;
    op /= (1.0f + tea / 10.0f);
    float ki = s.in[1], ko = s.ex[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    return k_out(s, P * flux * (m.dens * area) * op * kAtoPa, dt);
}

// Persistent Na+: Nav1.6 NaP (105, hh_biological.cpp:2951-2978) and the tadpole NaP (kTadpole,
// 195: Nav1.2 gamma, V1/2 -40 mV, slow inactivation in state[6], Na+ drawn from the extracellular
// pool, hh_biological.cpp:2980-3016).
template <bool kTadpole>
BC_DEV double ch_nap(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double mv = m.get(0);
    double s1 = m.get(1), s2 = m.get(2), s3 = m.get(3), s4 = m.get(4), s5 = m.get(5);
    double vh = (s1 != 0.0f) ? s1 : (kTadpole ? -40.0f : -58.7f);
    double km = (s2 > 0.0f) ? s2 : 6.0f;
    double gamma = (s3 > 0.0f) ? s3 : (kTadpole ? 20.8f : 20.7f);
    double tau = (s4 > 0.0f) ? s4 : 5.0f;
    double po = (s5 > 0.0f) ? s5 : 0.016f;
    float vhf = vh, kmf = km;
    float mi = 1.0f / (1.0f + bc::exp(-(V - vhf) / kmf));
    mv += (mi - mv) * dt / tau;
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m.set(0, mv);
    float nai = s.in[0], nao = s.ex[0];
    double P = p_single(gamma, 1.0f, kCbarNa);
    double flux = ghk_mono(s, V, nai, nao);
    if (!kTadpole) {
        float ina = (float)(P * flux * (m.dens * area) * (mv * po) * kAtoPa);
        s.in[0] -= (ina * dt) / s.cpm;
        floor7(s.in);
        return -ina;
    }
    float si = 1.0f / (1.0f + bc::exp((V - -50.0f) / 6.0f));
    double sv = m.get(6);
    sv += (si - sv) * dt / 500.0f;
    if (sv < 0.0f) sv = 0.0f;
    if (sv > 1.0f) sv = 1.0f;
    m.set(6, sv);
    double ina = P * flux * (m.dens * area) * (mv * po * sv) * kAtoPa;
    s.in[0] -= (ina * dt) / s.cpm;
    s.ex[0] += (ina * dt) / (s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -ina;
}

// Kir2.1 (106, hh_biological.cpp:3018-3037, modifier_helpers.h:870-891): Woodhull Mg2+ block and
// a three-site spermine block.
BC_DEV float kir_f_open(float V) {
    float vn = kFaraday * V * 0.001f / (kGas * kTemp);
    float kd_mg = 0.3f * bc::exp(-0.6f * 2.0f * vn);
    float mg = 0.5f / kd_mg;
    float kd_spm = 0.0775f * bc::exp(-0.5f * 4.0f * vn);
    float r = 0.001f / kd_spm;
    float spm = 0.0f, rk = r;
#pragma unroll
    for (int k = 0; k < 3; k++) { spm += rk; rk *= r; }
    return 1.0f / (1.0f + mg + spm);
}
BC_DEV double ch_kir(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s3 = m.get(3);
    double gamma = (s3 > 0.0f) ? s3 : 30.0f;
    float ki = s.in[1], ko = s.ex[1];
    float fo = kir_f_open(V);
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    return k_out(s, P * flux * (m.dens * area) * fo * kAtoPa, dt);
}

// Developmental Nav1.2 (193, hh_biological.cpp:314-357, modifier_helpers.h:221-241): m^3 with
// fast and slow inactivation (slow fraction state[4], default 0.7; slow recovery 500 ms), the
// inactivation V1/2 from state[2] (default 45: -45 mV), TTX and PKA/PKC block.
BC_DEV double ch_nav12dev(const ModSlot& m, SegRegs& s, float dt) {
    const NavK g = {16.0f, 6.0f, 41.5f, 66.5f, 45.0f};
    float area = eff_area(m, s);
    if (s.V < -200.0f) s.V = -200.0f;
    if (s.V > 200.0f) s.V = 200.0f;
    float V = s.V;
    float am = nav_am(g, V), bm = nav_bm(g, V);
    float minf = am / (am + bm);
    float mtau = 1.0f / (am + bm);
    double mv = minf + (m.get(0) - minf) * bc::exp(-dt / mtau);
    double s2 = m.get(2), s3 = m.get(3), s4 = m.get(4), s5 = m.get(5);
    double vhalf = (s2 > 0.0f) ? s2 : 45.0f;
    float vhf = vhalf;
    float hi = 1.0f / (1.0f + bc::exp((V + vhf) / 6.2f));
    float htau = 1.0f / (nav_ah(g, V) + nav_bh(g, V));
    double hf = hi + (m.get(1) - hi) * bc::exp(-dt / htau);
    double ts = (hi < s3) ? htau : 500.0f;
    double hs = hi + (s3 - hi) * bc::exp(-dt / ts);
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hf < 0.0f) hf = 0.0f;
    if (hf > 1.0f) hf = 1.0f;
    if (hs < 0.0f) hs = 0.0f;
    if (hs > 1.0f) hs = 1.0f;
    m.set(0, mv);
    m.set(1, hf);
    m.set(3, hs);
    double sf = (s4 > 0.0f) ? s4 : 0.7f;
    float he = (1.0f - sf) * hf + sf * hs;
    float op = mv * mv * mv * he;
    double ttx = (s5 > 0.0f) ? s5 : 0.0f;
    op /= (1.0f + ttx / 10.0f);
    op *= (1.0f - 0.7f * (s.pkc + s.pka) * 0.5f);
    if (op < 0.0f) op = 0.0f;
    double P = p_single(20.8f, 1.0f, kCbarNa);
    double flux = ghk_mono(s, V, s.in[0], s.ex[0]);
    double ina = P * flux * (m.dens * area) * op * kAtoPa;
    s.in[0] -= (ina * dt) / s.cpm;
    s.ex[0] += (ina * dt) / (s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -ina;
}

// Nav1.3 (177), Nav1.7 (169), Nav1.8 (170), Nav1.9 (171) (hh_biological.cpp nav13/17/18/19):
// m^3 h with Boltzmann steady states and fixed time constants, TTX block by state[5].
struct NavSimpleK { float vm, km, vh, kh, tm, th, ttx, gamma; };
BC_DEV double ch_nav_simple(const ModSlot& m, SegRegs& s, float dt, const NavSimpleK& g) {
    float V = s.V;
    float area = eff_area(m, s);
    float mi = 1.0f / (1.0f + bc::exp(-(V - g.vm) / g.km));
    float hi = 1.0f / (1.0f + bc::exp((V - g.vh) / g.kh));
    double g0 = m.get(0), g1 = m.get(1);
    g0 += (mi - g0) * (dt / g.tm);
    g1 += (hi - g1) * (dt / g.th);
    if (g0 < 0) g0 = 0;
    if (g0 > 1) g0 = 1;
    if (g1 < 0) g1 = 0;
    if (g1 > 1) g1 = 1;
    m.set(0, g0);
    m.set(1, g1);
    double op = g0 * g0 * g0 * g1;
    double s5 = m.get(5);
    double ttx = (s5 > 0.0f) ? s5 : 0.0f;
    op /= (1.0f + ttx / g.ttx);
    double P = p_single(g.gamma, 1.0f, kCbarNa);
    double flux = ghk_mono(s, V, s.in[0], s.ex[0]);
    float I = (float)(P * flux * (m.dens * area) * op * kAtoPa);
    s.in[0] -= (I * dt) / s.cpm;
    s.ex[0] += (I * dt) / (s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -I;
}

// Nav1.1 (199, hh_biological.cpp nav11_current, modifier_helpers.h:186-205): HH rates with the
// Nav1.1 offsets, inactivation slope 5 mV, TTX block; no PKA/PKC term.
BC_DEV double ch_nav11(const ModSlot& m, SegRegs& s, float dt) {
    const NavK g = {30.0f, 6.5f, 45.0f, 65.0f, 65.0f};
    float area = eff_area(m, s);
    if (s.V < -200.0f) s.V = -200.0f;
    if (s.V > 200.0f) s.V = 200.0f;
    float V = s.V;
    float am = nav_am(g, V), bm = nav_bm(g, V);
    float minf = am / (am + bm);
    float mtau = 1.0f / (am + bm);
    float hinf = 1.0f / (1.0f + bc::exp((V + 65.0f) / 5.0f));
    float htau = 1.0f / (nav_ah(g, V) + nav_bh(g, V));
    const double s0 = m.get(0), s1 = m.get(1), s5 = m.get(5);   // all loads before the stores
    double mv = minf + (s0 - minf) * bc::exp(-dt / mtau);
    double hv = hinf + (s1 - hinf) * bc::exp(-dt / htau);
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m.set(0, mv);
    m.set(1, hv);
    float op = mv * mv * mv * hv;
    double ttx = (s5 > 0.0f) ? s5 : 0.0f;
    op /= (1.0f + ttx / 10.0f);
    double P = p_single(17.0f, 1.0f, kCbarNa);
    double flux = ghk_mono(s, V, s.in[0], s.ex[0]);
    double ina = P * flux * (m.dens * area) * op * kAtoPa;
    s.in[0] -= (ina * dt) / s.cpm;
    s.ex[0] += (ina * dt) / (s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -ina;
}

// Kv1.2 (157), Kv1.3 (154), Kv1.4 (206), Kv1.6 (207), Kv2.2 (205) (hh_biological.cpp kv12/kv13/
// kv14/kv16/kv22, modifier_helpers.h:518-558, units.h:354-364): Kv1.2 and Kv1.3 m^4 h with the
// delayed-rectifier activation time constant (PKC block of Kv1.2; DTX/TEA and ShK toxin blocks),
// Kv1.4 m^3 h in ten substeps with a V-dependent recovery, Kv1.6 m^3, Kv2.2 m^3 h; pore blockers
// from the states.
template <bool kKv12>
BC_DEV double ch_kv1x(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    const float vh = kKv12 ? -27.0f : -30.0f, k = 6.0f, gamma = kKv12 ? 16.0f : 19.0f;
    float mi = 1.0f / (1.0f + bc::exp(-(V - vh) / (k + 1e-6f)));
    float tm = kdr_tau(V, vh);
    double g0 = m.get(0), g1 = m.get(1);
    g0 += (mi - g0) * dt / tm;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    float hi = 1.0f / (1.0f + bc::exp((V + 40.0f) / 6.0f));
    float th = kKv12 ? 400.0f : 3000.0f;
    g1 += (hi - g1) * dt / th;
    if (g1 < 0.0f) g1 = 0.0f;
    if (g1 > 1.0f) g1 = 1.0f;
    m.set(0, g0);
    m.set(1, g1);
    double op = g0 * g0 * g0 * g0 * g1;
    if (kKv12) {
        op *= (1.0f - 0.8f * s.pkc);
        if (op < 0.0f) op = 0.0f;
        double s5 = m.get(5), s6 = m.get(6);
        double dtx = (s5 > 0.0f) ? s5 : 0.0f;
        double tea = (s6 > 0.0f) ? s6 : 0.0f;
        op /= (1.0f + dtx / 5.0f) * (1.0f + tea / 170.0f);
    } else {
        double s5 = m.get(5);
        double shk = (s5 > 0.0f) ? s5 : 0.0f;
        op /= (1.0f + shk / 11.0f);
    }
    float ki = s.in[1], ko = s.ex[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    return k_out(s, P * flux * (m.dens * area) * op * kAtoPa, dt);
}
BC_DEV double ch_kv14(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2), s3 = m.get(3), s4 = m.get(4);
    double gamma = (s2 > 0.0f) ? s2 : 9.3f;
    double mv = m.get(0), hv = m.get(1);
    float mi = 1.0f / (1.0f + bc::exp(-(V - -21.7f) / 8.0f));
    float hi = 1.0f / (1.0f + bc::exp((V - -51.8f) / 3.2f));
    float tm = 2.0f;
    float th = 50.0f + (1500.0f - 50.0f) / (1.0f + bc::exp((V - -80.0f) / 10.0f));
    float sub = dt / 10.0f;
    for (int k = 0; k < 10; k++) {
        mv += (mi - mv) * sub / tm;
        hv += (hi - hv) * sub / th;
    }
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m.set(0, mv);
    m.set(1, hv);
    float op = mv * mv * mv * hv;
    double ap4 = (s3 > 0.0f) ? s3 : 0.0f;
    double tea = (s4 > 0.0f) ? s4 : 0.0f;
    op /= (1.0f + ap4 / 13.0f) * (1.0f + tea / 100.0f);
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, s.in[1], s.ex[1]);
    return k_out(s, P * flux * (m.dens * area) * op * kAtoPa, dt);
}
template <bool kKv22>
BC_DEV double ch_kv16_22(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2), s3 = m.get(3), s4 = m.get(4);
    double gamma = (s2 > 0.0f) ? s2 : (kKv22 ? 14.5f : 8.2f);
    double mv = m.get(0);
    float mi = kKv22 ? 1.0f / (1.0f + bc::exp(-(V - 13.2f) / 9.0f)) : 1.0f / (1.0f + bc::exp(-(V - -8.1f) / 13.0f));
    mv += (mi - mv) * dt / (kKv22 ? 32.0f : 20.0f);
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m.set(0, mv);
    float op;
    if (kKv22) {
        double hv = m.get(1);
        float hi = 1.0f / (1.0f + bc::exp((V - -16.6f) / 6.0f));
        hv += (hi - hv) * dt / 2000.0f;
        if (hv < 0.0f) hv = 0.0f;
        if (hv > 1.0f) hv = 1.0f;
        m.set(1, hv);
        op = mv * mv * mv * hv;
    } else {
        op = mv * mv * mv;
    }
    double tea = (s3 > 0.0f) ? s3 : 0.0f;
    double ap4 = (s4 > 0.0f) ? s4 : 0.0f;
    op /= (1.0f + tea / (kKv22 ? 2.6f : 1.7f)) * (1.0f + ap4 / (kKv22 ? 1.5f : 0.3f));
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, s.in[1], s.ex[1]);
    return k_out(s, P * flux * (m.dens * area) * op * kAtoPa, dt);
}

// Kv3.2 (203), Kv3.3 (204), Kv4.3 (202) (hh_biological.cpp kv32/kv33/kv43, modifier_helpers.h:
// 475-515): m^3 (h) with Boltzmann steady states, Kv4.3 in ten substeps with a V-dependent
// recovery; two pore blockers from state[3] and state[4].
struct KvGenK { float vm, km, vh, kh, tm, th, gamma, ic3, ic4; bool inact, sub; };
BC_DEV double ch_kvgen(const ModSlot& m, SegRegs& s, float dt, const KvGenK& g) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2), s3 = m.get(3), s4 = m.get(4);
    double gamma = (s2 > 0.0f) ? s2 : g.gamma;
    double mv = m.get(0), hv = m.get(1);
    float mi = 1.0f / (1.0f + bc::exp(-(V - g.vm) / g.km));
    float op;
    if (!g.inact) {
        mv += (mi - mv) * dt / g.tm;
        if (mv < 0.0f) mv = 0.0f;
        if (mv > 1.0f) mv = 1.0f;
        m.set(0, mv);
        op = mv * mv * mv;
    } else {
        float hi = 1.0f / (1.0f + bc::exp((V - g.vh) / g.kh));
        if (g.sub) {
            float th = 27.6f + (142.0f - 27.6f) / (1.0f + bc::exp((V - -60.0f) / 8.0f));
            float sub = dt / 10.0f;
            for (int k = 0; k < 10; k++) {
                mv += (mi - mv) * sub / g.tm;
                hv += (hi - hv) * sub / th;
            }
        } else {
            mv += (mi - mv) * dt / g.tm;
            hv += (hi - hv) * dt / g.th;
        }
        if (mv < 0.0f) mv = 0.0f;
        if (mv > 1.0f) mv = 1.0f;
        if (hv < 0.0f) hv = 0.0f;
        if (hv > 1.0f) hv = 1.0f;
        m.set(0, mv);
        m.set(1, hv);
        op = mv * mv * mv * hv;
    }
    double b3 = (s3 > 0.0f) ? s3 : 0.0f;
    double b4 = (s4 > 0.0f) ? s4 : 0.0f;
    op /= (1.0f + b3 / g.ic3) * (1.0f + b4 / g.ic4);
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, s.in[1], s.ex[1]);
    return k_out(s, P * flux * (m.dens * area) * op * kAtoPa, dt);
}

// Leak-like K+ channels (hh_biological.cpp girk/katp/thik1/thik2/kir22): GIRK (160) open as
// G-beta-gamma state[4], tertiapin block state[5]; KATP (161) closed by ATP state[4],
// glibenclamide state[5]; THIK-1 (166) Po 0.5 with halothane block state[5]; THIK-2 (167)
// Po 0.5; Kir2.2 (208) Woodhull block with Ba2+ state[4].
template <int T>
BC_DEV double ch_kleak(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    float po;
    double gamma;
    if (T == 160) {
        double gbg = m.get(4);
        if (gbg < 0.0f) gbg = 0.0f;
        if (gbg > 1.0f) gbg = 1.0f;
        po = gbg;
        double s5 = m.get(5);
        double tq = (s5 > 0.0f) ? s5 : 0.0f;
        po /= (1.0f + tq / 8.0f);
        gamma = 35.0f;
    } else if (T == 161) {
        double atp = m.get(4);
        if (atp < 0.0f) atp = 0.0f;
        po = 1.0f / (1.0f + atp / 15.0f);
        double s5 = m.get(5);
        double gl = (s5 > 0.0f) ? s5 : 0.0f;
        po /= (1.0f + gl / 3.0f);
        gamma = 75.0f;
    } else if (T == 166 || T == 167) {
        po = 0.5f;
        if (T == 166) {
            double s5 = m.get(5);
            double hal = (s5 > 0.0f) ? s5 : 0.0f;
            po /= (1.0f + hal / 2.8f);
        }
        gamma = 5.0f;
    } else {
        double s3 = m.get(3), s4 = m.get(4);
        gamma = (s3 > 0.0f) ? s3 : 36.0f;
        po = kir_f_open(V);
        double ba = (s4 > 0.0f) ? s4 : 0.0f;
        po /= (1.0f + ba / 2.8f);
    }
    double P = p_single(gamma, 1.0f, kCbarK);
    // Kir2.2 passes the concentrations rounded to float
    double flux = (T == 208) ? ghk_mono(s, V, (float)s.in[1], (float)s.ex[1]) : ghk_mono(s, V, s.in[1], s.ex[1]);
    return k_out(s, P * flux * (m.dens * area) * po * kAtoPa, dt);
}

// Kir7.1 (178, always open, 0.05 pS) and KCNQ4 (kKcnq4, 176: m with tau 212 ms, 2.31 pS)
// (hh_biological.cpp kir71/kcnq4): float current, both ion arrays floored.
template <bool kKcnq4>
BC_DEV double ch_kir71_kcnq4(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double op = 1.0f;
    if (kKcnq4) {
        float mi = 1.0f / (1.0f + bc::exp(-(V - (-78.0f)) / 14.0f));
        double g0 = m.get(0);
        g0 += (mi - g0) * (dt / 212.0f);
        if (g0 < 0) g0 = 0;
        if (g0 > 1) g0 = 1;
        m.set(0, g0);
        op = g0;
    }
    double P = p_single(kKcnq4 ? 2.31f : 0.05f, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, s.in[1], s.ex[1]);
    float I = (float)(P * flux * (m.dens * area) * op * kAtoPa);
    s.in[1] -= (I * dt) / s.cpm;
    s.ex[1] += (I * dt) / (s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -I;
}

// Cav2.2 N-type (168, hh_biological.cpp cav22_current): m^2 h, omega-conotoxin GVIA block
// state[5]; the intracellular Ca changes by the full current (no factor 2).
BC_DEV double ch_cav22(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double g0 = m.get(0), g1 = m.get(1), s5 = m.get(5);
    float mi = 1.0f / (1.0f + bc::exp(-(V - -20.0f) / (6.0f + 1e-6f)));
    g0 += (mi - g0) * dt / 1.7f;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    float hi = 1.0f / (1.0f + bc::exp((V + 60.0f) / 6.0f));
    g1 += (hi - g1) * dt / 150.0f;
    if (g1 < 0.0f) g1 = 0.0f;
    if (g1 > 1.0f) g1 = 1.0f;
    m.set(0, g0);
    m.set(1, g1);
    double op = g0 * g0 * g1;
    double gvia = (s5 > 0.0f) ? s5 : 0.0f;
    op /= (1.0f + gvia / 0.5f);
    double P = p_single(2.76f, 2.0f, kCbarCa);
    double flux = ghk_di(s, V, s.in[2], s.ex[2]);
    double ica = P * flux * (m.dens * area) * op * kAtoPa;
    s.in[2] -= (ica * dt) / s.cpm;
    s.ex[2] += (ica * dt) / (s.cpm * kEvr);
    floor7(s.ex);
    floor7(s.in);
    return -ica;
}

// T-type Cav3.1 (200), Cav3.2 (201), Cav3.3 (191) (hh_biological.cpp:2033-2143, modifier_helpers.h:
// 649-737, 1032-1036): m^3 h in ten substeps, a bell-shaped activation time constant and a
// V-dependent recovery, Ni2+/mibefradil/TTA-P2/Z944 blocks from states 3-6; the intracellular Ca
// changes with a float charge-per-mole factor and is not floored.
struct CavTK { float vm, km, vh, kh, gamma, thf, ths, tmp, ni, mib; };
BC_DEV double ch_cavt(const ModSlot& m, SegRegs& s, float dt, const CavTK& g) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2), s3 = m.get(3), s4 = m.get(4), s5 = m.get(5), s6 = m.get(6);
    double gamma = (s2 > 0.0f) ? s2 : g.gamma;
    double mv = m.get(0), hv = m.get(1);
    float mi = 1.0f / (1.0f + bc::exp(-(V - g.vm) / g.km));
    float hi = 1.0f / (1.0f + bc::exp((V - g.vh) / g.kh));
    float dv = V - -20.0f;
    float gg = bc::exp(-(dv * dv) / (2.0f * 22.0f * 22.0f));
    float tm = 1.25f + (g.tmp - 1.25f) * gg;
    float th = g.thf + (g.ths - g.thf) / (1.0f + bc::exp((V - -65.0f) / 8.0f));
    float sub = dt / 10.0f;
    for (int k = 0; k < 10; k++) {
        mv += (mi - mv) * sub / tm;
        hv += (hi - hv) * sub / th;
    }
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m.set(0, mv);
    m.set(1, hv);
    float op = mv * mv * mv * hv;
    double ni = (s3 > 0.0f) ? s3 : 0.0f;
    double mib = (s4 > 0.0f) ? s4 : 0.0f;
    double tta = (s5 > 0.0f) ? s5 : 0.0f;
    double z = (s6 > 0.0f) ? s6 : 0.0f;
    op /= (1.0f + ni / g.ni) * (1.0f + mib / g.mib) * (1.0f + tta / 82.0f) * (1.0f + z / 100.0f);
    double P = p_single(gamma, 2.0f, kCbarCa);
    double flux = ghk_di(s, V, s.in[2], s.ex[2]);
    double ica = P * flux * (m.dens * area) * op * kAtoPa;
    float vol = s.vol;
    float cpm = kFaraday * vol / 1000.0;
    s.in[2] -= (ica * dt) / (2.0f * cpm);
    return -ica;
}

// TRPC3 (210, DAG state[4], Pyr3 block state[5]), TRPC4 (162, Gq state[4], ML204 block
// state[5]) and TRPM7 (211, closed by intracellular Mg2+, NS8593 block state[5])
// (hh_biological.cpp trpc3/trpc4/trpm7): Na+ and Ca2+ currents from partial permeabilities; the
// Ca2+ change uses the charge-per-mole factor without the valence.
template <int T>
BC_DEV double ch_trp_naca(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    float po;
    double s5 = m.get(5);
    if (T == 211) {
        float mg = s.in[4];
        float r = mg / 0.7f;
        po = 1.0f / (1.0f + r * r);
        double ns = (s5 > 0.0f) ? s5 : 0.0f;
        po /= (1.0f + ns / 1.6f);
    } else {
        double a = m.get(4);
        if (a < 0.0f) a = 0.0f;
        if (a > 1.0f) a = 1.0f;
        po = a;
        double b = (s5 > 0.0f) ? s5 : 0.0f;
        po /= (1.0f + b / (T == 210 ? 0.7f : 2.0f));
    }
    const P3 pp = p_partial3(T == 211 ? 40.0f : T == 210 ? 60.0f : 40.0f, 1.0f, 0.0f,
                             T == 211 ? 0.34f : T == 210 ? 1.62f : 4.0f);
    double N = m.dens * area;
    double fna = ghk_mono(s, V, s.in[0], s.ex[0]);
    double fca = ghk_di(s, V, s.in[2], s.ex[2]);
    double ina = pp.na * fna * N * po * kAtoPa;
    double ica = pp.ca * fca * N * po * kAtoPa;
    s.in[0] -= (ina * dt) / s.cpm;
    s.ex[0] += (ina * dt) / (s.cpm * kEvr);
    s.in[2] -= (ica * dt) / s.cpm;
    s.ex[2] += (ica * dt) / (s.cpm * kEvr);
    floor7(s.ex);
    floor7(s.in);
    return -(ina + ica);
}

// TRPV3 (175, hh_biological.cpp trpv3_current): ligand state[4] (EC50 50) drives the gate with
// tau 20 ms; float Na+ and Ca2+ currents.
BC_DEV double ch_trpv3(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s4 = m.get(4);
    double lig = (s4 > 0.0f) ? s4 : 0.0f;
    float r = lig / (lig + 50.0f);
    double g0 = m.get(0);
    g0 += (r - g0) * (dt / 20.0f);
    if (g0 < 0) g0 = 0;
    if (g0 > 1) g0 = 1;
    m.set(0, g0);
    const P3 pp = p_partial3(185.0f, 0.5f, 0.0f, 0.5f);
    double N = m.dens * area;
    double fca = ghk_di(s, V, s.in[2], s.ex[2]);
    double fna = ghk_mono(s, V, s.in[0], s.ex[0]);
    float ica = (float)(pp.ca * fca * N * g0 * kAtoPa);
    float ina = (float)(pp.na * fna * N * g0 * kAtoPa);
    double cpm = s.cpm;
    s.in[2] -= (ica * dt) / (2.0f * cpm);
    s.in[0] -= (ina * dt) / cpm;
    s.ex[2] += (ica * dt) / (2.0f * cpm * kEvr);
    s.ex[0] += (ina * dt) / (cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -(ica + ina);
}

// ENaC (182, hh_biological.cpp enac_current): Po 0.5, amiloride block state[5] (IC50 100),
// 4.5 pS; float current.
BC_DEV double ch_enac(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    float op = 0.5f;
    double s5 = m.get(5);
    double am = (s5 > 0.0f) ? s5 : 0.0f;
    op /= (1.0f + am / 100.0f);
    double P = p_single(4.5f, 1.0f, kCbarNa);
    double flux = ghk_mono(s, V, s.in[0], s.ex[0]);
    float I = (float)(P * flux * (m.dens * area) * op * kAtoPa);
    s.in[0] -= (I * dt) / s.cpm;
    s.ex[0] += (I * dt) / (s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -I;
}

// Ligand-gated TRP channels with float Na+/Ca2+ currents (hh_biological.cpp trpv1_channel/
// trpm8/trpa1/trpc67): the ligand state[4] sets the gate target (EC50 ec), a Ca-driven
// desensitisation in state[1] (TRPV1 172, TRPM8 173, TRPA1 174) and, for TRPA1, Ca2+
// co-activation; TRPC6/7 (179) has DAG gating only.
struct TrpLigK { float ec, tau, tdes, gamma, fna, fca; bool des, ca_act; };
BC_DEV double ch_trp_lig(const ModSlot& m, SegRegs& s, float dt, const TrpLigK& g) {
    float V = s.V;
    float area = eff_area(m, s);
    double s4 = m.get(4);
    double lig = (s4 > 0.0f) ? s4 : 0.0f;
    float r = lig / (lig + g.ec);
    double g0 = m.get(0), op;
    if (g.des) {
        double g1 = m.get(1);
        float ca = s.in[2] * kMmToUm;
        float ca_act = ca / (ca + 6.0f);
        float d_inf = ca / (ca + 0.5f);
        g0 += (r - g0) * (dt / g.tau);
        g1 += (d_inf - g1) * (dt / g.tdes);
        if (g0 < 0) g0 = 0;
        if (g0 > 1) g0 = 1;
        if (g1 < 0) g1 = 0;
        if (g1 > 1) g1 = 1;
        m.set(1, g1);
        double act = g.ca_act ? (double)fminf(1.0f, g0 + ca_act) : g0;
        op = act * (1.0f - g1);
    } else {
        g0 += (r - g0) * (dt / g.tau);
        if (g0 < 0) g0 = 0;
        if (g0 > 1) g0 = 1;
        op = g0;
    }
    m.set(0, g0);
    const P3 pp = p_partial3(g.gamma, g.fna, 0.0f, g.fca);
    double N = m.dens * area;
    double fca = ghk_di(s, V, s.in[2], s.ex[2]);
    double fna = ghk_mono(s, V, s.in[0], s.ex[0]);
    float ica = (float)(pp.ca * fca * N * op * kAtoPa);
    float ina = (float)(pp.na * fna * N * op * kAtoPa);
    double cpm = s.cpm;
    s.in[2] -= (ica * dt) / (2.0f * cpm);
    s.in[0] -= (ina * dt) / cpm;
    s.ex[2] += (ica * dt) / (2.0f * cpm * kEvr);
    s.ex[0] += (ina * dt) / (cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -(ica + ina);
}

// TRPC5 (163, Gq state[4] x intracellular Ca2+ co-activation, HC-070 block state[5]) and TRPM4
// (kM4, 164: Ca2+-activated Na+/K+ channel, 9-phenanthrol block state[5]) (hh_biological.cpp
// trpc5/trpm4).
template <bool kM4>
BC_DEV double ch_trpc5_m4(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    float cu = s.in[2] * 1000.0f;
    double s5 = m.get(5);
    float po;
    if (kM4) {
        po = cu / (cu + 0.4f);
        double ph = (s5 > 0.0f) ? s5 : 0.0f;
        po /= (1.0f + ph / 18.0f);
    } else {
        double gq = m.get(4);
        if (gq < 0.0f) gq = 0.0f;
        if (gq > 1.0f) gq = 1.0f;
        float ca_act = cu / (cu + 0.5f);
        po = gq * ca_act;
        double hc = (s5 > 0.0f) ? s5 : 0.0f;
        po /= (1.0f + hc / 9.3f);
    }
    const P3 pp = kM4 ? p_partial3(25.0f, 1.0f, 1.0f, 0.0f) : p_partial3(50.0f, 1.0f, 0.0f, 5.0f);
    double N = m.dens * area;
    double fna = ghk_mono(s, V, s.in[0], s.ex[0]);
    double ina = pp.na * fna * N * po * kAtoPa;
    double f2 = kM4 ? ghk_mono(s, V, s.in[1], s.ex[1]) : ghk_di(s, V, s.in[2], s.ex[2]);
    double i2 = (kM4 ? pp.k : pp.ca) * f2 * N * po * kAtoPa;
    constexpr int k2 = kM4 ? 1 : 2;
    s.in[0] -= (ina * dt) / s.cpm;
    s.ex[0] += (ina * dt) / (s.cpm * kEvr);
    s.in[k2] -= (i2 * dt) / s.cpm;
    s.ex[k2] += (i2 * dt) / (s.cpm * kEvr);
    floor7(s.ex);
    floor7(s.in);
    return -(ina + i2);
}

// CALHM1/3 (180, hh_biological.cpp calhm13_current): depolarisation-activated gate (V1/2 30.5,
// tau 20 ms), float Na+/K+/Ca2+ currents.
BC_DEV double ch_calhm(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    float mi = 1.0f / (1.0f + bc::exp(-(V - 30.5f) / 8.0f));
    double g0 = m.get(0);
    g0 += (mi - g0) * (dt / 20.0f);
    if (g0 < 0) g0 = 0;
    if (g0 > 1) g0 = 1;
    m.set(0, g0);
    const P3 pp = p_partial3(25.0f, 1.0f, 1.17f, 10.0f);
    double N = m.dens * area;
    double fna = ghk_mono(s, V, s.in[0], s.ex[0]);
    double fk = ghk_mono(s, V, s.in[1], s.ex[1]);
    double fca = ghk_di(s, V, s.in[2], s.ex[2]);
    float ina = (float)(pp.na * fna * N * g0 * kAtoPa);
    float ik = (float)(pp.k * fk * N * g0 * kAtoPa);
    float ica = (float)(pp.ca * fca * N * g0 * kAtoPa);
    double cpv = s.cpm;
    s.in[0] -= (ina * dt) / cpv;
    s.in[1] -= (ik * dt) / cpv;
    s.in[2] -= (ica * dt) / (2.0f * cpv);
    s.ex[0] += (ina * dt) / (cpv * kEvr);
    s.ex[1] += (ik * dt) / (cpv * kEvr);
    s.ex[2] += (ica * dt) / (2.0f * cpv * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -(ina + ik + ica);
}

// Kv3.4 (159, hh_biological.cpp kv34_current, modifier_helpers.h:564-576): m h hs with PKC
// relief of the fast inactivation, 4-AP/TEA blocks from states 5 and 6; float K+ in the flux.
BC_DEV double ch_kv34(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double g0 = m.get(0), g1 = m.get(1), g2 = m.get(2), s5 = m.get(5), s6 = m.get(6);
    float mi = 1.0f / (1.0f + bc::exp(-(V - 19.0f) / 11.3f));
    float tm = 40.0f - 37.7f / (1.0f + bc::exp(-V / 28.0f));
    g0 += (mi - g0) * dt / tm;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    float hi = 1.0f / (1.0f + bc::exp((V + 15.0f) / 7.4f));
    float th = 12.0f + 165.6f * bc::exp(-V / 11.2f);
    if (th > 2000.0f) th = 2000.0f;
    g1 += (hi - g1) * dt / th;
    if (g1 < 0.0f) g1 = 0.0f;
    if (g1 > 1.0f) g1 = 1.0f;
    double he = g1 + (1.0f - g1) * s.pkc * 0.9f;
    float hsi = 1.0f / (1.0f + bc::exp((V + 30.0f) / 6.0f));
    g2 += (hsi - g2) * dt / 2000.0f;
    if (g2 < 0.0f) g2 = 0.0f;
    if (g2 > 1.0f) g2 = 1.0f;
    m.set(0, g0);
    m.set(1, g1);
    m.set(2, g2);
    double op = g0 * he * g2;
    double ap4 = (s5 > 0.0f) ? s5 : 0.0f;
    double tea = (s6 > 0.0f) ? s6 : 0.0f;
    op /= (1.0f + ap4 / 500.0f) * (1.0f + tea / 1000.0f);
    float ki = s.in[1], ko = s.ex[1];
    double P = p_single(20.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, ki, ko);
    return k_out(s, P * flux * (m.dens * area) * op * kAtoPa, dt);
}

// Ca2+/Na+ leak (101, hh_biological.cpp ca_leak_current): always open, 50 pS split 1:2 Na:Ca,
// float concentrations in the fluxes, the volume rounded to float, intracellular change only.
BC_DEV double ch_ca_leak(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    float V = s.V;
    float ci = s.in[2], co = s.ex[2];
    float ni = s.in[0], no = s.ex[0];
    const P3 pp = p_partial3(50.0f, 1.0f, 0.0f, 2.0f);
    double N = m.dens * area;
    double fna = ghk_mono(s, V, ni, no);
    double fca = ghk_di(s, V, ci, co);
    double ina = pp.na * fna * N * kAtoPa;
    double ica = pp.ca * fca * N * kAtoPa;
    float vol = s.vol;
    double cpv = kFaraday * vol / 1000.0f;
    s.in[0] -= (ina * dt) / cpv;
    s.in[2] -= (ica * dt) / (2.0f * cpv);
    floor7(s.in);
    return -(ina + ica);
}

// Post-inhibitory-rebound T-type Ca2+ (96, hh_biological.cpp pir_ttype_current): m^3 h in ten
// substeps, serotonin and dopamine reduce it by up to 70%; float Ca in the flux, the
// intracellular Ca changes with a float charge-per-mole factor and is not floored.
BC_DEV double ch_pir_ttype(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2);
    double gamma = (s2 > 0.0f) ? s2 : 7.2f;
    double mv = m.get(0), hv = m.get(1);
    float mi = 1.0f / (1.0f + bc::exp(-(V - -48.4f) / 5.0f));
    float hi = 1.0f / (1.0f + bc::exp((V - -75.6f) / 6.5f));
    float th = 15.0f + (150.0f - 15.0f) / (1.0f + bc::exp((V - -65.0f) / 8.0f));
    float sub = dt / 10.0f;
    for (int k = 0; k < 10; k++) {
        mv += (mi - mv) * sub / 3.0f;
        hv += (hi - hv) * sub / th;
    }
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m.set(0, mv);
    m.set(1, hv);
    float ci = s.in[2], co = s.ex[2];
    double P = p_single(gamma, 2.0f, kCbarCa);
    float ht = s.nt[3], da = s.nt[2];
    double mod = (1.0 - 0.7 / (1.0 + (ht + da) / 0.01f));
    double flux = ghk_di(s, V, ci, co);
    double ica = P * flux * (m.dens * area) * mod * (mv * mv * mv * hv) * kAtoPa;
    float vol = s.vol;
    float cpm = kFaraday * vol / 1000.0;
    s.in[2] -= (ica * dt) / (2.0f * cpm);
    return -ica;
}

// L-type Ca2+ Cav1.2 (84, hh_biological.cpp:599-646, modifier_helpers.h:624-626): m^2 times a
// slow voltage-dependent (state[1]) and a Ca-dependent (state[5]) inactivation; PKA doubles the
// current at full activity.
BC_DEV double ch_ca_l(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2), s3 = m.get(3), s4 = m.get(4);
    double vh = (s3 != 0.0f) ? -fabs(s3) : -11.0f;
    double k = (s4 > 0.0f) ? s4 : 6.9f;
    double gamma = (s2 > 0.0f) ? s2 : 25.0f;
    float mi = 1.0f / (1.0f + bc::exp(-(V - vh) / (k + 1e-6f)));
    double g0 = m.get(0), g1 = m.get(1), g5 = m.get(5);
    g0 += (mi - g0) * dt / 3.0f;
    float vi = 1.0f / (1.0f + bc::exp((V + 37.3f) / 8.0f));
    g1 += (vi - g1) * dt / 150.0f;
    float ca = s.in[2] * kMmToUm;
    float cdi = 1.0f / (1.0f + ca / 1.0f);
    g5 += (cdi - g5) * dt / 35.0f;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g1 < 0.0f) g1 = 0.0f;
    if (g5 < 0.0f) g5 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    if (g1 > 1.0f) g1 = 1.0f;
    if (g5 > 1.0f) g5 = 1.0f;
    m.set(0, g0);
    m.set(1, g1);
    m.set(5, g5);
    float ci = s.in[2], co = s.ex[2];
    double P = p_single(gamma, 2.0f, kCbarCa);
    double op = g0 * g0 * g1 * g5;
    float boost = 1.0f + 1.0f * s.pka;
    double flux = ghk_di(s, V, ci, co);
    double ica = P * flux * (m.dens * area) * op * boost * kAtoPa;
    s.in[2] -= (ica * dt) / (2.0f * s.cpm);
    s.ex[2] -= (ica * dt) / (2.0f * s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -ica;
}

// Cav1.4 (144, hh_biological.cpp:648-682, modifier_helpers.h:283-313): m^2 without inactivation;
// bound CaBP4 (s.cabp4) shifts V1/2 by -10 mV.
BC_DEV double ch_cav14(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2);
    double gamma = (s2 > 0.0f) ? s2 : 4.0f;
    float vh = -15.0f + -10.0f * s.cabp4;
    float mi = 1.0f / (1.0f + bc::exp(-(V - vh) / (7.0f + 1e-6f)));
    double g0 = m.get(0);
    g0 += (mi - g0) * dt / 2.0f;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    m.set(0, g0);
    float ci = s.in[2], co = s.ex[2];
    double P = p_single(gamma, 2.0f, kCbarCa);
    double flux = ghk_di(s, V, ci, co);
    double ica = P * flux * (m.dens * area) * (g0 * g0) * kAtoPa;
    s.in[2] -= (ica * dt) / (2.0f * s.cpm);
    s.ex[2] -= (ica * dt) / (2.0f * s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -ica;
}

// Metabotropic G protein (3, hh_biological.cpp metabotropic_current): driven by the sum of all
// 15 transmitters, effector state[1] = coupling x state[0], a 35 pS GIRK-like K+ GHK current.
BC_DEV double ch_metabotropic(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    float total = 0.0f;
#pragma unroll
    for (int n = 0; n < 15; n++) total += s.nt[n];
    double coupling = (m.get(2) > 0.0f) ? m.get(2) : 1.0f;
    double tau = (m.get(3) > 0.0f) ? m.get(3) : 100.0f;
    float act = total / (total + 1.0f);
    double s0 = m.get(0);
    s0 += (act - s0) * dt / tau;
    m.set(0, s0);
    double s1 = coupling * s0;
    m.set(1, s1);
    double P = p_single(35.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s, s.V, s.in[1], s.ex[1]);
    double N = m.dens * area;
    double ik = P * flux * N * s1 * kAtoPa;
    s.in[1] -= (ik * dt) / s.cpm;
    floor7(s.in);
    return -ik;
}
// Synaptic release site (5, hh_biological.cpp synaptic_release_current): Ca2+ sensor (Hill 2,
// K_d 0.5 uM) -> facilitation state[1] and depression state[2]; a 10 pS SK-like K+ GHK current
// opens with facilitation / (1 + depression).
BC_DEV double ch_release_site(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    float ca = s.in[2];
    float sensor = (ca * ca) / (ca * ca + 0.0005f * 0.0005f);
    double fac_tau = (m.get(4) > 0.0f) ? m.get(4) : 100.0f;
    double dep_tau = (m.get(5) > 0.0f) ? m.get(5) : 500.0f;
    double s0 = m.get(0), s1 = m.get(1), s2 = m.get(2);
    s0 += (sensor - s0) * dt / 1.0f;
    s1 += (s0 - s1) * dt / fac_tau;
    s2 += (s0 - s2) * dt / dep_tau;
    m.set(0, s0);
    m.set(1, s1);
    m.set(2, s2);
    double open = s1 / (1.0f + s2);
    double P = p_single(10.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s, s.V, s.in[1], s.ex[1]);
    double N = m.dens * area;
    double ik = P * flux * N * open * kAtoPa;
    s.in[1] -= (ik * dt) / s.cpm;
    floor7(s.in);
    return -ik;
}
// MRGPR (188, modifier_helpers.h mrgpr_signal): Gq activation state[0] by the ligand in state[4]
// (EC50 28 uM; tau 70 ms up, 1500 ms down), desensitisation state[1] (tau 70 s); no current.
BC_DEV double ch_mrgpr(const ModSlot& m, const SegRegs&, float dt) {
    float lig = (m.get(4) > 0.0f) ? m.get(4) : 0.0f;
    float occ = lig / (lig + 28.0f);
    float act = m.get(0);
    float tau = (occ > act) ? 70.0f : 1500.0f;
    act += (occ - act) * dt / tau;
    if (act < 0.0f) act = 0.0f;
    if (act > 1.0f) act = 1.0f;
    m.set(0, act);
    float dz = m.get(1);
    dz += (occ - dz) * dt / 70000.0f;
    if (dz < 0.0f) dz = 0.0f;
    if (dz > 1.0f) dz = 1.0f;
    m.set(1, dz);
    return 0.0f;
}
// Somatic ATP synthesis (77), thermoregulation (78) and CREB transcription (79)
// (hh_biological.cpp nb_soma_atp_current, nb_soma_thermal_current, nb_creb_transcription_current):
// ATP, temperature (s.temp) and the transcription signal (s.txn); no current.
BC_DEV double ch_soma_atp(const ModSlot& m, SegRegs& s, float dt) {
    double nadh = m.get(0);
    float atp = s.atp;
    float inh = 1.0f - atp / 5.0f;
    if (inh < 0.0f) inh = 0.0f;
    nadh += (0.001f - nadh * 0.01f) * dt;
    if (nadh < 0.0f) nadh = 0.0f;
    if (nadh > 1.0f) nadh = 1.0f;
    double rate = (m.get(1) > 0.0f) ? m.get(1) : 0.01f;
    s.atp += rate * nadh * inh * m.dens * dt;
    if (s.atp > 5.0f) s.atp = 5.0f;
    s.temp += 0.01f * nadh * inh * 0.5f * dt * 0.001f;
    if (s.temp > 42.0f) s.temp = 42.0f;
    m.set(0, nadh);
    return 0.0f;
}
BC_DEV double ch_soma_thermal(const ModSlot& m, SegRegs& s, float dt) {
    double stress = m.get(0);
    float T = s.temp;
    double tau = (m.get(1) > 0.0f) ? m.get(1) : 5000.0f;
    T += (37.0f - T) * dt / tau;
    if (T < 30.0f) T = 30.0f;
    if (T > 42.0f) T = 42.0f;
    if (T > 39.0f) stress += (T - 39.0f) * dt * 0.0001f;
    else stress -= stress * dt * 0.001f;
    if (stress < 0.0f) stress = 0.0f;
    if (stress > 1.0f) stress = 1.0f;
    s.temp = T;
    m.set(0, stress);
    return 0.0f;
}
BC_DEV double ch_creb(const ModSlot& m, SegRegs& s, float dt) {
    float ca = s.in[2];
    double p = m.get(0), late = m.get(1), st = m.get(2);
    double kd = (m.get(3) > 0.0f) ? m.get(3) : 0.5f;
    float can = ca * ca, kn = kd * kd;
    float camk = can / (can + kn);
    p += (camk * (1.0f - p) / 5000.0f - 0.0001f * p) * dt;
    late += (p * (1.0f - late) / 30000.0f - late * 0.00001f) * dt;
    st += (late * (1.0f - st) / 300000.0f - st * 0.000001f) * dt;
    if (p < 0.0f) p = 0.0f;
    if (p > 1.0f) p = 1.0f;
    if (late < 0.0f) late = 0.0f;
    if (late > 1.0f) late = 1.0f;
    if (st < 0.0f) st = 0.0f;
    if (st > 1.0f) st = 1.0f;
    m.set(0, p);
    m.set(1, late);
    m.set(2, st);
    s.txn = 0.5f * p + 0.3f * late + 0.2f * st;
    return 0.0f;
}

// TRPV1 as a transmitter receptor (16, hh_biological.cpp trpv1_current): heat (43 C, slope 0.5 C)
// or capsaicin (transmitter 14) opens it (state[0], 200 ms; desensitisation state[1], 5 s); a
// 90 pS Na+/K+/Ca2+ GHK current.
BC_DEV double ch_trpv1r(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    float V = s.V;
    float ft = 1.0f / (1.0f + bc::exp((43.0f - s.temp) / 0.5f));
    float cap = s.nt[14];
    if (cap < 0.0f) cap = 0.0f;
    float fc = cap / (cap + 0.5f);
    double r = m.get(0);
    float pinf = ft + fc - ft * fc;
    r += (pinf - r) * dt / 200.0f;
    if (r < 0.0f) r = 0.0f;
    if (r > 1.0f) r = 1.0f;
    m.set(0, r);
    double d = m.get(1);
    d += (pinf - d) * dt / 5000.0f;
    if (d < 0.0f) d = 0.0f;
    if (d > 1.0f) d = 1.0f;
    m.set(1, d);
    float open = r * (1.0f - d);
    const P3 pp = p_partial3(90.0f, 0.125f, 0.125f, 0.75f);
    double N = m.dens * area;
    double ina = pp.na * ghk_mono(s, V, s.in[0], s.ex[0]) * N * open * kAtoPa;
    float ik = (float)(pp.k * ghk_mono(s, V, s.in[1], s.ex[1]) * N * open * kAtoPa);
    double ica = pp.ca * ghk_di(s, V, s.in[2], s.ex[2]) * N * open * kAtoPa;
    s.in[0] -= (ina * dt) / s.cpm;
    s.in[1] -= (ik * dt) / s.cpm;
    s.in[2] -= (ica * dt) / (2.0f * s.cpm);
    floor7(s.in);
    return -(ina + ik + ica);
}
// Rod phototransduction (192, hh_biological.cpp rod_phototransduction_current): light state[4]
// -> R* -> transducin -> PDE -> cGMP (states 0-3); CNG current (Hill 3, K 0.3, 25 pS).
BC_DEV double ch_rod(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    double light = (m.get(4) > 0.0f) ? m.get(4) : 0.0f;
    double s0 = m.get(0), s1 = m.get(1), s2 = m.get(2), s3 = m.get(3);
    s0 += (light - s0) * dt / 1500.0f;
    s1 += (s0 - s1) * dt / 10.0f;
    s2 += (s1 - s2) * dt / 200.0f;
    s3 += (1.0f - s3) * dt / 500.0f - s2 * s3 * dt / 50.0f;
    double st[4] = {s0, s1, s2, s3};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (st[k] < 0) st[k] = 0;
        if (st[k] > 1) st[k] = 1;
    }
    m.set(0, st[0]);
    m.set(1, st[1]);
    m.set(2, st[2]);
    m.set(3, st[3]);
    double cg = st[3];
    float cg3 = cg * cg * cg;
    float kh = 0.3f, kh3 = kh * kh * kh;
    float po = cg3 / (kh3 + cg3);
    float area = eff_area(m, s);
    const P3 pp = p_partial3(25.0f, 0.5f, 0.3f, 0.2f);
    double N = m.dens * area * po;
    double ina = pp.na * ghk_mono(s, V, s.in[0], s.ex[0]) * N * kAtoPa;
    float ik = (float)(pp.k * ghk_mono(s, V, s.in[1], s.ex[1]) * N * kAtoPa);
    double ica = pp.ca * ghk_di(s, V, s.in[2], s.ex[2]) * N * kAtoPa;
    s.in[0] -= (ina * dt) / s.cpm;
    s.in[1] -= (ik * dt) / s.cpm;
    s.in[2] -= (ica * dt) / (2.0f * s.cpm);
    floor7(s.in);
    return -(ina + ik + ica);
}
BC_DEV void clamp01_4(double* st) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (st[k] < 0.0f) st[k] = 0.0f;
        if (st[k] > 1.0f) st[k] = 1.0f;
    }
}
// The intracellular cascades (hh_biological.cpp mapk_erk_current, ikk_nfkb_current,
// mtorc1_current, protein_synth_current, cofilin_current, mmp9_current): MAPK/ERK 87 (its raise
// of the segment's CREB states is done by the step-1 kernel after it runs), NF-kB 88 (adds to
// the transcription signal s.txn), mTORC1 89 (scaled by ATP), protein synthesis 90 (reads the
// last mTORC1, s.xa; consumes ATP), cofilin 91 (reads the last MAPK's ERK, s.xa), MMP-9 92 (reads
// the last MAPK's and NF-kB's nuclear states, s.xa and s.xb). Time is scaled by the density; no
// current.
BC_DEV double ch_mapk(const ModSlot& m, SegRegs& s, float dt) {
    float ca = s.in[2] * 1e3f;
    double kd = (m.get(4) > 0.0f) ? m.get(4) : 0.5f;
    double tau_erk = (m.get(5) > 0.0f) ? m.get(5) : 5000.0f;
    double tau_nuc = (m.get(6) > 0.0f) ? m.get(6) : 30000.0f;
    dt *= m.dens;
    float ca3 = ca * ca * ca;
    float kd3 = kd * kd * kd;
    float ras = ca3 / (ca3 + kd3 + 1e-9f);
    double st[4] = {m.get(0), m.get(1), m.get(2), m.get(3)};
    st[2] += (ras - st[2]) * dt / 500.0f;
    st[1] += (st[2] - st[1]) * dt / 1000.0f;
    double erk = st[1] * st[1];
    st[0] += (erk - st[0]) * dt / tau_erk;
    st[3] += (st[0] * 0.8f - st[3]) * dt / tau_nuc;
    clamp01_4(st);
#pragma unroll
    for (int k = 0; k < 4; k++) m.set(k, st[k]);
    return 0.0f;
}
BC_DEV double ch_nfkb(const ModSlot& m, SegRegs& s, float dt) {
    float ca = s.in[2] * 1e3f;
    double kd = (m.get(4) > 0.0f) ? m.get(4) : 0.3f;
    double tau_nfkb = (m.get(5) > 0.0f) ? m.get(5) : 10000.0f;
    double tau_ikb = (m.get(6) > 0.0f) ? m.get(6) : 5000.0f;
    dt *= m.dens;
    float ikk = ca / (ca + kd + 1e-9f);
    double st[4] = {m.get(0), m.get(1), m.get(2), m.get(3)};
    st[1] += (ikk - st[1]) * dt / 2000.0f;
    st[2] += (1.0f - st[1] * 0.9f - st[2]) * dt / tau_ikb;
    if (st[2] < 0.1f) st[2] = 0.1f;
    double fr = 1.0f - st[2];
    st[0] += (fr - st[0]) * dt / tau_nfkb;
    st[3] += (st[0] * 0.7f - st[3]) * dt / (tau_nfkb * 2.0f);
    clamp01_4(st);
#pragma unroll
    for (int k = 0; k < 4; k++) m.set(k, st[k]);
    s.txn += st[3] * 0.1f;
    if (s.txn > 1.0f) s.txn = 1.0f;
    return 0.0f;
}
BC_DEV double ch_mtorc1(const ModSlot& m, SegRegs& s, float dt) {
    double tau_mtor = (m.get(5) > 0.0f) ? m.get(5) : 15000.0f;
    double tau_akt = (m.get(6) > 0.0f) ? m.get(6) : 8000.0f;
    double kd = (m.get(4) > 0.0f) ? m.get(4) : 0.3f;
    dt *= m.dens;
    double aa = m.get(3);
    float rheb = aa / (aa + kd + 1e-9f);
    double s1 = m.get(1), s2 = m.get(2), s0 = m.get(0);
    s1 += (rheb - s1) * dt / 5000.0f;
    float akt = aa * 0.8f;
    s2 += (akt - s2) * dt / tau_akt;
    double inf = s1 * s2;
    float an = s.atp / 5.0f;
    if (an > 1.0f) an = 1.0f;
    if (an < 0.0f) an = 0.0f;
    inf *= an;
    s0 += (inf - s0) * dt / tau_mtor;
    if (s0 < 0.0f) s0 = 0.0f;
    if (s0 > 1.0f) s0 = 1.0f;
    if (s1 < 0.0f) s1 = 0.0f;
    if (s2 < 0.0f) s2 = 0.0f;
    m.set(0, s0);
    m.set(1, s1);
    m.set(2, s2);
    return 0.0f;
}
BC_DEV double ch_protein_synth(const ModSlot& m, SegRegs& s, float dt) {
    double tau_s6k = (m.get(4) > 0.0f) ? m.get(4) : 10000.0f;
    double tau_4ebp = (m.get(5) > 0.0f) ? m.get(5) : 8000.0f;
    double tau_tr = (m.get(6) > 0.0f) ? m.get(6) : 60000.0f;
    dt *= m.dens;
    float mt = s.xa;
    double st[4] = {m.get(0), m.get(1), m.get(2), m.get(3)};
    st[1] += (mt - st[1]) * dt / tau_s6k;
    st[2] += (mt - st[2]) * dt / tau_4ebp;
    double tr = st[1] * st[2];
    st[0] += (tr - st[0]) * dt / tau_tr;
    st[3] = st[0] * 0.9f;
    clamp01_4(st);
#pragma unroll
    for (int k = 0; k < 4; k++) m.set(k, st[k]);
    s.atp -= st[0] * 0.0001f * dt;
    if (s.atp < 0.0f) s.atp = 0.0f;
    return 0.0f;
}
BC_DEV double ch_cofilin(const ModSlot& m, const SegRegs& s, float dt) {
    double tau_cof = (m.get(4) > 0.0f) ? m.get(4) : 5000.0f;
    double tau_limk = (m.get(5) > 0.0f) ? m.get(5) : 8000.0f;
    double tau_actin = (m.get(6) > 0.0f) ? m.get(6) : 30000.0f;
    dt *= m.dens;
    float erk = s.xa;
    double st[4] = {m.get(0), m.get(1), m.get(2), m.get(3)};
    st[1] += (erk * 0.8f - st[1]) * dt / tau_limk;
    double pc = st[1] * 0.9f;
    st[2] += (pc - st[2]) * dt / tau_cof;
    double ct = 1.0f - st[2];
    st[0] += (ct - st[0]) * dt / tau_cof;
    float poly = 0.3f;
    double sev = st[0] * 0.1f;
    st[3] += (poly - sev * st[3]) * dt / tau_actin;
    clamp01_4(st);
#pragma unroll
    for (int k = 0; k < 4; k++) m.set(k, st[k]);
    return 0.0f;
}
BC_DEV double ch_mmp9(const ModSlot& m, const SegRegs& s, float dt) {
    double tau_mmp = (m.get(4) > 0.0f) ? m.get(4) : 20000.0f;
    double tau_timp = (m.get(5) > 0.0f) ? m.get(5) : 15000.0f;
    double tau_ecm = (m.get(6) > 0.0f) ? m.get(6) : 100000.0f;
    dt *= m.dens;
    float tr = (s.xa + s.xb) * 0.5f;
    double st[4] = {m.get(0), m.get(1), m.get(2), m.get(3)};
    st[0] += (tr - st[0]) * dt / tau_mmp;
    double timp = 1.0f - st[0] * 0.5f;
    st[1] += (timp - st[1]) * dt / tau_timp;
    double deg = st[0] * (1.0f - st[1]) * 0.01f;
    float syn = 0.005f;
    st[2] += (syn - deg * st[2]) * dt / tau_ecm;
    if (st[2] < 0.1f) st[2] = 0.1f;
    if (st[2] > 1.0f) st[2] = 1.0f;
    double integ = 1.0f - st[2];
    st[3] += (integ - st[3]) * dt / 10000.0f;
    clamp01_4(st);
#pragma unroll
    for (int k = 0; k < 4; k++) m.set(k, st[k]);
    return 0.0f;
}

// Sensors (hh_biological.cpp:40-52, 483-540, 2181-2440, 3122-3141). Adaptation states relax
// toward a float activation with a float time constant (no update below 1 ms), clamped to
// [0, 1]; the adaptation gain (1 + k_s S)(1 - k_h H) is a float expression.
BC_DEV double clamp01d(double v) {
    if (v < 0.0f) v = 0.0f;
    if (v > 1.0f) v = 1.0f;
    return v;
}
BC_DEV void sensor_relax(const ModSlot& m, int k, float act, float dt, float tau) {
    if (tau < 1.0f) return;
    double v = m.get(k);
    v += (act - v) * dt / tau;
    m.set(k, clamp01d(v));
}
BC_DEV float sensor_gain(float H, float S, float kh, float ks) { return (1.0f + ks * S) * (1.0f - kh * H); }
// GHK Na+ (double), K+ (float) and, with kCa, Ca2+ (double) currents at the segment's double V
// for N channels of partial permeabilities pp; updates the intracellular pools.
struct SensorI { double na; float k; double ca; };
template <bool kCa>
BC_DEV SensorI sensor_ghk(SegRegs& s, const P3& pp, double N, float dt) {
    SensorI r;
    r.na = pp.na * ghk_mono(s, s.V, s.in[0], s.ex[0]) * N * kAtoPa;
    r.k = (float)(pp.k * ghk_mono(s, s.V, s.in[1], s.ex[1]) * N * kAtoPa);
    r.ca = kCa ? pp.ca * ghk_di(s, s.V, s.in[2], s.ex[2]) * N * kAtoPa : 0.0;
    s.in[0] -= (r.na * dt) / s.cpm;
    s.in[1] -= (r.k * dt) / s.cpm;
    if (kCa) s.in[2] -= (r.ca * dt) / (2.0f * s.cpm);
    floor7(s.in);
    return r;
}
// Physical sensor (6): heat above 37 C, cold below 33 C and stretch of segments longer than
// 50 um (s.len), scaled by state[1], drive state[0] (10 ms); habituation state[2] (tau state[6]
// or 2 s), sensitisation state[3] (30 s); 50 pS Na+/K+/Ca2+ 0.4/0.3/0.3.
BC_DEV double ch_physical_sensor(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    double sens = (m.get(1) > 0.0f) ? m.get(1) : 1.0f;
    float T = s.temp;
    float hot = (T > 37.0f) ? (T - 37.0f) / 6.0f : 0.0f;
    float cold = (T < 33.0f) ? (33.0f - T) / 8.0f : 0.0f;
    float stretch = (s.len - 50.0f) / 50.0f;
    if (stretch < 0.0f) stretch = 0.0f;
    float stim = sens * ((hot + cold) + stretch * 0.5f);
    double a = m.get(0);
    a = clamp01d(a + (stim - a) * dt / 10.0f);
    m.set(0, a);
    double tau_h = (m.get(6) > 0.0f) ? m.get(6) : 2000.0f;
    sensor_relax(m, 2, a, dt, tau_h);
    sensor_relax(m, 3, a, dt, 30000.0f);
    double kh = (m.get(4) > 0.0f) ? m.get(4) : 0.6f;
    double ks = (m.get(5) > 0.0f) ? m.get(5) : 0.2f;
    double g = sensor_gain(m.get(2), m.get(3), kh, ks);
    constexpr P3 pp = p_partial3(50.0f, 0.4f, 0.3f, 0.3f);
    const SensorI I = sensor_ghk<true>(s, pp, m.dens * area * a * g, dt);
    return -(I.na - I.k + I.ca);
}
// Hunger sensor (94): ATP / 5 mM below 0.3 (sigmoid slope 10) drives state[0] (500 ms);
// habituation state[1] (300 s), sensitisation state[2] (600 s); 50 pS Na+/K+ 0.6/0.3.
BC_DEV double ch_hunger_sensor(const ModSlot& m, SegRegs& s, float dt) {
    float an = s.atp / 5.0f;
    if (an > 1.0f) an = 1.0f;
    if (an < 0.0f) an = 0.0f;
    float h = 1.0f / (1.0f + bc::exp(10.0f * (an - 0.3f)));
    double a = m.get(0);
    a = clamp01d(a + (h - a) * dt / 500.0f);
    m.set(0, a);
    sensor_relax(m, 1, a, dt, 300000.0f);
    sensor_relax(m, 2, a, dt, 600000.0f);
    double kh = (m.get(3) > 0.0f) ? m.get(3) : 0.5f;
    double ks = (m.get(4) > 0.0f) ? m.get(4) : 0.4f;
    double g = sensor_gain(m.get(1), m.get(2), kh, ks);
    float area = eff_area(m, s);
    constexpr P3 pp = p_partial3(50.0f, 0.6f, 0.3f, 0.0f);
    const SensorI I = sensor_ghk<false>(s, pp, m.dens * area * a * g, dt);
    return -(I.na - I.k);
}
// Light sensor (95): intensity state[1]; habituation state[2] rises (2 s) above 0.1 and decays
// (5 s) below, sensitisation state[3] (60 s); the adapted intensity sets state[0] (50 ms); the
// CNG channel (25 pS, Na+/K+/Ca2+ 0.5/0.3/0.2) is open in the dark, 1 - state[0].
BC_DEV double ch_light_sensor(const ModSlot& m, SegRegs& s, float dt) {
    double I0 = m.get(1);
    if (I0 > 0.1f) sensor_relax(m, 2, 1.0f, dt, 2000.0f);
    else m.set(2, m.get(2) - dt * m.get(2) / 5000.0f);
    m.set(2, clamp01d(m.get(2)));
    sensor_relax(m, 3, I0, dt, 60000.0f);
    float kh = 0.7f;
    double ks = (m.get(4) > 0.0f) ? m.get(4) : 0.5f;
    double ad = I0 * (1.0f - kh * m.get(2)) * (1.0f + ks * m.get(3));
    if (ad < 0.0f) ad = 0.0f;
    double a = m.get(0);
    a = clamp01d(a + (ad - a) * dt / 50.0f);
    m.set(0, a);
    float area = eff_area(m, s);
    double open = 1.0f - a;
    constexpr P3 pp = p_partial3(25.0f, 0.5f, 0.3f, 0.2f);
    const SensorI I = sensor_ghk<true>(s, pp, m.dens * area * open, dt);
    return -(I.na + I.k + I.ca);
}
// Olfactory receptor (108): ligand state[1] (sigmoid at 0.01, slope 200) -> state[0] (50 ms);
// habituation state[2] (2 s), sensitisation state[3] (10 s); the Gs path closes up to 80% of a
// 35 pS K+ conductance and the current is the closed part.
BC_DEV double ch_olfactory(const ModSlot& m, SegRegs& s, float dt) {
    double L = m.get(1);
    if (L < 0.0f) L = 0.0f;
    if (L > 1.0f) L = 1.0f;
    float act = 1.0f / (1.0f + bc::exp(-2.0f * (L - 0.01f) * 100.0f));
    double a = m.get(0);
    a = clamp01d(a + (act - a) * dt / 50.0f);
    m.set(0, a);
    sensor_relax(m, 2, a, dt, 2000.0f);
    sensor_relax(m, 3, a, dt, 10000.0f);
    float kh = 0.7f;
    double ks = (m.get(4) > 0.0f) ? m.get(4) : 0.3f;
    double g = sensor_gain(m.get(2), m.get(3), kh, ks);
    float area = eff_area(m, s);
    double e = a * g;
    constexpr double P = p_single(35.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s, s.V, s.in[1], s.ex[1]);
    double N = m.dens * area;
    float basal = (float)(P * flux * N * kAtoPa);
    float modulated = (float)(P * (1.0f - e * 0.8f) * flux * N * kAtoPa);
    float eff = basal - modulated;
    s.in[1] -= (eff * dt) / s.cpm;
    floor7(s.in);
    return -eff;
}
// A double converted to uint32_t the way the reference's x86-64 build does it: truncation to a
// 64-bit integer (cvttsd2si) keeping the low 32 bits; outside the int64 range, or NaN, gives 0.
BC_DEV uint32_t u32_trunc(double x) {
    if (!(x > -9223372036854775808.0 && x < 9223372036854775808.0)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)x;
}
// Thermal noise (111): xorshift32 of (segment id s.sid, state[1], V) -> uniform r; sigma state[0]
// (default 0.05 mV) x 1.732; the equivalent current through the segment's capacitance.
BC_DEV double ch_thermal_noise(const ModSlot& m, const SegRegs& s, float dt) {
    double sigma = (m.get(0) > 0.0f) ? m.get(0) : 0.05f;
    uint32_t x = s.sid * 7919u + u32_trunc(m.get(1) * 131.0f);
    x ^= u32_trunc(s.V * 1000.0f);
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    float r = (float)(x & 0x7FFFFFFFu) / (float)0x7FFFFFFFu;
    float noise = (r * 2.0f - 1.0f) * sigma * 1.732f;
    float cm = eff_area(m, s) * 0.01f;
    if (cm < 1e-4f) cm = 1e-4f;
    return noise * cm / dt;
}
// Vibration sensor (115): intensity state[1] (sigmoid at 0.3, slope 30) -> state[0] (30 ms);
// habituation state[2] (3 s), sensitisation state[3] (10 s); 150 pS MET, Na+/K+ 0.6/0.4.
BC_DEV double ch_vibration(const ModSlot& m, SegRegs& s, float dt) {
    double I0 = m.get(1);
    if (I0 < 0.0f) I0 = 0.0f;
    if (I0 > 1.0f) I0 = 1.0f;
    float act = 1.0f / (1.0f + bc::exp(-3.0f * (I0 - 0.3f) * 10.0f));
    double a = m.get(0);
    a = clamp01d(a + (act - a) * dt / 30.0f);
    m.set(0, a);
    sensor_relax(m, 2, a, dt, 3000.0f);
    sensor_relax(m, 3, a, dt, 10000.0f);
    double ks = (m.get(4) > 0.0f) ? m.get(4) : 0.25f;
    double g = sensor_gain(m.get(2), m.get(3), 0.6f, ks);
    double e = a * g;
    float area = eff_area(m, s);
    constexpr P3 pp = p_partial3(150.0f, 0.6f, 0.4f, 0.0f);
    const SensorI I = sensor_ghk<false>(s, pp, m.dens * area * e, dt);
    return -(I.na + I.k);
}
// Otolith hair cell (116): shear state[1] along the preferred direction (cosine state[5]; -99
// or below means 1), reduced by Ca2+ adaptation state[2] (30 ms), sets the MET open probability
// state[0] (tonic 0.15, 1 ms); 150 pS Na+/K+/Ca2+ 0.6/0.3/0.1.
BC_DEV double ch_otolith(const ModSlot& m, SegRegs& s, float dt) {
    double sh = m.get(1);
    if (sh < 0.0f) sh = 0.0f;
    if (sh > 1.0f) sh = 1.0f;
    double c = (m.get(5) > -99.0f) ? m.get(5) : 1.0f;
    if (c < -1.0f) c = -1.0f;
    if (c > 1.0f) c = 1.0f;
    float es = sh * c;
    if (es < 0.0f) es = 0.0f;
    double adapt = m.get(2);
    float ad = es * (1.0f - 0.7f * adapt);
    double k = (m.get(4) > 0.0f) ? m.get(4) : 5.0f;
    float target = 0.15f + 0.85f / (1.0f + bc::exp(-k * (ad - 0.15f) * 10.0f));
    double a = m.get(0);
    a = clamp01d(a + (target - a) * dt / 1.0f);
    m.set(0, a);
    m.set(2, clamp01d(adapt + (a - adapt) * dt / 30.0f));
    float area = eff_area(m, s);
    constexpr P3 pp = p_partial3(150.0f, 0.60f, 0.30f, 0.10f);
    const SensorI I = sensor_ghk<true>(s, pp, m.dens * area * a, dt);
    float tot = I.na - I.k + I.ca;
    return -tot;
}
// Semicircular-canal cupula (117): a damped torsion pendulum driven by angular acceleration
// state[1] (theta state[0], omega state[3]; omega0^2 state[4] or 1, 2 zeta omega0 state[5] or
// 0.9) with an 8 s elastic return; |theta| opens the hair cell, adapted by state[2] (10 ms);
// 150 pS Na+/K+/Ca2+ 0.55/0.35/0.10.
BC_DEV double ch_cupula(const ModSlot& m, SegRegs& s, float dt) {
    double al = m.get(1);
    if (al < -1.0f) al = -1.0f;
    if (al > 1.0f) al = 1.0f;
    double w2 = (m.get(4) > 0.0f) ? m.get(4) : 1.0f;
    double zw = (m.get(5) > 0.0f) ? m.get(5) : 0.9f;
    float gam = 1.0f;
    double om = m.get(3), th = m.get(0);
    float dom = al * gam - zw * om - w2 * th;
    om += dom * dt / 1000.0f;
    th += om * dt / 1000.0f;
    th *= bc::exp(-dt / 8000.0f);
    om *= bc::exp(-dt / 8000.0f);
    if (th < -1.0f) th = -1.0f;
    if (th > 1.0f) th = 1.0f;
    m.set(0, th);
    m.set(3, om);
    float defl = fabsf((float)th);
    double hs = (m.get(4) > 0.0f) ? m.get(4) * 0.4f : 2.5f;
    float target = 1.0f / (1.0f + bc::exp(-hs * (defl - 0.1f) * 15.0f));
    double ca = m.get(2);
    ca = clamp01d(ca + (target - ca) * dt / 10.0f);
    m.set(2, ca);
    double open = target * (1.0f - 0.6f * ca);
    float area = eff_area(m, s);
    constexpr P3 pp = p_partial3(150.0f, 0.55f, 0.35f, 0.10f);
    const SensorI I = sensor_ghk<true>(s, pp, m.dens * area * open, dt);
    float tot = I.na - I.k + I.ca;
    return -tot;
}
// Cryptochrome magnetic sensor (120): radical-pair yield light x sin^2(theta) x wavelength x
// sensitivity (states 1, 2, 5, 4) -> state[0] (50 ms). Its boost of the segment's first light
// sensor (state[0] x (1 + 0.3 state[0]), capped at 1) is applied by the step-1 kernel after it
// runs. No current.
BC_DEV double ch_cryptochrome(const ModSlot& m, const SegRegs& s, float dt) {
    double L = m.get(1);
    if (L < 0.0f) L = 0.0f;
    if (L > 1.0f) L = 1.0f;
    double s2 = m.get(2);
    if (s2 < 0.0f) s2 = 0.0f;
    if (s2 > 1.0f) s2 = 1.0f;
    double wf = (m.get(5) > 0.0f) ? m.get(5) : 1.0f;
    double sens = (m.get(4) > 0.0f) ? m.get(4) : 1.0f;
    float rp = L * s2 * wf * sens;
    double a = m.get(0);
    m.set(0, clamp01d(a + (rp - a) * dt / 50.0f));
    return 0.0f;
}

// ER Ca2+ store (hh_biological.cpp:3473-3591): SERCA (126) pumps cytosolic Ca2+ into the
// segment's ER pool s.er (Segment.ER_Ca2), IP3R (127: Li-Rinzel gating, IP3 = s.ip3) and RyR
// (128: Keizer-Levine gating with adaptation w) release it; no membrane current. Constants
// units.h:130-156.
BC_DEV double ch_serca(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    double km = (m.get(1) > 0.0f) ? m.get(1) : 0.00027f;
    float ca = s.in[2];
    if (ca < 1e-9f) ca = 1e-9f;
    float fca = (ca * ca) / (ca * ca + km * km);
    float atp = s.atp;
    if (atp < 0.0f) atp = 0.0f;
    float fatp = atp / (atp + 0.1f);
    double N = m.dens * area;
    float flux = N * 100.0f * fca * fatp * 2.0f * dt * 1e-3f / 6.02214076e23f;
    float dconc = flux / (s.vol * 1e-18f);
    float avail = s.in[2] - 1e-9f;
    if (avail < 0.0f) avail = 0.0f;
    float moved = (dconc < avail) ? dconc : avail;
    s.in[2] -= moved;
    s.er += moved;
    floor7(s.in);
    if (s.er < 0.0f) s.er = 0.0f;
    s.atp -= flux * 1e6f;
    if (s.atp < 0.0f) s.atp = 0.0f;
    return 0.0f;
}
BC_DEV double ch_ip3r(const ModSlot& m, SegRegs& s, float dt) {
    float ip3 = s.ip3;
    double h = m.get(1);
    float ca = s.in[2];
    if (ca < 1e-9f) ca = 1e-9f;
    float er = s.er;
    if (er < 0.0f) er = 0.0f;
    float minf = ip3 / (ip3 + 0.00013f);
    float ninf = ca / (ca + 0.0002f);
    float hinf = 0.001f / (0.001f + ca);
    if (h <= 0.0f) h = hinf;
    h += (hinf - h) * dt / 100.0f;
    if (h < 0.0f) h = 0.0f;
    if (h > 1.0f) h = 1.0f;
    float m3 = minf * minf * minf;
    float n3 = ninf * ninf * ninf;
    float h3 = h * h * h;
    float popen = m3 * n3 * h3;
    float drive = er - ca;
    if (drive < 0.0f) drive = 0.0f;
    double dconc = 5.0f * m.dens * popen * drive * dt * 1e-3f;
    float avail = er - 1e-9f;
    if (avail < 0.0f) avail = 0.0f;
    float moved = (dconc < avail) ? dconc : avail;
    s.er -= moved;
    s.in[2] += moved;
    floor7(s.in);
    if (s.er < 0.0f) s.er = 0.0f;
    m.set(1, h);
    return 0.0f;
}
BC_DEV double ch_ryr(const ModSlot& m, SegRegs& s, float dt) {
    double w = m.get(0);
    float ca = s.in[2];
    if (ca < 1e-9f) ca = 1e-9f;
    float er = s.er;
    if (er < 0.0f) er = 0.0f;
    float ckb = ca / 0.000636f;
    float ckb2 = ckb * ckb;
    float ckb3 = ckb2 * ckb;
    float kac = 0.000372f / ca;
    float kac4 = kac * kac * kac * kac;
    float winf = (1.0f + kac4 + ckb3) / (1.0f + 1.0f / 0.057f + kac4 + ckb2);
    if (w <= 0.0f) w = winf;
    float popen = w * (1.0f + ckb3) / (1.0f + kac4 + ckb3);
    float tauw = 10000.0f * winf;
    if (tauw < 1.0f) tauw = 1.0f;
    w += (winf - w) * dt / tauw;
    if (w < 0.0f) w = 0.0f;
    if (w > 1.0f) w = 1.0f;
    float drive = er - ca;
    if (drive < 0.0f) drive = 0.0f;
    double dconc = 20.0f * m.dens * popen * drive * dt * 1e-3f;
    float avail = er - 1e-9f;
    if (avail < 0.0f) avail = 0.0f;
    float moved = (dconc < avail) ? dconc : avail;
    s.er -= moved;
    s.in[2] += moved;
    floor7(s.in);
    if (s.er < 0.0f) s.er = 0.0f;
    m.set(0, w);
    return 0.0f;
}
// PLC (136, modifier_helpers.h:1553-1580): the Gq receptors' activation s.gq drives IP3
// (state[0]), DAG (state[1]) and a desensitisation (state[2]); constants units.h:342-349.
BC_DEV double ch_plc(const ModSlot& m, const SegRegs& s, float dt) {
    float gq = s.gq;
    if (gq > 1.0f) gq = 1.0f;
    float vip3 = (m.get(3) > 0.0f) ? m.get(3) : 0.0005f;
    float vdag = (m.get(4) > 0.0f) ? m.get(4) : 0.5f;
    float drive = 0.02f + gq;
    float ip3 = m.get(0);
    ip3 += (vip3 * m.dens * drive) * dt * 0.001f - ip3 * dt / 1000.0f;
    if (ip3 < 0.0f) ip3 = 0.0f;
    m.set(0, ip3);
    float dag = m.get(1);
    dag += (vdag * m.dens * drive) * dt * 0.001f - dag * dt / 2000.0f;
    if (dag < 0.0f) dag = 0.0f;
    m.set(1, dag);
    float dz = m.get(2);
    dz += (gq - dz) * dt / 70000.0f;
    if (dz < 0.0f) dz = 0.0f;
    if (dz > 1.0f) dz = 1.0f;
    m.set(2, dz);
    return 0.0f;
}

// eCB synthesis (86, hh_biological.cpp:1746-1764): 2-AG from Ca (Hill 2, Kd 0.5 uM) scaled by
// 1 + DAG (s.dag), degraded at rate state[2] (0.005 if not positive), synthesis rate state[1]
// (0.01 if not positive), dt scaled by the density, state[0] clamped to [0, 10]; no current.
// Step 3 deposits the level into the clefts of the segment's incoming synapses.
BC_DEV double ch_ecb(const ModSlot& m, const SegRegs& s, float dt) {
    float ca = s.in[2] * kMmToUm;
    double synth_rate = (m.get(1) > 0.0f) ? m.get(1) : 0.01f;
    double degrad = (m.get(2) > 0.0f) ? m.get(2) : 0.005f;
    float dag = s.dag;
    dt *= m.dens;
    float can = ca * ca;
    float synth = synth_rate * can / (can + 0.25f + 1e-9f) * (1.0f + dag);
    double e = m.get(0);
    e += (synth - e * degrad) * dt;
    if (e < 0.0f) e = 0.0f;
    if (e > 10.0f) e = 10.0f;
    m.set(0, e);
    return 0.0;
}

// CaMKII (82, hh_biological.cpp:545-570): Ca2+ (uM, Hill 4, K_d state[2] or 1) drives state[0]
// with time constant state[3] (2000 ms); autophosphorylation state[1] at rate state[5] (0.0005)
// decays at 1e-4 per ms; dt scaled by the density; no current.
BC_DEV double ch_camkii(const ModSlot& m, const SegRegs& s, float dt) {
    float ca = s.in[2];
    float cu = ca * kMmToUm;
    const double s0 = m.get(0), s1 = m.get(1), s2 = m.get(2), s3 = m.get(3), s5 = m.get(5);
    double kd = (s2 > 0.0f) ? s2 : 1.0f;
    double tau = (s3 > 0.0f) ? s3 : 2000.0f;
    double ar = (s5 > 0.0f) ? s5 : 0.0005f;
    dt *= m.dens;
    float c4 = cu * cu * cu * cu;
    float inf = c4 / (c4 + kd * kd * kd * kd + 1e-9f);
    double a = s0 + (inf - s0) * dt / tau;
    double b = s1 + (a * ar - s1 * 0.0001f) * dt;
    if (a < 0.0f) a = 0.0f;
    if (b < 0.0f) b = 0.0f;
    if (a > 1.0f) a = 1.0f;
    if (b > 1.0f) b = 1.0f;
    m.set(0, a);
    m.set(1, b);
    return 0.0;
}

// Calcineurin (83, hh_biological.cpp:575-591): Ca2+ (uM, Hill 2, K_d state[1] or 0.05) drives
// state[0] with time constant state[2] (300 ms); dt scaled by the density; no current.
BC_DEV double ch_calcineurin(const ModSlot& m, const SegRegs& s, float dt) {
    float ca = s.in[2];
    float cu = ca * kMmToUm;
    const double s0 = m.get(0), s1 = m.get(1), s2 = m.get(2);
    double kd = (s1 > 0.0f) ? s1 : 0.05f;
    double tau = (s2 > 0.0f) ? s2 : 300.0f;
    dt *= m.dens;
    float c2 = cu * cu;
    float inf = c2 / (c2 + kd * kd + 1e-9f);
    double a = s0 + (inf - s0) * dt / tau;
    if (a < 0.0f) a = 0.0f;
    if (a > 1.0f) a = 1.0f;
    m.set(0, a);
    return 0.0;
}

// PP1 / DARPP-32 (113, hh_biological.cpp:3222-3253): the PKA activity (s.pka: state[1] of the
// segment's first cAMP/PKA modifier) drives p-DARPP-32 (state[0]) and p-inhibitor-1 (state[1])
// with time constant state[4] (5000 ms), dt scaled by the density; PP1 activity state[3] =
// basal (state[5] or 0.5) / (1 + 10 state[0] + 5 state[1] + 1e-9); state[6] = PKA; no current.
BC_DEV double ch_pp1(const ModSlot& m, const SegRegs& s, float dt) {
    const double s0 = m.get(0), s1 = m.get(1), s4 = m.get(4), s5 = m.get(5);
    double tau = (s4 > 0.0f) ? s4 : 5000.0f;
    double basal = (s5 > 0.0f) ? s5 : 0.5f;
    dt *= m.dens;
    float pka = s.pka;
    double a = s0 + (pka - s0) * dt / tau;
    double b = s1 + (pka - s1) * dt / tau;
    if (a < 0.0f) a = 0.0f;
    if (b < 0.0f) b = 0.0f;
    if (a > 1.0f) a = 1.0f;
    if (b > 1.0f) b = 1.0f;
    m.set(0, a);
    m.set(1, b);
    double inh = 10.0f * a + 5.0f * b;
    m.set(3, basal / (1.0f + inh + 1e-9f));
    m.set(6, pka);
    return 0.0;
}

// PKC (137, modifier_helpers.h:1584-1593): state[0] = DAG / (DAG + 0.1) * Ca / (Ca + 0.0005 mM),
// DAG = s.dag (state[1] of the segment's first PLC/IP3/DAG modifier), Ca floored at 0; no current.
BC_DEV double ch_pkc(const ModSlot& m, const SegRegs& s) {
    float dag = s.dag;
    float ca = s.in[2];
    if (ca < 0.0f) ca = 0.0f;
    float df = dag / (dag + 0.1f);
    float cf = ca / (ca + 0.0005f);
    m.set(0, df * cf);
    return 0.0;
}

// CaBP4 (145, hh_biological.cpp:684-693): no current; state[0] binds towards 1 with tau 1 ms.
BC_DEV double ch_cabp4(const ModSlot& m, float dt) {
    double g0 = m.get(0);
    g0 += (1.0f - g0) * dt / 1.0f;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    m.set(0, g0);
    return 0.0;
}

// CaCC / ANO2 (146, hh_biological.cpp:695-722, modifier_helpers.h:318-327): Ca-activated Cl-
// channel, Hill 2 with EC50 4 uM (the reference's pow(x, 2) of a float is the exact square in
// double), GHK anion current; only the intracellular Cl changes.
BC_DEV double ch_cacc(const ModSlot& m, SegRegs& s, float dt) {
    float V = s.V;
    float area = eff_area(m, s);
    double s2 = m.get(2);
    double gamma = (s2 > 0.0f) ? s2 : 2.0f;
    float cu = s.in[2] * kMmToUm;
    if (cu < 0.0f) cu = 0.0f;
    float cn = (double)cu * (double)cu;
    float mi = cn / (cn + 16.0f);
    double g0 = m.get(0);
    g0 += (mi - g0) * dt / 2.0f;
    if (g0 < 0.0f) g0 = 0.0f;
    if (g0 > 1.0f) g0 = 1.0f;
    m.set(0, g0);
    float ci = s.in[3], co = s.ex[3];
    double P = p_single(gamma, 1.0f, kCbarCl);
    double flux = ghk_an(s, V, ci, co);
    double icl = P * flux * (m.dens * area) * g0 * kAtoPa;
    s.in[3] += (icl * dt) / s.cpm;
    floor7(s.in);
    return -icl;
}

// CAN / TRPM4-5 (152, hh_biological.cpp:3643-3682): Ca-activated (Hill 2, Kd 0.5 uM) Na+/K+
// channel with slow Ca-dependent desensitisation in state[1]; float currents.
BC_DEV double ch_can(const ModSlot& m, SegRegs& s, float dt) {
    constexpr P3 pp = p_partial3(30.0f, 1.0f, 1.0f, 0.0f);
    float V = s.V;
    float area = eff_area(m, s);
    float ca = s.in[2];
    const float kd = 5e-4f;
    float po = (ca * ca) / (ca * ca + kd * kd);
    float di = ca / (ca + 0.01f);
    double g1 = m.get(1);
    g1 += (di - g1) * dt / 2000.0f;
    if (g1 < 0.0f) g1 = 0.0f;
    if (g1 > 1.0f) g1 = 1.0f;
    m.set(1, g1);
    po *= (1.0f - g1);
    double N = m.dens * area;
    double fna = ghk_mono(s, V, s.in[0], s.ex[0]);
    double fk = ghk_mono(s, V, s.in[1], s.ex[1]);
    float ina = (float)(pp.na * fna * N * po * kAtoPa);
    float ik = (float)(pp.k * fk * N * po * kAtoPa);
    s.in[0] -= (ina * dt) / s.cpm;
    s.in[1] -= (ik * dt) / s.cpm;
    s.ex[0] += (ina * dt) / (s.cpm * kEvr);
    s.ex[1] += (ik * dt) / (s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -(ina - ik);
}

// TMC1 (190, hh_biological.cpp tmc1_current): hair-cell transduction channel, open fraction
// relaxes to 0.1 + 0.9 stim/(stim + 0.3) (stim = state[4]), fast (state[1]) and slow (state[2])
// Ca-dependent adaptation; Ca2+ and Na+ currents from partial permeabilities, as floats.
BC_DEV double ch_tmc1(const ModSlot& m, SegRegs& s, float dt) {
    constexpr P3 pp = p_partial3(80.0f, 0.21f, 0.0f, 0.79f);
    float V = s.V;
    float area = eff_area(m, s);
    double s4 = m.get(4);
    double stim = (s4 > 0.0f) ? s4 : 0.0f;
    float pi = 0.1f + 0.9f * (stim / (stim + 0.3f));
    double g0 = m.get(0), g1 = m.get(1), g2 = m.get(2);
    g0 += (pi - g0) * (dt / 1.0f);
    float ca = s.in[2] * kMmToUm;
    float ad = ca / (ca + 5.0f);
    g1 += (ad - g1) * (dt / 2.0f);
    g2 += (ad - g2) * (dt / 45.0f);
    g0 = g0 < 0 ? 0.0 : (g0 > 1 ? 1.0 : g0);
    g1 = g1 < 0 ? 0.0 : (g1 > 1 ? 1.0 : g1);
    g2 = g2 < 0 ? 0.0 : (g2 > 1 ? 1.0 : g2);
    m.set(0, g0);
    m.set(1, g1);
    m.set(2, g2);
    double open = g0 * (1.0f - g1) * (1.0f - g2);
    double N = m.dens * area;
    double fca = ghk_di(s, V, s.in[2], s.ex[2]);
    double fna = ghk_mono(s, V, s.in[0], s.ex[0]);
    float ica = (float)(pp.ca * fca * N * open * kAtoPa);
    float ina = (float)(pp.na * fna * N * open * kAtoPa);
    s.in[2] -= (ica * dt) / (2.0f * s.cpm);
    s.in[0] -= (ina * dt) / s.cpm;
    s.ex[2] += (ica * dt) / (2.0f * s.cpm * kEvr);
    s.ex[0] += (ina * dt) / (s.cpm * kEvr);
    floor7(s.in);
    floor7(s.ex);
    return -(ica + ina);
}

// AQP4 (186, hh_biological.cpp:4114-4129, modifier_helpers.h:395-396): water follows the osmotic
// difference of the summed (float) concentrations and dilutes every intracellular ion by a
// factor clamped to [0.95, 1]; no current, no floor.
BC_DEV double ch_aqp4(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    float si = 0.0f, se = 0.0f;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        si += s.in[i];
        se += s.ex[i];
    }
    float dpi = si - se;
    double J = m.dens * 24e-14f * area * dpi * 1e-6f;
    float f = 1.0f / (1.0f + J * dt);
    if (f < 0.95f) f = 0.95f;
    if (f > 1.0f) f = 1.0f;
#pragma unroll
    for (int i = 0; i < 7; i++) s.in[i] *= f;
    return 0.0;
}

// PMCA (98, hh_biological.cpp:2574-2614): Hill Ca2+ extrusion, ATP-dependent. The reference
// takes the Hill powers with pow; for the default n = 2 (the only one the loader accepts as
// covered) both are exact squares in double.
BC_DEV double ch_pmca(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    double s1 = m.get(1);
    double km = (s1 > 0.0f) ? s1 : 0.0003f;
    float atp = s.atp;
    if (atp < 0.0f) atp = 0.0f;
    float f_atp = atp / (atp + 0.2f);
    float ca = s.in[2];
    if (ca < 1e-9f) ca = 1e-9f;
    float ca_pow = (double)ca * (double)ca;
    float f_ca = ca_pow / (ca_pow + km * km);
    double N = m.dens * area;
    float ip = N * 100.0f * f_ca * f_atp * 1.602176634e-19f * 2.0f * kAtoPa;
    float vol = s.vol;
    double cpv = kFaraday * vol / 1000.0;
    float tr = ip * dt / (2.0f * cpv);
    float ca_av = s.in[2] - 1e-9f;
    if (ca_av < 0.0f) ca_av = 0.0f;
    if (tr > ca_av) tr = ca_av;
    s.in[2] -= tr;
    s.ex[2] += tr / kEvr;
    floor7(s.in);
    s.atp -= ip * dt * 1e-8f;
    if (s.atp < 0.0f) s.atp = 0.0f;
    return -ip;
}

// NCX (99, hh_biological.cpp:2616-2646, modifier_helpers.h:1181-1201): 3 Na+ in / 1 Ca2+ out,
// voltage-dependent (eta 0.35), allosteric Ca2+ activation, either direction.
BC_DEV double ch_ncx(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    float na_i = s.in[0]; if (na_i < 1e-9f) na_i = 1e-9f;
    float na_o = s.ex[0]; if (na_o < 1e-9f) na_o = 1e-9f;
    float ca_i = s.in[2]; if (ca_i < 1e-9f) ca_i = 1e-9f;
    float ca_o = s.ex[2]; if (ca_o < 1e-9f) ca_o = 1e-9f;
    float vv = (float)s.V * 1e-3f;
    float frt = kFaraday / (kGas * kTemp);
    float e_eta = bc::exp(0.35f * frt * vv);
    float e_eta1 = bc::exp((0.35f - 1.0f) * frt * vv);
    float nai3 = na_i * na_i * na_i;
    float nao3 = na_o * na_o * na_o;
    float fwd = e_eta1 * nao3 * ca_i;
    float rev = e_eta * nai3 * ca_o;
    float den = 1.0f + 0.27f * e_eta1;
    float cr = 0.000125f / ca_i;
    float allo = 1.0f / (1.0f + cr * cr);
    float fn = allo * (fwd - rev) / (den * 530.2f);
    double N = m.dens * area;
    float I = N * 500.0f * fn * 1.602176634e-19f * kAtoPa;
    float vol = s.vol;
    double cpv = kFaraday * vol / 1000.0;
    float tr = I * dt / cpv;
    if (tr > ca_i) tr = ca_i;
    if (tr < -na_i / 3.0f) tr = -na_i / 3.0f;
    s.in[2] -= tr;
    s.ex[2] += tr / kEvr;
    s.in[0] += 3.0f * tr;
    s.ex[0] -= 3.0f * tr / kEvr;
    floor7(s.in);
    return I;
}

// KCC2 (100, hh_biological.cpp:2648-2703): electroneutral K+-Cl- cotransport down the product
// gradient, limited by what the losing side holds. No current.
BC_DEV double ch_kcc2(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    double s1 = m.get(1), s2 = m.get(2);
    double km_k = (s1 > 0.0f) ? s1 : 6.5f;
    double km_cl = (s2 > 0.0f) ? s2 : 6.5f;
    float k_i = s.in[1]; if (k_i < 1e-9f) k_i = 1e-9f;
    float cl_i = s.in[3]; if (cl_i < 1e-9f) cl_i = 1e-9f;
    float k_o = s.ex[1]; if (k_o < 1e-9f) k_o = 1e-9f;
    float cl_o = s.ex[3]; if (cl_o < 1e-9f) cl_o = 1e-9f;
    float f_k = k_i / (k_i + km_k);
    float f_cl = cl_i / (cl_i + km_cl);
    float drv = 1.0f - (k_o * cl_o) / (k_i * cl_i);
    if (drv > 1.0f) drv = 1.0f;
    if (drv < -1.0f) drv = -1.0f;
    float flux = 5.0f * 1e-20f * m.dens * area * f_k * f_cl * drv * dt;
    float dc = flux / (s.vol * 1e-18f);
    if (dc > 0.0f) {
        float ka = k_i - 1e-9f, cla = cl_i - 1e-9f;
        float mn = (ka < cla) ? ka : cla;
        if (mn < 0.0f) mn = 0.0f;
        if (dc > mn) dc = mn;
    } else {
        float ka = (k_o - 1e-9f) * kEvr, cla = (cl_o - 1e-9f) * kEvr;
        float mn = (ka < cla) ? ka : cla;
        if (mn < 0.0f) mn = 0.0f;
        float mag = -dc;
        if (mag > mn) mag = mn;
        dc = -mag;
    }
    s.in[1] -= dc;
    s.in[3] -= dc;
    s.ex[1] += dc / kEvr;
    s.ex[3] += dc / kEvr;
    floor7(s.in);
    return 0.0;
}

// NKCC1 (122, hh_biological.cpp:2705-2760): Na+-K+-2Cl- uptake, trans-stimulated by external
// Na+ and trans-inhibited by external K+. No current.
BC_DEV double ch_nkcc1(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    double s1 = m.get(1), s2 = m.get(2), s3 = m.get(3), s4 = m.get(4), s5 = m.get(5);
    double km_nai = (s1 > 0.0f) ? s1 : 10.0f;
    double km_ki = (s2 > 0.0f) ? s2 : 15.0f;
    double km_cli = (s3 > 0.0f) ? s3 : 50.0f;
    double km_nao = (s4 > 0.0f) ? s4 : 0.5f;
    double km_ko = (s5 > 0.0f) ? s5 : 0.2f;
    float na_i = s.in[0], k_i = s.in[1], cl_i = s.in[3], na_o = s.ex[0], k_o = s.ex[1];
    if (na_i < 1e-9f) na_i = 1e-9f;
    if (k_i < 1e-9f) k_i = 1e-9f;
    if (cl_i < 1e-9f) cl_i = 1e-9f;
    if (na_o < 1e-9f) na_o = 1e-9f;
    if (k_o < 1e-9f) k_o = 1e-9f;
    float fna = na_i / (na_i + km_nai);
    float fk = k_i / (k_i + km_ki);
    float fcl = cl_i / (cl_i + km_cli);
    float stim = na_o / (na_o + km_nao);
    float inhib = 1.0f / (1.0f + k_o / km_ko);
    float flux = 7.0f * 1e-20f * m.dens * area * fna * fk * fcl * stim * inhib * dt;
    float dc = flux / (s.vol * 1e-18f);
    float naa = (s.ex[0] - 1e-9f) * kEvr;
    float ka = (s.ex[1] - 1e-9f) * kEvr;
    float cla = (s.ex[3] - 1e-9f) * kEvr / 2.0f;
    float mn = (naa < ka) ? naa : ka;
    if (cla < mn) mn = cla;
    if (mn < 0.0f) mn = 0.0f;
    if (dc > mn) dc = mn;
    s.in[0] += dc;
    s.in[1] += dc;
    s.in[3] += 2.0f * dc;
    s.ex[0] -= dc / kEvr;
    s.ex[1] -= dc / kEvr;
    s.ex[3] -= 2.0f * dc / kEvr;
    floor7(s.in);
    return 0.0;
}

// Ca2+ buffer (102, hh_biological.cpp:2825-2865): first-order binding to a pool of capacity
// state[0] (0.15 mM) x density; state[1] is the bound Ca (set to equilibrium when not positive).
BC_DEV double ch_ca_buffer(const ModSlot& m, SegRegs& s, float dt) {
    double s0 = m.get(0), s2 = m.get(2), s3 = m.get(3), s4 = m.get(4);
    double bref = (s0 > 0.0f) ? s0 : 0.15f;
    double btot = bref * m.dens;
    double kd = (s2 > 0.0f) ? s2 : 0.0005f;
    double kon = (s3 > 0.0f) ? s3 : 0.1f;
    double koff = (s4 > 0.0f) ? s4 : (kd * kon);
    double bound = m.get(1);
    if (bound <= 0.0f && btot > 0.0f) {
        float c0 = s.in[2];
        bound = btot * c0 / (c0 + kd);
    }
    float cf = s.in[2];
    if (cf < 1e-9f) cf = 1e-9f;
    float bf = btot - bound;
    if (bf < 0.0f) bf = 0.0f;
    float db = (kon * cf * bf - koff * bound) * dt;
    bound += db;
    if (bound < 0.0f) bound = 0.0f;
    if (bound > btot) bound = btot;
    s.in[2] -= db;
    if (s.in[2] < 1e-9f) {
        float def = 1e-9f - s.in[2];
        s.in[2] = 1e-9f;
        bound -= def;
        if (bound < 0.0f) bound = 0.0f;
    }
    m.set(1, bound);
    return 0.0;
}

// Na+-coupled transmitter transporters (modifier_helpers.h:418-455, hh_biological.cpp:725-780):
// Michaelis-Menten uptake of transmitter NT from the segment's extracellular pool with NA Na+
// (and CL Cl-) in and KO K+ out per cycle, NET charges; the turnover is limited by the
// transmitter, the extracellular Na+ and, with K+ countertransport, the intracellular K+.
// Returns the transporter current (pA). The charge-per-mole factor uses the volume rounded to
// float, as the reference's helper does.
template <int NT, int NA, int CL, int KO, int NET>
BC_DEV double ch_transporter(const ModSlot& m, SegRegs& s, float dt, float km, float rel_per_mm) {
    constexpr float kUaPerCm2ToPa = 1e-8f * 1e6f;
    const float area = eff_area(m, s);
    float nt = s.nt[NT];
    if (nt < 0.0f) nt = 0.0f;
    const float frac = nt / (nt + km);
    // the density is a double in the record, so the product runs in double
    const float I = (float)(2.0f * m.dens * area * kUaPerCm2ToPa * frac);
    const double cpm = kFaraday * (double)(float)s.vol / 1000.0;
    float tr = (float)((double)(I * dt) / ((double)(float)NET * cpm));
    const float nt_av = nt / rel_per_mm;
    const float na_av = (float)(s.ex[0] * (double)kEvr / (double)(float)NA);
    if (tr > nt_av) tr = nt_av;
    if (tr > na_av) tr = na_av;
    if (KO && tr > s.in[1]) tr = (float)s.in[1];
    if (tr < 0.0f) tr = 0.0f;
    s.in[0] += (float)NA * tr;
    s.ex[0] -= (float)NA * tr / kEvr;
    if (CL) {
        s.in[3] += (float)CL * tr;
        s.ex[3] -= (float)CL * tr / kEvr;
    }
    if (KO) {
        s.in[1] -= (float)KO * tr;
        s.ex[1] += (float)KO * tr / kEvr;
    }
    float v = s.nt[NT] - tr * rel_per_mm;
    if (v < 0.0f) v = 0.0f;
    s.nt[NT] = v;
    floor7(s.in);
    floor7(s.ex);
    return I;
}

// Modifier dispatch (kernel_dispatch.h:337-570) over the covered IDs; unknown IDs give 0
// current and leave their state untouched, as the reference's device switch does.
// The network-builder NMDA variants (63, 72; receptors_biological.cpp:222-229, 312-319): NMDA
// gating without Ca-dependent inactivation, Mg2+ block by the extracellular Mg2+ (1 mM when below
// 0.01).
BC_DEV double rec_nmda_mg(const ModSlot& m, SegRegs& s, float dt) {
    const IonoK k = {50.0f, 1.0f, 5.0f, 50.0f, 0, 0.08f, 0.08f, 0.84f, false, 0.0f};
    double I = iono(m, s, dt, k);
    float mg = s.ex[4];
    if (mg < 0.01f) mg = 1.0f;
    float block = 1.0f / (1.0f + (mg / 3.57f) * bc::exp(-0.062f * s.V));
    return I * block;
}

// The MSN receptors' mod_scale (receptors_biological.cpp:338-345, 358-365): state[1] of the
// segment's first D1 (65, for 75) or D2 (66, for 76) modulator as it stands, floored at 0, else 1.
BC_DEV float msn_scale(const uint8_t* mtype, const double* mstate, size_t P, uint32_t i, uint32_t cnt,
                       uint8_t t) {
    const uint8_t want = t == 75 ? 65 : 66;
    for (uint32_t q = 0; q < cnt; q++)
        if (mtype[q * P + i] == want) {
            float v = mstate[((size_t)q * 7 + 1) * P + i];
            if (v < 0.0f) v = 0.0f;
            return v;
        }
    return 1.0f;
}

// Gi/o-coupled receptors (17-40, modifier_helpers.h:1332-1363): occupancy of transmitter NT
// drives a GIRK gate (state[0]) with desensitisation (state[1]); EC50 and gamma from
// state[active_vars] and state[active_vars + 1] when positive.
template <int NT>
BC_DEV double ch_gio(const ModSlot& m, SegRegs& s, float dt) {
    float area = eff_area(m, s);
    double V = s.V;
    const double sa = state_any(m, m.av), sb = state_any(m, m.av + 1u);
    float ec50 = (sa > 0.0f) ? sa : 1.0f;
    float gamma = (sb > 0.0f) ? sb : 35.0f;
    float nt = s.nt[NT];
    float act = nt / (nt + ec50);
    float r = m.get(0);
    float tau = (act > r) ? 70.0f : 1500.0f;
    r += (act - r) * dt / tau;
    if (r < 0.0f) r = 0.0f;
    if (r > 1.0f) r = 1.0f;
    m.set(0, r);
    float d = m.get(1);
    d += (act - d) * dt / 40000.0f;
    if (d < 0.0f) d = 0.0f;
    if (d > 1.0f) d = 1.0f;
    m.set(1, d);
    float fo = kir_f_open(V) * r * (1.0f - 0.55f * d);
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(s, V, s.in[1], s.ex[1]);
    double ik = P * flux * (m.dens * area) * fo * kAtoPa;
    s.in[1] -= (ik * dt) / s.cpm;
    floor7(s.in);
    return -ik;
}

// Gs-coupled receptors (41-51, modifier_helpers.h:1379-1398): activation (state[0]) and
// desensitisation (state[1]) signals of transmitter NT; no current.
template <int NT>
BC_DEV double ch_gs(const ModSlot& m, SegRegs& s, float dt) {
    const double sa = state_any(m, m.av), sb = state_any(m, m.av + 1u);
    float ec50 = (sa > 0.0f) ? sa : 1.0f;
    float tdes = (sb > 0.0f) ? sb : 70000.0f;
    float nt = s.nt[NT];
    float occ = nt / (nt + ec50);
    float act = m.get(0);
    float tau = (occ > act) ? 70.0f : 1500.0f;
    act += (occ - act) * dt / tau;
    if (act < 0.0f) act = 0.0f;
    if (act > 1.0f) act = 1.0f;
    m.set(0, act);
    float d = m.get(1);
    d += (occ - d) * dt / tdes;
    if (d < 0.0f) d = 0.0f;
    if (d > 1.0f) d = 1.0f;
    m.set(1, d);
    return 0.0;
}

// The covered IDs outside the core set, in a switch of their own: the core networks' kernels
// never see these cases in their register allocation.
BC_DEV double dispatch_ext(uint8_t type, const ModSlot& m, SegRegs& s, float dt) {
    switch (type) {
        case M_CX36: case M_CX43: return ch_connexin(m, s, dt);
        case M_HCN2: return ch_hcn1<2>(m, s, dt);
        case 131: return ch_hcn1<3>(m, s, dt);
        case 132: return ch_hcn1<4>(m, s, dt);
        case 15: return iono(m, s, dt, IonoK{15.0f, 0.5f, 1.0f, 20.0f, 10, 0.22f, 0.22f, 0.56f, false, 20.0f});
        case 134: return iono(m, s, dt, IonoK{15.0f, 0.5f, 1.0f, 20.0f, 10, 0.22f, 0.22f, 0.56f, false, 5000.0f});
        case 64: case 71: return iono(m, s, dt, IonoK{8.0f, 0.5f, 0.5f, 3.0f, 0, 0.5f, 0.5f, 0.0f, false, 3.0f});
        case 73: return iono(m, s, dt, IonoK{28.0f, 1.0f, 1.0f, 5.0f, 1, 0.0f, 0.0f, 0.0f, true, 2000.0f});
        case 67: return iono(m, s, dt, IonoK{40.0f, 0.5f, 0.5f, 3.0f, 4, 0.12f, 0.17f, 0.71f, false, 50.0f});
        case 68: return iono(m, s, dt, IonoK{30.0f, 0.5f, 1.0f, 10.0f, 4, 0.48f, 0.48f, 0.04f, false, 2000.0f});
        case 74: return ch_gio<1>(m, s, dt);
        case 63: case 72: return rec_nmda_mg(m, s, dt);
        case 75: {   // MSN NMDA (receptors_biological.cpp:337-352): scaled by s.xa, the D1 term
            const IonoK k = {50.0f, 1.0f, 5.0f, 50.0f, 0, 0.08f, 0.08f, 0.84f, false, 0.0f};
            double I = iono(m, s, dt, k, s.xa);
            float mg = s.ex[4];
            if (mg < 0.01f) mg = 1.0f;
            float block = 1.0f / (1.0f + (mg / 3.57f) * bc::exp(-0.062f * s.V));
            return I * block;
        }
        case 76:   // MSN AMPA (receptors_biological.cpp:357-369): scaled by s.xa, the D2 term
            return iono(m, s, dt, IonoK{8.0f, 0.5f, 0.5f, 3.0f, 0, 0.5f, 0.5f, 0.0f, false, 3.0f}, s.xa);
        case 52: case 53: case 54: return ch_gs<3>(m, s, dt);   // Gq: the Gs signal's form
        case 55: case 56: return ch_gs<6>(m, s, dt);
        case 57: case 58: case 59: case 70: return ch_gs<4>(m, s, dt);
        case 60: case 61: return ch_gs<0>(m, s, dt);
        case 62: return ch_gs<8>(m, s, dt);
        case 10: return iono(m, s, dt, IonoK{3.0f, 0.5f, 1.0f, 10.0f, 0, 0.5f, 0.5f, 0.0f, false, 5.0f});
        case 12: return iono(m, s, dt, IonoK{60.0f, 0.5f, 0.3f, 5.0f, 7, 0.0f, 0.0f, 0.0f, true, 2000.0f});
        case 13: return iono(m, s, dt, IonoK{30.0f, 0.5f, 1.0f, 10.0f, 4, 0.48f, 0.48f, 0.04f, false, 2000.0f});
        case 14: return iono(m, s, dt, IonoK{2.0f, 0.3f, 1.0f, 10.0f, 3, 0.40f, 0.40f, 0.20f, false, 800.0f});
        case 97: return iono(m, s, dt, IonoK{40.0f, 0.5f, 0.5f, 3.0f, 4, 0.12f, 0.17f, 0.71f, false, 50.0f});
        case 194: return rec_nmda<false, true>(m, s, dt);
        case 17: return ch_gio<1>(m, s, dt);
        case 18: case 19: case 20: case 21: return ch_gio<3>(m, s, dt);
        case 22: case 23: case 24: return ch_gio<2>(m, s, dt);
        case 25: case 26: return ch_gio<4>(m, s, dt);
        case 27: case 28: return ch_gio<5>(m, s, dt);
        case 29: case 30: case 31: case 32: case 33: case 34: return ch_gio<0>(m, s, dt);
        case 35: return ch_gio<6>(m, s, dt);
        case 36: case 37: return ch_gio<9>(m, s, dt);
        case 38: return ch_gio<8>(m, s, dt);
        case 39: case 40: return ch_gio<11>(m, s, dt);
        case 41: case 42: case 43: return ch_gs<3>(m, s, dt);
        case 44: case 45: case 51: return ch_gs<2>(m, s, dt);
        case 46: case 47: return ch_gs<5>(m, s, dt);
        case 48: case 49: return ch_gs<6>(m, s, dt);
        case 50: return ch_gs<8>(m, s, dt);
        case M_CAV13: return ch_cav13(m, s, dt);
        case M_CLC2: return ch_cl(m, s, dt);
        case M_KV42: return ch_kv42(m, s, dt);
        case M_KV21: return ch_kv21(m, s, dt);
        case M_NAP: return ch_nap<false>(m, s, dt);
        case M_NAP_TAD: return ch_nap<true>(m, s, dt);
        case M_KIR: return ch_kir(m, s, dt);
        case M_NAV12DEV: return ch_nav12dev(m, s, dt);
        case 177: return ch_nav_simple(m, s, dt, NavSimpleK{-22.0f, 6.5f, -64.0f, 6.5f, 0.5f, 0.65f, 4.0f, 20.8f});
        case 169: return ch_nav_simple(m, s, dt, NavSimpleK{-24.0f, 5.0f, -71.0f, 6.0f, 0.5f, 1.2f, 10.0f, 19.5f});
        case 170: return ch_nav_simple(m, s, dt, NavSimpleK{-20.0f, 7.0f, -30.0f, 5.0f, 1.5f, 8.0f, 60000.0f, 20.0f});
        case 171: return ch_nav_simple(m, s, dt, NavSimpleK{-47.0f, 5.5f, -44.0f, 6.0f, 10.0f, 43.0f, 39000.0f, 16.0f});
        case 199: return ch_nav11(m, s, dt);
        case 157: return ch_kv1x<true>(m, s, dt);
        case 154: return ch_kv1x<false>(m, s, dt);
        case 206: return ch_kv14(m, s, dt);
        case 207: return ch_kv16_22<false>(m, s, dt);
        case 205: return ch_kv16_22<true>(m, s, dt);
        case 203: return ch_kvgen(m, s, dt, KvGenK{10.0f, 8.0f, 0.0f, 1.0f, 12.0f, 0.0f, 15.0f, 63.0f, 25.0f, false, false});
        case 204: return ch_kvgen(m, s, dt, KvGenK{1.5f, 8.0f, -27.9f, 8.0f, 4.0f, 200.0f, 35.0f, 67.5f, 100.0f, true, false});
        case 202: return ch_kvgen(m, s, dt, KvGenK{-31.7f, 8.0f, -47.4f, 5.9f, 1.7f, 0.0f, 4.4f, 488.0f, 81.0f, true, true});
        case 160: return ch_kleak<160>(m, s, dt);
        case 161: return ch_kleak<161>(m, s, dt);
        case 166: return ch_kleak<166>(m, s, dt);
        case 167: return ch_kleak<167>(m, s, dt);
        case 208: return ch_kleak<208>(m, s, dt);
        case 178: return ch_kir71_kcnq4<false>(m, s, dt);
        case 176: return ch_kir71_kcnq4<true>(m, s, dt);
        case 168: return ch_cav22(m, s, dt);
        case 200: return ch_cavt(m, s, dt, CavTK{-45.0f, 5.5f, -77.0f, 6.0f, 7.5f, 25.0f, 250.0f, 15.0f, 250.0f, 1.6f});
        case 201: return ch_cavt(m, s, dt, CavTK{-46.0f, 6.0f, -72.0f, 6.0f, 9.0f, 16.0f, 180.0f, 15.0f, 13.0f, 3.0f});
        case 191: return ch_cavt(m, s, dt, CavTK{-43.8f, 6.9f, -72.5f, 6.6f, 11.0f, 97.0f, 400.0f, 30.0f, 200.0f, 1.6f});
        case 210: return ch_trp_naca<210>(m, s, dt);
        case 162: return ch_trp_naca<162>(m, s, dt);
        case 211: return ch_trp_naca<211>(m, s, dt);
        case 175: return ch_trpv3(m, s, dt);
        case 182: return ch_enac(m, s, dt);
        case 172: return ch_trp_lig(m, s, dt, TrpLigK{0.7f, 10.0f, 150.0f, 70.0f, 0.09f, 0.91f, true, false});
        case 173: return ch_trp_lig(m, s, dt, TrpLigK{30.0f, 20.0f, 200.0f, 63.0f, 0.5f, 0.5f, true, false});
        case 174: return ch_trp_lig(m, s, dt, TrpLigK{30.0f, 30.0f, 300.0f, 90.0f, 0.54f, 0.46f, true, true});
        case 179: return ch_trp_lig(m, s, dt, TrpLigK{10.0f, 50.0f, 0.0f, 35.0f, 0.17f, 0.83f, false, false});
        case 163: return ch_trpc5_m4<false>(m, s, dt);
        case 164: return ch_trpc5_m4<true>(m, s, dt);
        case 180: return ch_calhm(m, s, dt);
        case 159: return ch_kv34(m, s, dt);
        case 101: return ch_ca_leak(m, s, dt);
        case 96: return ch_pir_ttype(m, s, dt);
        case 84: return ch_ca_l(m, s, dt);
        case 144: return ch_cav14(m, s, dt);
        case 145: return ch_cabp4(m, dt);
        case 86: return ch_ecb(m, s, dt);
        case 126: return ch_serca(m, s, dt);
        case 127: return ch_ip3r(m, s, dt);
        case 128: return ch_ryr(m, s, dt);
        case 136: return ch_plc(m, s, dt);
        case 82: return ch_camkii(m, s, dt);
        case 83: return ch_calcineurin(m, s, dt);
        case 113: return ch_pp1(m, s, dt);
        case 137: return ch_pkc(m, s);
        case 3: return ch_metabotropic(m, s, dt);
        case 5: return ch_release_site(m, s, dt);
        case 188: return ch_mrgpr(m, s, dt);
        case 77: return ch_soma_atp(m, s, dt);
        case 78: return ch_soma_thermal(m, s, dt);
        case 79: return ch_creb(m, s, dt);
        case 16: return ch_trpv1r(m, s, dt);
        case 192: return ch_rod(m, s, dt);
        case 87: return ch_mapk(m, s, dt);
        case 88: return ch_nfkb(m, s, dt);
        case 89: return ch_mtorc1(m, s, dt);
        case 90: return ch_protein_synth(m, s, dt);
        case 91: return ch_cofilin(m, s, dt);
        case 92: return ch_mmp9(m, s, dt);
        case 6: return ch_physical_sensor(m, s, dt);
        case 94: return ch_hunger_sensor(m, s, dt);
        case 95: return ch_light_sensor(m, s, dt);
        case 108: return ch_olfactory(m, s, dt);
        case 111: return ch_thermal_noise(m, s, dt);
        case 115: return ch_vibration(m, s, dt);
        case 116: return ch_otolith(m, s, dt);
        case 117: return ch_cupula(m, s, dt);
        case 120: return ch_cryptochrome(m, s, dt);
        case 146: return ch_cacc(m, s, dt);
        case 152: return ch_can(m, s, dt);
        case 186: return ch_aqp4(m, s, dt);
        case 190: return ch_tmc1(m, s, dt);
        case M_PUMP_A3: return ch_pump<PUMP_A3>(m, s, dt);
        case M_PUMP_GLIA: return ch_pump<PUMP_GLIA>(m, s, dt);
        case M_PMCA: return ch_pmca(m, s, dt);
        case M_NCX: return ch_ncx(m, s, dt);
        case M_KCC2: return ch_kcc2(m, s, dt);
        case M_NKCC1: return ch_nkcc1(m, s, dt);
        case M_CA_BUFFER: return ch_ca_buffer(m, s, dt);
        case M_GLAST: return ch_transporter<0, 3, 0, 1, 2>(m, s, dt, 10.0f, 667.0f);
        case M_GLYT1: return ch_transporter<7, 2, 1, 0, 1>(m, s, dt, 30.0f, 667.0f);
        case M_GAT3: return ch_transporter<1, 2, 1, 0, 1>(m, s, dt, 0.5f, 667.0f);
        case M_EAAT3: return ch_transporter<0, 3, 0, 1, 2>(m, s, dt, 20.0f, 667.0f);
        case M_GLYT2: return ch_transporter<7, 3, 1, 0, 2>(m, s, dt, 55.7f, 667.0f);
        case M_GAT1: return ch_transporter<1, 3, 1, 0, 2>(m, s, dt, 4.7f, 667.0f);
        default: return 0.0;
    }
}
// kNtMods = false: the caller has no modifier that reads transmitters (no ionotropic receptor
// or transporter; see modifier_reads_nt), and those cases are left out. kExt = false: only the
// core channel set runs (see modifier_core), which keeps the kernel's register budget for the
// common networks.
template <bool kNtMods = true, bool kExt = true>
BC_DEV double dispatch(uint8_t type, const ModSlot& m, SegRegs& s, float dt) {
    if (!kNtMods && modifier_reads_nt(type)) return 0.0;
    if (kExt && !modifier_core(type)) return dispatch_ext(type, m, s, dt);
    switch (type) {
        case M_PUMP: return ch_pump(m, s, dt);
        case M_AMPA: {
            const IonoK k = {8.0f, 0.5f, 0.5f, 3.0f, 0, 0.5f, 0.5f, 0.0f, false, 3.0f};
            return iono(m, s, dt, k);
        }
        case M_NMDA: return rec_nmda(m, s, dt);
        case M_GABAA: {
            const IonoK k = {28.0f, 1.0f, 1.0f, 5.0f, 1, 0.0f, 0.0f, 0.0f, true, 2000.0f};
            return iono(m, s, dt, k);
        }
        case M_SK: return ch_sk(m, s, dt);
        case M_BK: return ch_bk(m, s, dt);
        case M_K2P: return ch_k2p(m, s, dt);
        case M_KV7: return ch_kv7(m, s, dt);
        case M_KVDR: return ch_kvdr(m, s, dt);
        case M_HCN1: return ch_hcn1(m, s, dt);
        case M_NALCN: return ch_nalcn(m, s, dt);
        case M_CAV21: return ch_cav21(m, s, dt);
        case M_NAV16: {
            const NavK g = {43.0f, 6.0f, 50.0f, 75.0f, 72.0f};
            return ch_nav(m, s, dt, g, 20.7f, 5.0f, 0.2f);
        }
        case M_NAV12: {
            const NavK g = {23.0f, 6.0f, 41.5f, 66.5f, 63.5f};
            return ch_nav(m, s, dt, g, 20.8f, 10.0f, 0.7f);
        }
        case M_KV31: return ch_kv31(m, s, dt);
        case M_KV11: return ch_kv11(m, s, dt);
        default: return 0.0;
    }
}

// ---- electrodiffusion helpers (diffusion_kernel.h:108-113) ----
template <class Tab = ExpTabLds>
BC_DEV float bernoulli(float x, Tab tab = Tab{}) {
    if (fabsf(x) < 1e-3f) return 1.0f - 0.5f * x;
    if (x > 50.0f) return 0.0f;
    if (x < -50.0f) return -x;
    return x / (bc::exp_t(x, tab) - 1.0f);
}

BC_DEV float xorshift_draw(uint32_t seed) {   // simulation_host.cpp:150-156
    uint32_t x = seed;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return (float)(x & 0x7FFFFFFFu) / (float)0x7FFFFFFFu;
}

}  // namespace bc
