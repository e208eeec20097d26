// cpu_ref.cpp — the oracle: a single-threaded CPU restatement of BioNeuronCore's per-tick step
// (core_run_tick, src/simulation_host.cpp:174-657) for the channel/receptor set the hot path
// covers. TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it as the checker and CPU baseline; the product (libbiocore_hip_backend.so) never
// links, loads or calls it.
//
// It reproduces the reference's arithmetic operation by operation: the same float/double
// types per expression, the same evaluation order, no FMA contraction (-ffp-contract=off),
// and glibc's ::exp/::pow, so results are bit-identical to the reference library. That
// identity is checked by tests/test_oracle_vs_reference.py against oracle/_ref (the reference
// compiled from its own sources) and against the committed golden fixtures.
//
// Scope: the modifier and receptor IDs dispatch() names; chemical synapses with the reference's
// twelve built-in presynaptic release modulators, eCB retrograde deposits, electrical synapses
// (gap junctions) and glial transmitter capture into GLAST-carrying segments; the protein ->
// release mapping of consolidation; no mechano stimuli or plugins. Anything outside that set is
// reported by oracle_unsupported_count() and computed as the reference's "unknown ID" default
// (0 current, state untouched).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "../include/biocore_abi.h"
// The record fields' own float type; f64.h (the FP64 build) promotes every other float.
#ifndef ORC_F64
typedef float abi_float;
#endif

namespace orc {

// ---- constants (reference include/types.h, units.h, run_tick_common.h, modifier_helpers.h) ----
constexpr double kFaraday = 96485.0;
constexpr float kGas = 8.314f;
constexpr float kTemp = 310.15f;
constexpr float kPi = 3.14159265f;
constexpr float kEvr = 0.2f / (1.0f - 0.2f);   // EXTRA_VOLUME_RATIO
constexpr float kAtoPa = 1e12f;
constexpr float kPsToS = 1e-12f;
constexpr float kMmToUm = 1e3f;
constexpr float kCbarNa = 55.964f, kCbarK = 40.514f, kCbarCa = 0.12775f, kCbarCl = 36.20f;
constexpr float kSigmaE = 3.612f;
constexpr float kFrtMv = 0.03742f;
constexpr float kChargeToPa = 96.485f;
constexpr float kZ[7] = {1.0f, 1.0f, 2.0f, -1.0f, 2.0f, -1.0f, 1.0f};
constexpr float kDiff[7] = {1.33f, 1.96f, 0.22f, 2.03f, 0.10f, 0.0f, 9.31f};
constexpr float kBath[7] = {140.0f, 3.0f, 1.2f, 130.0f, 1.0f, 0.0f, 4.0e-5f};
constexpr float kNtMass[15] = {147.0f, 103.0f, 153.0f, 176.0f, 146.0f, 267.0f, 169.0f, 75.0f,
                               111.0f, 348.0f, 507.0f, 232.0f, 24.0f, 162.0f, 0.0f};

enum : uint8_t {
    T_MYELIN = 4, T_PUMP = 7, T_AMPA = 8, T_NMDA = 9, T_GABAA = 11, T_GLYCINE = 12, T_GABAB = 17,
    T_5HT3 = 14, T_NICOTINIC = 13, T_D1 = 44, T_D2 = 22, T_5HT7 = 43, T_CAMKII = 82,
    T_CALCINEURIN = 83, T_ECB = 86, T_SK = 103, T_BK = 104, T_K2P = 107, T_KV7 = 112,
    T_PP1 = 113, T_KVDR = 123, T_HCN1 = 129, T_NALCN = 133, T_CAMP_PKA = 135, T_PKC = 137,
    T_CAV21 = 140, T_NAV16 = 141, T_NAV12 = 142, T_GLAST = 147, T_KV31 = 150, T_KV11 = 156,
    T_CX36 = 114, T_CX43 = 198, T_GLYT1 = 148, T_GAT3 = 149, T_EAAT3 = 183, T_GLYT2 = 184,
    T_GAT1 = 185, T_HCN2 = 130, T_PUMP_A3 = 197, T_PUMP_GLIA = 110, T_PMCA = 98, T_NCX = 99,
    T_KCC2 = 100, T_NKCC1 = 122, T_CA_BUFFER = 102, T_CAV13 = 1, T_CL = 2, T_KV42 = 85,
    T_KV21 = 196, T_NAP = 105, T_NAP_TAD = 195, T_KIR = 106, T_NAV12DEV = 193,
};

static uint64_t g_unsupported = 0;

static inline float nt_diffusion(int nt) {
    float mass = kNtMass[nt];
    if (mass < 1.0f) mass = 1.0f;
    return 1.33f * powf(23.0f / mass, 1.0f / 3.0f);
}

// ---- geometry (run_tick_common.h:209-264) ----
static inline bool is_sphere(const Segment* s) {
    return s->segment_type == 1 || s->segment_type == 3 || s->segment_type == 4;
}
static inline float membrane_area(const Segment* s) {
    return is_sphere(s) ? 4.0f * kPi * s->radius * s->radius : 2.0f * kPi * s->radius * s->length;
}
static inline double volume(const Segment* s) {
    if (is_sphere(s))
        return (4.0 / 3.0) * kPi * (double)s->radius * (double)s->radius * (double)s->radius;
    return kPi * (double)s->radius * (double)s->radius * (double)s->length;
}
static inline float capacitance(const Segment* s) {
    float c = membrane_area(s) * 0.01f;
    if (c < 1e-4f) c = 1e-4f;
    for (uint8_t j = 0; j < s->modifier_count && j < 16; j++) {
        if (s->modifiers[j].type == T_MYELIN) {
            float gr = s->modifiers[j].state[0];
            if (gr <= 0.0f || gr >= 1.0f) gr = 0.6f;
            float factor = 100.0f * (1.0f / gr - 1.0f) / (1.0f / 0.6f - 1.0f);
            c /= factor;
            break;
        }
    }
    return c;
}
static inline float eff_area(const Modifier* m, const Segment* s) {
    if (m->state[7] > 0.0f) return m->state[7];
    return membrane_area(s);
}
static inline void floor_ions(double* a) {
    for (int i = 0; i < 7; i++)
        if (a[i] < 1e-9) a[i] = 1e-9;
}
static inline bool has_myelin(const Segment* s) {
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == T_MYELIN) return true;
    return false;
}
static inline float first_state_of(const Segment* s, uint8_t type, int k, float dflt) {
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == type) return s->modifiers[j].state[k];
    return dflt;
}
// Ionotropic + blocked-by-myelin predicates (kernel_dispatch.h:283-309), restricted to the IDs
// the covered set can contain; the full list matters only for myelinated segments.
static inline bool is_ionotropic(uint8_t t) {
    return t <= 2 || (t >= 8 && t <= 16) || t == 97 || t == 134 || t == 194;
}

// ---- GHK (run_tick_common.h:274-312, modifier_helpers.h:48-75) ----
static inline double ghk_mono(double V, double ci, double co) {
    double vn = kFaraday * V * 0.001 / (kGas * kTemp);
    if (fabs(vn) < 1e-6) return kFaraday * (ci - co);
    double k = kFaraday / (kGas * kTemp);
    if (vn > 20.0) return k * kFaraday * V * 0.001 * ci;
    if (vn < -20.0) return k * kFaraday * V * 0.001 * co;
    return k * kFaraday * V * 0.001 * (ci - co * exp(-vn)) / (1.0 - exp(-vn));
}
static inline double ghk_di(double V, double ci, double co) {
    double vn = 2.0 * kFaraday * V * 0.001 / (kGas * kTemp);
    if (fabs(vn) < 1e-6) return 2.0 * kFaraday * (ci - co);
    double k = 4.0 * kFaraday / (kGas * kTemp);
    if (vn > 20.0) return k * kFaraday * V * 0.001 * ci;
    if (vn < -20.0) return k * kFaraday * V * 0.001 * co;
    return k * kFaraday * V * 0.001 * (ci - co * exp(-vn)) / (1.0 - exp(-vn));
}
static inline double ghk_an(double V, double ci, double co) {
    double vn = kFaraday * V * 0.001 / (kGas * kTemp);
    if (fabs(vn) < 1e-6) return kFaraday * (co - ci);
    double k = kFaraday / (kGas * kTemp);
    if (vn > 20.0) return k * kFaraday * V * 0.001 * co;
    if (vn < -20.0) return k * kFaraday * V * 0.001 * ci;
    return k * kFaraday * V * 0.001 * (co - ci * exp(-vn)) / (1.0 - exp(-vn));
}
static inline double p_single(float gamma_ps, float z, float cbar) {
    double g = gamma_ps * kPsToS;
    double z2f2 = z * z * kFaraday * kFaraday;
    double rt = kGas * kTemp;
    return g * rt / (z2f2 * cbar);
}
static inline void p_partial(float gamma_ps, float fna, float fk, float fca, double* pna,
                             double* pk, double* pca) {
    double w = fna * kCbarNa + fk * kCbarK + fca * 4.0 * kCbarCa;
    *pna = p_single((float)(gamma_ps * fna * kCbarNa / w), 1.0f, kCbarNa);
    *pk = p_single((float)(gamma_ps * fk * kCbarK / w), 1.0f, kCbarK);
    *pca = p_single((float)(gamma_ps * fca * 4.0 * kCbarCa / w), 2.0f, kCbarCa);
}
static inline void p_nalcn(float gamma_ps, double* pna, double* pk, double* pca) {
    const double fna = 1.3, fk = 1.2, fca = 0.5;
    double w = fna * kCbarNa + fk * kCbarK + fca * 4.0 * kCbarCa;
    *pna = p_single((float)(gamma_ps * fna * kCbarNa / w), 1.0f, kCbarNa);
    *pk = p_single((float)(gamma_ps * fk * kCbarK / w), 1.0f, kCbarK);
    *pca = p_single((float)(gamma_ps * fca * 4.0 * kCbarCa / w), 2.0f, kCbarCa);
}

// ---- gate helpers (modifier_helpers.h:132-271, 909-1013) ----
static inline float vtrap(float x, float y) {
    if (fabsf(x / y) < 1e-6f) return y * (1.0f - x / (2.0f * y));
    return x / (1.0f - exp(-x / y));
}
struct NavGates { float am_off, am_k, ah_off, bh_off, h_off; };
static constexpr NavGates kNav16 = {43.0f, 6.0f, 50.0f, 75.0f, 72.0f};
static constexpr NavGates kNav12 = {23.0f, 6.0f, 41.5f, 66.5f, 63.5f};
static inline float nav_am(const NavGates& g, float V) { return 0.182f * vtrap(V + g.am_off, g.am_k); }
static inline float nav_bm(const NavGates& g, float V) { return 0.124f * vtrap(-(V + g.am_off), g.am_k); }
static inline float nav_ah(const NavGates& g, float V) { return 0.024f * vtrap(V + g.ah_off, 5.0f); }
static inline float nav_bh(const NavGates& g, float V) { return 0.0091f * vtrap(-(V + g.bh_off), 5.0f); }
static inline float nav_minf(const NavGates& g, float V) {
    float a = nav_am(g, V), b = nav_bm(g, V);
    return a / (a + b);
}
static inline float nav_mtau(const NavGates& g, float V) { return 1.0f / (nav_am(g, V) + nav_bm(g, V)); }
static inline float nav_hinf(const NavGates& g, float V) { return 1.0f / (1.0f + exp((V + g.h_off) / 6.2f)); }
static inline float nav_htau(const NavGates& g, float V) { return 1.0f / (nav_ah(g, V) + nav_bh(g, V)); }

static inline float cav21_minf(float V) { return 1.0f / (1.0f + exp(-(V - -15.0f) / (6.0f + 1e-6f))); }
static inline float cav21_hinf(float V, float ca_um) {
    float h = 1.0f / (1.0f + exp((V - -50.0f) / 8.0f));
    h *= 1.0f / (1.0f + ca_um / 3.0f);
    return h;
}
static inline float kdr_tau(float V, float vh) {
    float tau = 14.6f * exp(-(V - vh) / 30.1f);
    if (tau < 0.3f) tau = 0.3f;
    if (tau > 200.0f) tau = 200.0f;
    return tau;
}
static inline float sk_pinf(float ca) {
    if (ca < 1e-9f) ca = 1e-9f;
    float r = 0.0004f / ca;
    return 1.0f / (1.0f + pow(r, 4.0f));
}
static inline float bk_po(float V, float ca) {
    float vn = kFaraday * V * 1e-3f / (kGas * kTemp);
    float J = 0.104f * exp(0.55f * vn);
    float L = 3.7e-8f * exp(0.30f * vn);
    if (ca < 0.0f) ca = 0.0f;
    float K = ca / 0.044f;
    float o = 1.0f + K * 23.0f + J * 32.0f + J * K * 23.0f * 32.0f * 3.7f;
    float c = 1.0f + K + J + J * K * 3.7f;
    o *= o; o *= o;
    c *= c; c *= c;
    float num = L * o;
    return num / (num + c);
}
static inline float pump_vnorm(float V, float na, float ko, float atp_level, float kmna = 10.0f,
                               float kmk = 1.5f) {
    float na3 = na * na * na, k2 = ko * ko;
    float kna3 = kmna * kmna * kmna, kk2 = kmk * kmk;
    float fna = na3 / (na3 + kna3), fk = k2 / (k2 + kk2);
    float atp = atp_level;
    if (atp < 0.0f) atp = 0.0f;
    float fatp = atp / (atp + 0.5f);
    float eta = 0.8f * kFaraday * (V * 1e-3f) / (kGas * kTemp);
    float a2 = 60.0f * exp(0.5f * eta);
    float b2 = 6.0f * exp(-(1.0f - 0.5f) * eta);
    float a1 = 200.0f * fna * fatp;
    float a3 = 200.0f * fk;
    float a4 = 60.0f;
    float denom = a3 / a1 + (b2 + a3) / a2 + 1.0f + a3 / a4;
    float turnover = a3 / denom;
    float k3 = 200.0f;
    float vmax = k3 / (k3 / 200.0f + 1.0f + k3 / 60.0f);
    return turnover / vmax;
}

// ---- channel currents (return pA, negative = outward) ----
// Na/K-ATPase (7), its alpha3 isoform (197: K_m Na 25 mM, ouabain block by state[5] nM, IC50
// 20 nM) and the glial alpha2 pump (110: K_m K 1 mM) (hh_biological.cpp:57-141, 3414-3445).
enum PumpKind { PUMP_A1, PUMP_A3, PUMP_GLIA };
static double chan_pump(Modifier* m, Segment* s, float dt, PumpKind kind = PUMP_A1) {
    float area = eff_area(m, s);
    float vn = pump_vnorm(s->V_m, s->intra_ions[0], s->extra_ions[1], s->ATP_level,
                          kind == PUMP_A3 ? 25.0f : 10.0f, kind == PUMP_GLIA ? 1.0f : 1.5f);
    if (kind == PUMP_A3) {
        double oua = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        vn /= (1.0f + oua / 20.0f);
    }
    double N = m->density * area;
    float ip = N * vn * 75.0f * 1.602176634e-19f * kAtoPa;
    float vol = volume(s);
    double cpm = kFaraday * vol / 1000.0;
    float tr = ip * dt / cpm;
    float na_av = (s->intra_ions[0] - 1e-9f) / 3.0f;
    float k_av = (s->extra_ions[1] - 1e-9f) * kEvr / 2.0f;
    if (na_av < 0.0f) na_av = 0.0f;
    if (k_av < 0.0f) k_av = 0.0f;
    float mx = (na_av < k_av) ? na_av : k_av;
    if (tr > mx) tr = mx;
    s->intra_ions[0] -= 3.0f * tr;
    s->extra_ions[0] += 3.0f * tr / kEvr;
    s->intra_ions[1] += 2.0f * tr;
    s->extra_ions[1] -= 2.0f * tr / kEvr;
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    s->ATP_level -= ip * dt * 1e-8f;
    if (s->ATP_level < 0.0f) s->ATP_level = 0.0f;
    return -ip;
}

static double chan_nav(Modifier* m, Segment* s, float dt, const NavGates& g, float gamma,
                       float ttx_ic50, float pkx) {   // hh_biological.cpp:193-277
    float V = s->V_m;
    float area = eff_area(m, s);
    if (s->V_m < -200.0f) s->V_m = -200.0f;
    if (s->V_m > 200.0f) s->V_m = 200.0f;
    V = s->V_m;
    double mv = nav_minf(g, V) + (m->state[0] - nav_minf(g, V)) * exp(-dt / nav_mtau(g, V));
    double hv = nav_hinf(g, V) + (m->state[1] - nav_hinf(g, V)) * exp(-dt / nav_htau(g, V));
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m->state[0] = mv;
    m->state[1] = hv;
    float op = mv * mv * mv * hv;
    double ttx = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    op /= (1.0f + ttx / ttx_ic50);
    float pkc = first_state_of(s, T_PKC, 0, 0.0f);
    float pka = first_state_of(s, T_CAMP_PKA, 1, 0.0f);
    op *= (1.0f - pkx * (pkc + pka) * 0.5f);
    if (op < 0.0f) op = 0.0f;
    double P = p_single(gamma, 1.0f, kCbarNa);
    double flux = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double ia = P * flux * (m->density * area) * op;
    double ina = ia * kAtoPa;
    double vol = volume(s);
    double cpv = kFaraday * vol / 1000.0;
    s->intra_ions[0] -= (ina * dt) / cpv;
    s->extra_ions[0] += (ina * dt) / (cpv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -ina;
}

static double chan_cav21(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:364-388
    float V = s->V_m;
    float mi = cav21_minf(V);
    m->state[0] += (mi - m->state[0]) * dt / 1.0f;
    float ca = s->intra_ions[2] * kMmToUm;
    float hi = cav21_hinf(V, ca);
    m->state[1] += (hi - m->state[1]) * dt / 300.0f;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    return 0.0f;
}

// K+ efflux bookkeeping shared by the K channels: flux in double, extra floor only.
static inline double k_out(Segment* s, double ik, float dt) {
    double vol = volume(s);
    double cpv = kFaraday * vol / 1000.0;
    s->intra_ions[1] -= (ik * dt) / cpv;
    s->extra_ions[1] += (ik * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    return -ik;
}

static double chan_kv31(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:1043-1100
    float V = s->V_m;
    float area = eff_area(m, s);
    float mi = 1.0f / (1.0f + exp(-(V - 10.0f) / (10.0f + 1e-6f)));
    float tau_m = 1.0f;
    m->state[0] += (mi - m->state[0]) * dt / tau_m;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float hi = 1.0f / (1.0f + exp((V + 30.0f) / 6.0f));
    float tau_h = 2000.0f;
    m->state[1] += (hi - m->state[1]) * dt / tau_h;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(27.0f, 1.0f, kCbarK);
    double op = m->state[0] * m->state[0] * m->state[0] * m->state[1];
    double tea = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    op /= (1.0f + tea / 0.2f);
    float pkc = first_state_of(s, T_PKC, 0, 0.0f);
    float kp = 5.6e-6f, kd = 1e-4f;
    m->state[2] += (kp * pkc * (1.0f - m->state[2]) - kd * m->state[2]) * dt;
    if (m->state[2] < 0.0f) m->state[2] = 0.0f;
    if (m->state[2] > 1.0f) m->state[2] = 1.0f;
    op *= (1.0f - 0.27f * m->state[2]);
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * op;
    return k_out(s, ia * kAtoPa, dt);
}

static double chan_kv11(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:1189-1231
    float V = s->V_m;
    float area = eff_area(m, s);
    float vh = -30.0f, kk = 6.0f, gamma = 10.0f;
    float mi = 1.0f / (1.0f + exp(-(V - vh) / (kk + 1e-6f)));
    float tau_m = kdr_tau(V, vh);
    m->state[0] += (mi - m->state[0]) * dt / tau_m;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float hi = 1.0f / (1.0f + exp((V + 40.0f) / 6.0f));
    float tau_h = 3000.0f;
    m->state[1] += (hi - m->state[1]) * dt / tau_h;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    double op = m->state[0] * m->state[0] * m->state[0] * m->state[0] * m->state[1];
    double dtx = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    double tea = (m->state[6] > 0.0f) ? m->state[6] : 0.0f;
    op /= (1.0f + dtx / 2.0f) * (1.0f + tea / 170.0f);
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * op;
    return k_out(s, ia * kAtoPa, dt);
}

static double chan_kvdr(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:2763-2798
    float V = s->V_m;
    float area = eff_area(m, s);
    double mv = m->state[0];
    double vh = (m->state[1] != 0.0f) ? m->state[1] : -30.0f;
    double kk = (m->state[2] > 0.0f) ? m->state[2] : 6.0f;
    double tau = (m->state[4] > 0.0f) ? m->state[4] : kdr_tau(V, vh);
    float mi = 1.0f / (1.0f + exp(-(V - vh) / (kk + 1e-6f)));
    mv += (mi - mv) * dt / (tau + 1e-6f);
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m->state[0] = mv;
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(12.0f, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * mv;
    double ik = ia * kAtoPa;
    double vol = volume(s);
    float cpmv = kFaraday * vol / 1000.0;   // float in the reference
    s->intra_ions[1] -= (ik * dt) / cpmv;
    s->extra_ions[1] += (ik * dt) / (cpmv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -ik;
}

static double chan_sk(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:2859-2888
    float V = s->V_m;
    float area = eff_area(m, s);
    double op = m->state[0];
    double gamma = (m->state[3] > 0.0f) ? m->state[3] : 10.0f;
    double t_open = (m->state[4] > 0.0f) ? m->state[4] : 5.0f;
    double t_close = (m->state[5] > 0.0f) ? m->state[5] : 50.0f;
    float ca = s->intra_ions[2];
    float pi = sk_pinf(ca);
    float tau = (pi > op) ? t_open : t_close;
    op += (pi - op) * dt / tau;
    if (op < 0.0f) op = 0.0f;
    if (op > 1.0f) op = 1.0f;
    m->state[0] = op;
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * op;
    float ik = (float)(ia * kAtoPa);
    double vol = volume(s);
    double cpv = kFaraday * vol / 1000.0f;
    s->intra_ions[1] -= (ik * dt) / cpv;
    s->extra_ions[1] += (ik * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    return -ik;
}

static double chan_bk(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:2922-2949
    float V = s->V_m;
    float area = eff_area(m, s);
    double op = m->state[0];
    double gamma = (m->state[1] > 0.0f) ? m->state[1] : 250.0f;
    double tau = (m->state[2] > 0.0f) ? m->state[2] : 1.0f;
    float ca = s->intra_ions[2];
    float pi = bk_po(V, ca);
    op += (pi - op) * dt / tau;
    if (op < 0.0f) op = 0.0f;
    if (op > 1.0f) op = 1.0f;
    m->state[0] = op;
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * op;
    float ik = (float)(ia * kAtoPa);
    double vol = volume(s);
    double cpv = kFaraday * vol / 1000.0f;
    s->intra_ions[1] -= (ik * dt) / cpv;
    s->extra_ions[1] += (ik * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    return -ik;
}

static double chan_k2p(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:3063-3085
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[0] > 0.0f) ? m->state[0] : 100.0f;
    double po = (m->state[1] > 0.0f) ? m->state[1] : 0.012f;
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * po;
    return k_out(s, ia * kAtoPa, dt);
}

static double chan_kv7(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:3158-3220
    float V = s->V_m;
    float area = eff_area(m, s);
    double mv = m->state[0];
    double vh = (m->state[1] != 0.0f) ? m->state[1] : -38.0f;
    double ks = (m->state[2] > 0.0f) ? m->state[2] : 10.0f;
    double t_act = (m->state[4] > 0.0f) ? m->state[4] : 100.0f;
    double t_deact = (m->state[5] > 0.0f) ? m->state[5] : 60.0f;
    double pip2 = (m->state[6] >= 0.0f) ? m->state[6] : 1.0f;
    float mi = 1.0f / (1.0f + exp(-(V - vh) / ks));
    float tau_m = t_deact + (t_act - t_deact) * mi;
    if (tau_m < 1.0f) tau_m = 1.0f;
    mv += (mi - mv) * dt / tau_m;
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m->state[0] = mv;
    if (pip2 < 0.01f) return 0.0f;
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(7.0f, 1.0f, kCbarK);
    float pkc = first_state_of(s, T_PKC, 0, 0.0f);
    float sup = 1.0f - 0.8f * pkc;
    if (sup < 0.0f) sup = 0.0f;
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * (mv * pip2 * sup);
    return k_out(s, ia * kAtoPa, dt);
}

// HCN1 (129: V1/2 -69 mV, tau 84 ms, 0.84 pS) and HCN2 (130: -95 mV, 557 ms, 1.54 pS)
// (modifier_helpers.h:1442-1484, hh_biological.cpp:3598-3599)
static double chan_hcn1(Modifier* m, Segment* s, float dt, float vh0 = -69.0f, float topen0 = 84.0f,
                        float gamma0 = 0.84f) {
    double V = s->V_m;
    float area = eff_area(m, s);
    float vh = (m->state[1] != 0) ? m->state[1] : vh0;
    float k = (m->state[2] > 0) ? m->state[2] : 6.0f;
    float gamma = (m->state[3] > 0) ? m->state[3] : gamma0;
    float t_open = (m->state[4] > 0) ? m->state[4] : topen0;
    float t_close = (m->state[5] > 0) ? m->state[5] : 300.0f;
    float camp = first_state_of(s, T_CAMP_PKA, 0, 0.1f);
    float ve = vh + 15.0f * camp / (camp + 0.5f);
    float mv = m->state[0];
    float vf = V;
    float mi = 1.0f / (1.0f + exp((vf - ve) / k));
    float tau = (V < ve) ? (t_open * exp((V - ve) / 15.0f)) : t_close;
    mv += (mi - mv) * dt / tau;
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m->state[0] = mv;
    double pna, pk, pca;
    p_partial(gamma, 0.2f, 0.8f, 0.0f, &pna, &pk, &pca);
    double N = m->density * area;
    double fna = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double fk = ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    double in = pna * fna * N * mv * kAtoPa;
    double ik = pk * fk * N * mv * kAtoPa;
    double vol = volume(s);
    double cpm = kFaraday * vol / 1000.0;
    s->intra_ions[0] -= (in * dt) / cpm;
    s->intra_ions[1] -= (ik * dt) / cpm;
    s->extra_ions[0] += (in * dt) / (cpm * kEvr);
    s->extra_ions[1] += (ik * dt) / (cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -(in + ik);
}

static double chan_nalcn(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:3609-3641
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[0] > 0.0f) ? m->state[0] : 27.0f;
    double pna, pk, pca;
    p_nalcn(gamma, &pna, &pk, &pca);
    double N = m->density * area;
    float ina = (float)(pna * ghk_mono(V, s->intra_ions[0], s->extra_ions[0]) * N * kAtoPa);
    float ik = (float)(pk * ghk_mono(V, s->intra_ions[1], s->extra_ions[1]) * N * kAtoPa);
    float ica = (float)(pca * ghk_di(V, s->intra_ions[2], s->extra_ions[2]) * N * kAtoPa);
    double vol = volume(s);
    double cpm = kFaraday * vol / 1000.0;
    s->intra_ions[0] -= (ina * dt) / cpm;
    s->intra_ions[1] -= (ik * dt) / cpm;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    s->extra_ions[0] += (ina * dt) / (cpm * kEvr);
    s->extra_ions[1] += (ik * dt) / (cpm * kEvr);
    s->extra_ions[2] += (ica * dt) / (2.0f * cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -(ina - ik + ica);
}

// Ionotropic receptor (modifier_helpers.h:1213-1276).
struct Iono { float gamma, ec50, t_rise, t_decay; int nt; float pna, pk, pca; bool anion; float t_desens; };
static constexpr Iono kAmpa = {8.0f, 0.5f, 0.5f, 3.0f, 0, 0.5f, 0.5f, 0.0f, false, 3.0f};
static constexpr Iono kNmda = {50.0f, 1.0f, 5.0f, 50.0f, 0, 0.08f, 0.08f, 0.84f, false, 0.0f};
static constexpr Iono kGabaa = {28.0f, 1.0f, 1.0f, 5.0f, 1, 0.0f, 0.0f, 0.0f, true, 2000.0f};
// receptors_biological.cpp:57-100 with units.h:160-275: kainate (10), glycine (12), nicotinic
// alpha4beta2 (13), 5-HT3 (14), alpha7 (97); GluN2B NMDA (194) is NMDA with slower kinetics.
static constexpr Iono kKainate = {3.0f, 0.5f, 1.0f, 10.0f, 0, 0.5f, 0.5f, 0.0f, false, 5.0f};
static constexpr Iono kGlycine = {60.0f, 0.5f, 0.3f, 5.0f, 7, 0.0f, 0.0f, 0.0f, true, 2000.0f};
static constexpr Iono kNicotinic = {30.0f, 0.5f, 1.0f, 10.0f, 4, 0.48f, 0.48f, 0.04f, false, 2000.0f};
static constexpr Iono k5ht3 = {2.0f, 0.3f, 1.0f, 10.0f, 3, 0.40f, 0.40f, 0.20f, false, 800.0f};
static constexpr Iono kAlpha7 = {40.0f, 0.5f, 0.5f, 3.0f, 4, 0.12f, 0.17f, 0.71f, false, 50.0f};
static constexpr Iono kNmda2b = {50.0f, 1.0f, 10.0f, 150.0f, 0, 0.08f, 0.08f, 0.84f, false, 0.0f};
// P2X (15) and P2X2 (134) on ATP (receptors_biological.cpp:102-111, units.h:173, 275-292).
static constexpr Iono kP2x = {15.0f, 0.5f, 1.0f, 20.0f, 10, 0.22f, 0.22f, 0.56f, false, 20.0f};
static constexpr Iono kP2x2 = {15.0f, 0.5f, 1.0f, 20.0f, 10, 0.22f, 0.22f, 0.56f, false, 5000.0f};

static double iono_current(Modifier* m, Segment* s, float dt, const Iono& p, float scale = 1.0f) {
    float area = eff_area(m, s);
    double V = s->V_m;
    double vol = volume(s);
    double cpm = kFaraday * vol / 1000.0;
    float gamma = (m->active_vars > 0 && m->state[m->active_vars] > 0.0f) ? m->state[m->active_vars] : p.gamma;
    float nt = s->neurotransmitter[p.nt];
    float r_inf = nt / (nt + p.ec50);
    float r = m->state[0];
    float tau = (r_inf > r) ? p.t_rise : p.t_decay;
    m->state[0] += (r_inf - r) * dt / tau;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float op = m->state[0];
    if (p.t_desens > 0.0f) {
        float kd = p.ec50;
        float d = m->state[1];
        float d_inf = nt / (nt + kd);
        d += (d_inf - d) * dt / p.t_desens;
        if (d < 0.0f) d = 0.0f;
        if (d > 1.0f) d = 1.0f;
        m->state[1] = d;
        op *= (1.0f - d);
    }
    double N = m->density * area;
    if (p.anion) {
        double pcl = p_single(gamma, 1.0f, kCbarCl);
        double fcl = ghk_an(V, s->intra_ions[3], s->extra_ions[3]);
        double icl = pcl * fcl * N * op * scale * kAtoPa;
        s->intra_ions[3] += (icl * dt) / cpm;
        floor_ions(s->intra_ions);
        return -icl;
    }
    double pna, pk, pca;
    p_partial(gamma, p.pna, p.pk, p.pca, &pna, &pk, &pca);
    double ina = pna * ghk_mono(V, s->intra_ions[0], s->extra_ions[0]) * N * op * scale * kAtoPa;
    double ik = pk * ghk_mono(V, s->intra_ions[1], s->extra_ions[1]) * N * op * scale * kAtoPa;
    double itot = ina + ik;
    s->intra_ions[0] -= (ina * dt) / cpm;
    s->intra_ions[1] -= (ik * dt) / cpm;
    if (p.pca > 0.0f) {
        double ica = pca * ghk_di(V, s->intra_ions[2], s->extra_ions[2]) * N * op * scale * kAtoPa;
        itot += ica;
        s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    }
    floor_ions(s->intra_ions);
    return -itot;
}
static double nmda_cdi(Modifier* m, Segment* s, float dt) {   // modifier_helpers.h:1281-1291
    double ca_um = s->intra_ions[2] * kMmToUm;
    double cdi_inf = 1.0 - 0.8f * ca_um / (ca_um + 1.0f);
    float tau = (cdi_inf < m->state[2]) ? 500.0f : 9000.0f;
    m->state[2] += (cdi_inf - m->state[2]) * dt / tau;
    if (m->state[2] < 0.0f) m->state[2] = 0.0f;
    if (m->state[2] > 1.0f) m->state[2] = 1.0f;
    float coag = 10.0f / (10.0f + 1.0f);
    return m->state[2] * coag;
}
static double rec_nmda(Modifier* m, Segment* s, float dt, const Iono& k = kNmda) {   // receptors_biological.cpp:48-62
    double I = iono_current(m, s, dt, k);
    float mg = 1.0f / (1.0f + (1.0f / 3.57f) * exp(-0.062f * s->V_m));
    double r = I * mg;
    return r * nmda_cdi(m, s, dt);
}

// The network-builder NMDA variants (63, 72; receptors_biological.cpp:222-229, 312-319): NMDA
// gating without the Ca-dependent inactivation, Mg2+ block by the extracellular Mg2+ (1 mM when
// below 0.01).
// MSN NMDA (75, receptors_biological.cpp:337-352) and MSN AMPA (76, :357-369): NMDA with the Mg
// block and AMPA, their current scaled by state[1] of the segment's first D1 (65) or D2 (66)
// modulator, floored at 0 (1 without one).
static float msn_scale(const Segment* s, uint8_t want) {
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == want) {
            float v = s->modifiers[j].state[1];
            if (v < 0.0f) v = 0.0f;
            return v;
        }
    return 1.0f;
}
static double rec_nmda_mg(Modifier* m, Segment* s, float dt, float scale = 1.0f) {
    double I = iono_current(m, s, dt, kNmda, scale);
    float mg = s->extra_ions[4];
    if (mg < 0.01f) mg = 1.0f;
    float block = 1.0f / (1.0f + (mg / 3.57f) * exp(-0.062f * s->V_m));
    return I * block;
}

// Gi/o-coupled receptors (17-40, modifier_helpers.h:1332-1363, receptors_biological.cpp:163-186):
// transmitter occupancy drives a GIRK gate (state[0]) with desensitisation (state[1]); EC50 and
// gamma come from state[active_vars] and state[active_vars + 1] when positive. The transmitter
// each one reads: GABA-B GABA; 5-HT1A/B/D/F serotonin; D2-4 dopamine; M2/M4 acetylcholine;
// A1/A3 adenosine; mGluR2-8 glutamate; alpha2 noradrenaline; CB1/2 endocannabinoid; H3
// histamine; MT1/2 melatonin.
static int gio_nt(uint8_t t) {
    if (t == 17) return 1;
    if (t <= 21) return 3;
    if (t <= 24) return 2;
    if (t <= 26) return 4;
    if (t <= 28) return 5;
    if (t <= 34) return 0;
    if (t == 35) return 6;
    if (t <= 37) return 9;
    if (t == 38) return 8;
    return 11;
}
static float kir_f_open(float V);
static double chan_gio(Modifier* m, Segment* s, float dt, int nt_index) {
    float area = eff_area(m, s);
    double V = s->V_m;
    double cpm = kFaraday * volume(s) / 1000.0;
    if (m->active_vars >= 7) g_unsupported++;
    float ec50 = (m->state[m->active_vars] > 0.0f) ? m->state[m->active_vars] : 1.0f;
    float gamma = (m->state[m->active_vars + 1] > 0.0f) ? m->state[m->active_vars + 1] : 35.0f;
    float nt = s->neurotransmitter[nt_index];
    float act = nt / (nt + ec50);
    float r = m->state[0];
    float tau = (act > r) ? 70.0f : 1500.0f;
    r += (act - r) * dt / tau;
    if (r < 0.0f) r = 0.0f;
    if (r > 1.0f) r = 1.0f;
    m->state[0] = r;
    float d = m->state[1];
    d += (act - d) * dt / 40000.0f;
    if (d < 0.0f) d = 0.0f;
    if (d > 1.0f) d = 1.0f;
    m->state[1] = d;
    float fo = kir_f_open(V) * r * (1.0f - 0.55f * d);
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    double ia = P * flux * (m->density * area) * fo;
    double ik = ia * kAtoPa;
    s->intra_ions[1] -= (ik * dt) / cpm;
    floor_ions(s->intra_ions);
    return -ik;
}

// Gs-coupled receptors (41-51, modifier_helpers.h:1379-1398): occupancy drives the activation
// (state[0]) and desensitisation (state[1]) signals; no current. 5-HT4/6/7 read serotonin,
// D1/D5 and TAAR1 dopamine, A2A/A2B adenosine, beta1/2 noradrenaline, H2 histamine.
static int gs_nt(uint8_t t) {
    if (t <= 43) return 3;
    if (t <= 45) return 2;
    if (t <= 47) return 5;
    if (t <= 49) return 6;
    if (t == 50) return 8;
    return 2;
}
static double chan_gs(Modifier* m, Segment* s, float dt, int nt_index) {
    if (m->active_vars >= 7) g_unsupported++;
    float ec50 = (m->state[m->active_vars] > 0.0f) ? m->state[m->active_vars] : 1.0f;
    float tdes = (m->state[m->active_vars + 1] > 0.0f) ? m->state[m->active_vars + 1] : 70000.0f;
    float nt = s->neurotransmitter[nt_index];
    float occ = nt / (nt + ec50);
    float act = m->state[0];
    float tau = (occ > act) ? 70.0f : 1500.0f;
    act += (occ - act) * dt / tau;
    if (act < 0.0f) act = 0.0f;
    if (act > 1.0f) act = 1.0f;
    m->state[0] = act;
    float d = m->state[1];
    d += (occ - d) * dt / tdes;
    if (d < 0.0f) d = 0.0f;
    if (d > 1.0f) d = 1.0f;
    m->state[1] = d;
    return 0.0f;
}

// Connexin Cx36 (114) and Cx43 (198) (hh_biological.cpp:3266-3300, 3310-3341, identical
// bodies): connexon turnover, CaMKII/PKA phosphorylation, pH-gated open probability. They
// carry no current; gap junctions read state[0] (connexons) and state[4] (open probability).
static double chan_connexin(Modifier* m, Segment* s, float dt) {
    double synth = (m->state[2] > 0.0f) ? m->state[2] : 0.001f;
    double degr = (m->state[3] > 0.0f) ? m->state[3] : 0.0005f;
    double pH = (m->state[5] > 4.0f) ? m->state[5] : 7.2f;
    m->state[0] += (synth - degr * m->state[0]) * dt / 3600000.0f;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float camkii = 0.0f;
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == T_CAMKII) camkii = s->modifiers[j].state[0];
    double pka = m->state[6];
    float p_target = camkii * 0.6f + pka * 0.3f;
    m->state[1] += (p_target - m->state[1]) * dt / 5000.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    float ph_f = (pH < 6.8f) ? (pH - 6.0f) / 0.8f : 1.0f;
    if (ph_f < 0.0f) ph_f = 0.0f;
    double p_cam = 1.0f - m->state[1] * 0.5f;
    double p_pka = 1.0f + m->state[1] * 0.3f;
    m->state[4] = (p_cam * 0.5f + p_pka * 0.5f) * ph_f;
    if (m->state[4] < 0.05f) m->state[4] = 0.05f;
    if (m->state[4] > 2.0f) m->state[4] = 2.0f;
    return 0.0f;
}

// Cav1.3 (1, hh_biological.cpp:142-185): L-type m^2 h, fixed time constants, Ca-dependent h.
// Extracellular Ca decreases with the intracellular change, as the reference writes it.
static double chan_cav13(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double vh = (m->state[3] != 0.0f) ? -fabs(m->state[3]) : -35.0f;
    double k = (m->state[4] > 0.0f) ? m->state[4] : 7.0f;
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 25.0f;
    float mi = 1.0f / (1.0f + exp(-(V - vh) / (k + 1e-6f)));
    m->state[0] += (mi - m->state[0]) * dt / 2.0f;
    float ca = s->intra_ions[2] * kMmToUm;
    float hi = 1.0f / (1.0f + exp((V + 52.0f) / 8.0f));
    hi *= 1.0f / (1.0f + ca / 3.0f);
    m->state[1] += (hi - m->state[1]) * dt / 100.0f;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    float ci = s->intra_ions[2], co = s->extra_ions[2];
    double P = p_single(gamma, 2.0f, kCbarCa);
    double op = m->state[0] * m->state[0] * m->state[1];
    double flux = ghk_di(V, ci, co);
    double ia = P * flux * (m->density * area) * op;
    double ica = ia * kAtoPa;
    double cpm = kFaraday * volume(s) / 1000.0;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    s->extra_ions[2] -= (ica * dt) / (2.0f * cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -ica;
}

// ClC-2 leak (2, hh_biological.cpp:390-414): open state[0] forced to 1 when below 0.5, GHK
// anion current; the charge-per-mole factor is a float here.
static double chan_cl(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    if (m->state[0] < 0.5f) m->state[0] = 1.0f;
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 2.5f;
    float ci = s->intra_ions[3], co = s->extra_ions[3];
    double P = p_single(gamma, 1.0f, kCbarCl);
    double flux = ghk_an(V, ci, co);
    double ia = P * flux * (m->density * area) * m->state[0];
    double icl = ia * kAtoPa;
    float vol = volume(s);
    float cpm = kFaraday * vol / 1000.0;
    s->intra_ions[3] += (icl * dt) / cpm;
    s->extra_ions[3] -= (icl * dt) / (cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -icl;
}

// Kv4.2 A-type (85, hh_biological.cpp:789-836, modifier_helpers.h:463-472): m^3 h, V-dependent
// recovery, CaMKII (state[0] + state[1] of the last CaMKII modifier) phosphorylation in state[3].
static double chan_kv42(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double vh = (m->state[4] != 0.0f) ? -fabs(m->state[4]) : -10.0f;
    double k = (m->state[5] > 0.0f) ? m->state[5] : 13.5f;
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 5.0f;
    float mi = 1.0f / (1.0f + exp(-(V - vh) / (k + 1e-6f)));
    m->state[0] += (mi - m->state[0]) * dt / 1.5f;
    float hi = 1.0f / (1.0f + exp((V + 65.0f) / 8.0f));
    float th = 15.0f + (170.0f - 15.0f) / (1.0f + exp((V - -60.0f) / 8.0f));
    m->state[1] += (hi - m->state[1]) * dt / th;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    float ck = 0.0f;
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == T_CAMKII) ck = s->modifiers[j].state[0] + s->modifiers[j].state[1];
    m->state[3] += (ck * 0.5f - m->state[3]) * dt / 5000.0f;
    if (m->state[3] < 0.0f) m->state[3] = 0.0f;
    if (m->state[3] > 1.0f) m->state[3] = 1.0f;
    double geff = gamma * (1.0f - m->state[3] * 0.7f);
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(geff, 1.0f, kCbarK);
    double op = m->state[0] * m->state[0] * m->state[0] * m->state[1];
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * op;
    double ik = ia * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0;
    s->intra_ions[1] -= (ik * dt) / cpv;
    s->extra_ions[1] += (ik * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    return -ik;
}

// Kv2.1 delayed rectifier (196, hh_biological.cpp:1392-1428, modifier_helpers.h:584-620): m^4 h
// with a U-shaped inactivation, TEA block by state[5].
static double chan_kv21(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 8.0f;
    float mi = 1.0f / (1.0f + exp(-(V - 0.0f) / 9.4f));
    float tm = 15.0f + (73.0f - 15.0f) / (1.0f + exp((V - 0.0f) / 10.0f));
    m->state[0] += (mi - m->state[0]) * dt / tm;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float left = 1.0f / (1.0f + exp(-(V - -26.5f) / 5.2f));
    float right = 1.0f / (1.0f + exp((V - 15.0f) / 10.0f));
    float hi = 1.0f - (1.0f - 0.2f) * left * right;
    float th = 4000.0f + (300.0f - 4000.0f) / (1.0f + exp((V - -80.0f) / 15.0f));
    m->state[1] += (hi - m->state[1]) * dt / th;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    double op = m->state[0] * m->state[0] * m->state[0] * m->state[0] * m->state[1];
    double tea = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    op /= (1.0f + tea / 10.0f);
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * op;
    double ik = ia * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0;
    s->intra_ions[1] -= (ik * dt) / cpv;
    s->extra_ions[1] += (ik * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    return -ik;
}

// Persistent Na+ (105: Nav1.6 NaP, hh_biological.cpp:2951-2978) and the tadpole NaP (195:
// Nav1.2 gamma, V1/2 -40 mV, slow inactivation in state[6], Na+ taken from the extracellular
// pool, hh_biological.cpp:2980-3016).
static double chan_nap(Modifier* m, Segment* s, float dt, bool tadpole) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double mv = m->state[0];
    double vh = (m->state[1] != 0.0f) ? m->state[1] : (tadpole ? -40.0f : -58.7f);
    double km = (m->state[2] > 0.0f) ? m->state[2] : 6.0f;
    double gamma = (m->state[3] > 0.0f) ? m->state[3] : (tadpole ? 20.8f : 20.7f);
    double tau = (m->state[4] > 0.0f) ? m->state[4] : 5.0f;
    double po = (m->state[5] > 0.0f) ? m->state[5] : 0.016f;
    float vhf = vh, kmf = km;
    float mi = 1.0f / (1.0f + exp(-(V - vhf) / kmf));
    mv += (mi - mv) * dt / tau;
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m->state[0] = mv;
    float nai = s->intra_ions[0], nao = s->extra_ions[0];
    double P = p_single(gamma, 1.0f, kCbarNa);
    double flux = ghk_mono(V, nai, nao);
    if (!tadpole) {
        double ia = P * flux * (m->density * area) * (mv * po);
        float ina = (float)(ia * kAtoPa);
        double cpv = kFaraday * volume(s) / 1000.0f;
        s->intra_ions[0] -= (ina * dt) / cpv;
        floor_ions(s->intra_ions);
        return -ina;
    }
    float si = 1.0f / (1.0f + exp((V - -50.0f) / 6.0f));
    double sv = m->state[6];
    sv += (si - sv) * dt / 500.0f;
    if (sv < 0.0f) sv = 0.0f;
    if (sv > 1.0f) sv = 1.0f;
    m->state[6] = sv;
    double ia = P * flux * (m->density * area) * (mv * po * sv);
    double ina = ia * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[0] -= (ina * dt) / cpv;
    s->extra_ions[0] += (ina * dt) / (cpv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -ina;
}

// Kir2.1 (106, hh_biological.cpp:3018-3037, modifier_helpers.h:870-891): Woodhull Mg2+ and
// three-site spermine block.
static inline float kir_f_open(float V) {
    float vn = kFaraday * V * 0.001f / (kGas * kTemp);
    float kd_mg = 0.3f * exp(-0.6f * 2.0f * vn);
    float mg = 0.5f / kd_mg;
    float kd_spm = 0.0775f * exp(-0.5f * 4.0f * vn);
    float r = 0.001f / kd_spm;
    float spm = 0.0f, rk = r;
    for (int k = 0; k < 3; k++) { spm += rk; rk *= r; }
    return 1.0f / (1.0f + mg + spm);
}
static double chan_kir(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[3] > 0.0f) ? m->state[3] : 30.0f;
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    float fo = kir_f_open(V);
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    double ia = P * flux * (m->density * area) * fo;
    double ik = ia * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[1] -= (ik * dt) / cpv;
    s->extra_ions[1] += (ik * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    return -ik;
}

// Developmental Nav1.2 (193, hh_biological.cpp:314-357, modifier_helpers.h:221-241): m^3 with
// fast and slow inactivation (fraction state[4], default 0.7; slow recovery 500 ms), V1/2 of
// inactivation from state[2] (default 45, i.e. -45 mV), TTX and PKA/PKC block.
static double chan_nav12dev(Modifier* m, Segment* s, float dt) {
    static constexpr NavGates kDev = {16.0f, 6.0f, 41.5f, 66.5f, 45.0f};
    float V = s->V_m;
    float area = eff_area(m, s);
    if (s->V_m < -200.0f) s->V_m = -200.0f;
    if (s->V_m > 200.0f) s->V_m = 200.0f;
    V = s->V_m;
    double mv = nav_minf(kDev, V) + (m->state[0] - nav_minf(kDev, V)) * exp(-dt / nav_mtau(kDev, V));
    double vhalf = (m->state[2] > 0.0f) ? m->state[2] : 45.0f;
    float vhf = vhalf;
    float hi = 1.0f / (1.0f + exp((V + vhf) / 6.2f));
    double hf = hi + (m->state[1] - hi) * exp(-dt / nav_htau(kDev, V));
    double ts = (hi < m->state[3]) ? nav_htau(kDev, V) : 500.0f;
    double hs = hi + (m->state[3] - hi) * exp(-dt / ts);
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hf < 0.0f) hf = 0.0f;
    if (hf > 1.0f) hf = 1.0f;
    if (hs < 0.0f) hs = 0.0f;
    if (hs > 1.0f) hs = 1.0f;
    m->state[0] = mv;
    m->state[1] = hf;
    m->state[3] = hs;
    double sf = (m->state[4] > 0.0f) ? m->state[4] : 0.7f;
    float he = (1.0f - sf) * hf + sf * hs;
    float op = mv * mv * mv * he;
    double ttx = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    op /= (1.0f + ttx / 10.0f);
    float pkc = first_state_of(s, T_PKC, 0, 0.0f);
    float pka = first_state_of(s, T_CAMP_PKA, 1, 0.0f);
    op *= (1.0f - 0.7f * (pkc + pka) * 0.5f);
    if (op < 0.0f) op = 0.0f;
    double P = p_single(20.8f, 1.0f, kCbarNa);
    double flux = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double ia = P * flux * (m->density * area) * op;
    double ina = ia * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0;
    s->intra_ions[0] -= (ina * dt) / cpv;
    s->extra_ions[0] += (ina * dt) / (cpv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -ina;
}

// Nav1.3 (177), Nav1.7 (169), Nav1.8 (170), Nav1.9 (171) (hh_biological.cpp nav13/17/18/19):
// m^3 h with Boltzmann steady states and fixed time constants, TTX block by state[5].
struct NavSimple { float vm, km, vh, kh, tm, th, ttx, gamma; };
static constexpr NavSimple kNav13 = {-22.0f, 6.5f, -64.0f, 6.5f, 0.5f, 0.65f, 4.0f, 20.8f};
static constexpr NavSimple kNav17 = {-24.0f, 5.0f, -71.0f, 6.0f, 0.5f, 1.2f, 10.0f, 19.5f};
static constexpr NavSimple kNav18 = {-20.0f, 7.0f, -30.0f, 5.0f, 1.5f, 8.0f, 60000.0f, 20.0f};
static constexpr NavSimple kNav19 = {-47.0f, 5.5f, -44.0f, 6.0f, 10.0f, 43.0f, 39000.0f, 16.0f};
static double chan_nav_simple(Modifier* m, Segment* s, float dt, const NavSimple& g) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float mi = 1.0f / (1.0f + exp(-(V - g.vm) / g.km));
    float hi = 1.0f / (1.0f + exp((V - g.vh) / g.kh));
    m->state[0] += (mi - m->state[0]) * (dt / g.tm);
    m->state[1] += (hi - m->state[1]) * (dt / g.th);
    if (m->state[0] < 0) m->state[0] = 0;
    if (m->state[0] > 1) m->state[0] = 1;
    if (m->state[1] < 0) m->state[1] = 0;
    if (m->state[1] > 1) m->state[1] = 1;
    double op = m->state[0] * m->state[0] * m->state[0] * m->state[1];
    double ttx = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    op /= (1.0f + ttx / g.ttx);
    double P = p_single(g.gamma, 1.0f, kCbarNa);
    double flux = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double ia = P * flux * (m->density * area) * op;
    float I = (float)(ia * kAtoPa);
    double cpv = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[0] -= (I * dt) / cpv;
    s->extra_ions[0] += (I * dt) / (cpv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -I;
}

// Nav1.1 (199, hh_biological.cpp nav11_current, modifier_helpers.h:186-205): HH rates with the
// Nav1.1 offsets, inactivation slope 5 mV, TTX block; no PKA/PKC term.
static double chan_nav11(Modifier* m, Segment* s, float dt) {
    static constexpr NavGates k11 = {30.0f, 6.5f, 45.0f, 65.0f, 65.0f};
    float V = s->V_m;
    float area = eff_area(m, s);
    if (s->V_m < -200.0f) s->V_m = -200.0f;
    if (s->V_m > 200.0f) s->V_m = 200.0f;
    V = s->V_m;
    float hinf = 1.0f / (1.0f + exp((V + 65.0f) / 5.0f));
    double mv = nav_minf(k11, V) + (m->state[0] - nav_minf(k11, V)) * exp(-dt / nav_mtau(k11, V));
    double hv = hinf + (m->state[1] - hinf) * exp(-dt / nav_htau(k11, V));
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m->state[0] = mv;
    m->state[1] = hv;
    float op = mv * mv * mv * hv;
    double ttx = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    op /= (1.0f + ttx / 10.0f);
    double P = p_single(17.0f, 1.0f, kCbarNa);
    double flux = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double ia = P * flux * (m->density * area) * op;
    double ina = ia * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0;
    s->intra_ions[0] -= (ina * dt) / cpv;
    s->extra_ions[0] += (ina * dt) / (cpv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -ina;
}

// Kv1.2 (157), Kv1.3 (154), Kv1.4 (206), Kv1.6 (207), Kv2.2 (205) (hh_biological.cpp kv12/kv13/
// kv14/kv16/kv22, modifier_helpers.h:518-558, units.h:354-364): Kv1.2 and Kv1.3 m^4 h with the
// delayed-rectifier activation time constant (PKC block of Kv1.2; DTX/TEA and ShK toxin blocks),
// Kv1.4 m^3 h in ten substeps with a V-dependent recovery, Kv1.6 m^3, Kv2.2 m^3 h; pore blockers
// from the states.
static double k_efflux(Segment* s, double ik, float dt) {
    double cpv = kFaraday * volume(s) / 1000.0;
    s->intra_ions[1] -= (ik * dt) / cpv;
    s->extra_ions[1] += (ik * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    return -ik;
}
static double chan_kv1x(Modifier* m, Segment* s, float dt, bool kv12) {
    float V = s->V_m;
    float area = eff_area(m, s);
    const float vh = kv12 ? -27.0f : -30.0f, k = 6.0f, gamma = kv12 ? 16.0f : 19.0f;
    float mi = 1.0f / (1.0f + exp(-(V - vh) / (k + 1e-6f)));
    float tm = kdr_tau(V, vh);
    m->state[0] += (mi - m->state[0]) * dt / tm;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float hi = 1.0f / (1.0f + exp((V + 40.0f) / 6.0f));
    float th = kv12 ? 400.0f : 3000.0f;
    m->state[1] += (hi - m->state[1]) * dt / th;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    double op = m->state[0] * m->state[0] * m->state[0] * m->state[0] * m->state[1];
    if (kv12) {
        float pkc = first_state_of(s, T_PKC, 0, 0.0f);
        op *= (1.0f - 0.8f * pkc);
        if (op < 0.0f) op = 0.0f;
        double dtx = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        double tea = (m->state[6] > 0.0f) ? m->state[6] : 0.0f;
        op /= (1.0f + dtx / 5.0f) * (1.0f + tea / 170.0f);
    } else {
        double shk = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        op /= (1.0f + shk / 11.0f);
    }
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    return k_efflux(s, P * flux * (m->density * area) * op * kAtoPa, dt);
}
static double chan_kv14(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 9.3f;
    double mv = m->state[0], hv = m->state[1];
    float mi = 1.0f / (1.0f + exp(-(V - -21.7f) / 8.0f));
    float hi = 1.0f / (1.0f + exp((V - -51.8f) / 3.2f));
    float tm = 2.0f;
    float th = 50.0f + (1500.0f - 50.0f) / (1.0f + exp((V - -80.0f) / 10.0f));
    float sub = dt / 10.0f;
    for (int k = 0; k < 10; k++) {
        mv += (mi - mv) * sub / tm;
        hv += (hi - hv) * sub / th;
    }
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m->state[0] = mv;
    m->state[1] = hv;
    float op = mv * mv * mv * hv;
    double ap4 = (m->state[3] > 0.0f) ? m->state[3] : 0.0f;
    double tea = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    op /= (1.0f + ap4 / 13.0f) * (1.0f + tea / 100.0f);
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    return k_efflux(s, P * flux * (m->density * area) * op * kAtoPa, dt);
}
static double chan_kv16_22(Modifier* m, Segment* s, float dt, bool kv22) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : (kv22 ? 14.5f : 8.2f);
    double mv = m->state[0], hv = m->state[1];
    float mi = kv22 ? 1.0f / (1.0f + exp(-(V - 13.2f) / 9.0f)) : 1.0f / (1.0f + exp(-(V - -8.1f) / 13.0f));
    mv += (mi - mv) * dt / (kv22 ? 32.0f : 20.0f);
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    m->state[0] = mv;
    float op;
    if (kv22) {
        float hi = 1.0f / (1.0f + exp((V - -16.6f) / 6.0f));
        hv += (hi - hv) * dt / 2000.0f;
        if (hv < 0.0f) hv = 0.0f;
        if (hv > 1.0f) hv = 1.0f;
        m->state[1] = hv;
        op = mv * mv * mv * hv;
    } else {
        op = mv * mv * mv;
    }
    double tea = (m->state[3] > 0.0f) ? m->state[3] : 0.0f;
    double ap4 = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    op /= (1.0f + tea / (kv22 ? 2.6f : 1.7f)) * (1.0f + ap4 / (kv22 ? 1.5f : 0.3f));
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    return k_efflux(s, P * flux * (m->density * area) * op * kAtoPa, dt);
}

// Kv3.2 (203), Kv3.3 (204), Kv4.3 (202) (hh_biological.cpp kv32/kv33/kv43, modifier_helpers.h:
// 475-515): m^3 (h) with Boltzmann steady states, Kv4.3 in ten substeps with a V-dependent
// recovery; two pore blockers from state[3] and state[4].
struct KvGen { float vm, km, vh, kh, tm, th, gamma, ic3, ic4; bool inact, sub; };
static constexpr KvGen kKv32 = {10.0f, 8.0f, 0.0f, 1.0f, 12.0f, 0.0f, 15.0f, 63.0f, 25.0f, false, false};
static constexpr KvGen kKv33 = {1.5f, 8.0f, -27.9f, 8.0f, 4.0f, 200.0f, 35.0f, 67.5f, 100.0f, true, false};
static constexpr KvGen kKv43 = {-31.7f, 8.0f, -47.4f, 5.9f, 1.7f, 0.0f, 4.4f, 488.0f, 81.0f, true, true};
static double chan_kvgen(Modifier* m, Segment* s, float dt, const KvGen& g) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : g.gamma;
    double mv = m->state[0], hv = m->state[1];
    float mi = 1.0f / (1.0f + exp(-(V - g.vm) / g.km));
    float op;
    if (!g.inact) {
        mv += (mi - mv) * dt / g.tm;
        if (mv < 0.0f) mv = 0.0f;
        if (mv > 1.0f) mv = 1.0f;
        m->state[0] = mv;
        op = mv * mv * mv;
    } else {
        float hi = 1.0f / (1.0f + exp((V - g.vh) / g.kh));
        if (g.sub) {
            float th = 27.6f + (142.0f - 27.6f) / (1.0f + exp((V - -60.0f) / 8.0f));
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
        m->state[0] = mv;
        m->state[1] = hv;
        op = mv * mv * mv * hv;
    }
    double b3 = (m->state[3] > 0.0f) ? m->state[3] : 0.0f;
    double b4 = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    op /= (1.0f + b3 / g.ic3) * (1.0f + b4 / g.ic4);
    double P = p_single(gamma, 1.0f, kCbarK);
    double flux = ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    return k_efflux(s, P * flux * (m->density * area) * op * kAtoPa, dt);
}

// Leak-like K+ channels (hh_biological.cpp girk/katp/thik1/thik2/kir22): GIRK (160) open as
// G-beta-gamma state[4], tertiapin block state[5]; KATP (161) closed by ATP state[4],
// glibenclamide state[5]; THIK-1 (166) Po 0.5 with halothane block state[5]; THIK-2 (167)
// Po 0.5; Kir2.2 (208) Woodhull block with Ba2+ state[4].
static double chan_kleak(Modifier* m, Segment* s, float dt, uint8_t t) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float po;
    double gamma;
    if (t == 160) {
        double gbg = m->state[4];
        if (gbg < 0.0f) gbg = 0.0f;
        if (gbg > 1.0f) gbg = 1.0f;
        po = gbg;
        double tq = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        po /= (1.0f + tq / 8.0f);
        gamma = 35.0f;
    } else if (t == 161) {
        double atp = m->state[4];
        if (atp < 0.0f) atp = 0.0f;
        po = 1.0f / (1.0f + atp / 15.0f);
        double gl = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        po /= (1.0f + gl / 3.0f);
        gamma = 75.0f;
    } else if (t == 166 || t == 167) {
        po = 0.5f;
        if (t == 166) {
            double hal = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
            po /= (1.0f + hal / 2.8f);
        }
        gamma = 5.0f;
    } else {   // 208
        gamma = (m->state[3] > 0.0f) ? m->state[3] : 36.0f;
        po = kir_f_open(V);
        double ba = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
        po /= (1.0f + ba / 2.8f);
    }
    double P = p_single(gamma, 1.0f, kCbarK);
    // Kir2.2 passes the concentrations rounded to float
    double flux = (t == 208) ? ghk_mono(V, (float)s->intra_ions[1], (float)s->extra_ions[1])
                             : ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    return k_efflux(s, P * flux * (m->density * area) * po * kAtoPa, dt);
}

// Kir7.1 (178, always open, 0.05 pS) and KCNQ4 (176, m with tau 212 ms, 2.31 pS)
// (hh_biological.cpp kir71/kcnq4): float current, both ion arrays floored.
static double chan_kir71_kcnq4(Modifier* m, Segment* s, float dt, bool kcnq4) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double op = 1.0f;
    if (kcnq4) {
        float mi = 1.0f / (1.0f + exp(-(V - (-78.0f)) / 14.0f));
        m->state[0] += (mi - m->state[0]) * (dt / 212.0f);
        if (m->state[0] < 0) m->state[0] = 0;
        if (m->state[0] > 1) m->state[0] = 1;
        op = m->state[0];
    }
    double P = p_single(kcnq4 ? 2.31f : 0.05f, 1.0f, kCbarK);
    double flux = ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    float I = (float)(P * flux * (m->density * area) * op * kAtoPa);
    double cpv = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[1] -= (I * dt) / cpv;
    s->extra_ions[1] += (I * dt) / (cpv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -I;
}

// Cav2.2 N-type (168, hh_biological.cpp cav22_current): m^2 h, omega-conotoxin GVIA block
// state[5]; the intracellular Ca changes by the full current (no factor 2).
static double chan_cav22(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float mi = 1.0f / (1.0f + exp(-(V - -20.0f) / (6.0f + 1e-6f)));
    m->state[0] += (mi - m->state[0]) * dt / 1.7f;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float hi = 1.0f / (1.0f + exp((V + 60.0f) / 6.0f));
    m->state[1] += (hi - m->state[1]) * dt / 150.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    double op = m->state[0] * m->state[0] * m->state[1];
    double gvia = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    op /= (1.0f + gvia / 0.5f);
    double P = p_single(2.76f, 2.0f, kCbarCa);
    double flux = ghk_di(V, s->intra_ions[2], s->extra_ions[2]);
    double ica = P * flux * (m->density * area) * op * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0;
    s->intra_ions[2] -= (ica * dt) / cpv;
    s->extra_ions[2] += (ica * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    floor_ions(s->intra_ions);
    return -ica;
}

// T-type Cav3.1 (200), Cav3.2 (201), Cav3.3 (191) (hh_biological.cpp:2033-2143, modifier_helpers.h:
// 649-737, 1032-1036): m^3 h in ten substeps, a bell-shaped activation time constant and a
// V-dependent recovery, Ni2+/mibefradil/TTA-P2/Z944 blocks from states 3-6; the intracellular Ca
// changes with a float charge-per-mole factor and is not floored.
struct CavT { float vm, km, vh, kh, gamma, thf, ths, tmp, ni, mib; };
static constexpr CavT kCav31 = {-45.0f, 5.5f, -77.0f, 6.0f, 7.5f, 25.0f, 250.0f, 15.0f, 250.0f, 1.6f};
static constexpr CavT kCav32 = {-46.0f, 6.0f, -72.0f, 6.0f, 9.0f, 16.0f, 180.0f, 15.0f, 13.0f, 3.0f};
static constexpr CavT kCav33 = {-43.8f, 6.9f, -72.5f, 6.6f, 11.0f, 97.0f, 400.0f, 30.0f, 200.0f, 1.6f};
static double chan_cavt(Modifier* m, Segment* s, float dt, const CavT& g) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : g.gamma;
    double mv = m->state[0], hv = m->state[1];
    float mi = 1.0f / (1.0f + exp(-(V - g.vm) / g.km));
    float hi = 1.0f / (1.0f + exp((V - g.vh) / g.kh));
    float dv = V - -20.0f;
    float gg = exp(-(dv * dv) / (2.0f * 22.0f * 22.0f));
    float tm = 1.25f + (g.tmp - 1.25f) * gg;
    float th = g.thf + (g.ths - g.thf) / (1.0f + exp((V - -65.0f) / 8.0f));
    float sub = dt / 10.0f;
    for (int k = 0; k < 10; k++) {
        mv += (mi - mv) * sub / tm;
        hv += (hi - hv) * sub / th;
    }
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m->state[0] = mv;
    m->state[1] = hv;
    float op = mv * mv * mv * hv;
    double ni = (m->state[3] > 0.0f) ? m->state[3] : 0.0f;
    double mib = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    double tta = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    double z = (m->state[6] > 0.0f) ? m->state[6] : 0.0f;
    op /= (1.0f + ni / g.ni) * (1.0f + mib / g.mib) * (1.0f + tta / 82.0f) * (1.0f + z / 100.0f);
    double P = p_single(gamma, 2.0f, kCbarCa);
    double flux = ghk_di(V, s->intra_ions[2], s->extra_ions[2]);
    double ica = P * flux * (m->density * area) * op * kAtoPa;
    float vol = volume(s);
    float cpm = kFaraday * vol / 1000.0;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    return -ica;
}

// TRPC3 (210, DAG state[4], Pyr3 block state[5]), TRPC4 (162, Gq state[4], ML204 block
// state[5]) and TRPM7 (211, closed by intracellular Mg2+, NS8593 block state[5])
// (hh_biological.cpp trpc3/trpc4/trpm7): Na+ and Ca2+ currents from partial permeabilities; the
// Ca2+ change uses the charge-per-mole factor without the valence.
static double chan_trp_naca(Modifier* m, Segment* s, float dt, uint8_t t) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float po;
    float gamma, pca;
    if (t == 211) {
        float mg = s->intra_ions[4];
        float r = mg / 0.7f;
        po = 1.0f / (1.0f + r * r);
        double ns = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        po /= (1.0f + ns / 1.6f);
        gamma = 40.0f;
        pca = 0.34f;
    } else {
        double a = m->state[4];
        if (a < 0.0f) a = 0.0f;
        if (a > 1.0f) a = 1.0f;
        po = a;
        double b = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        po /= (1.0f + b / (t == 210 ? 0.7f : 2.0f));
        gamma = t == 210 ? 60.0f : 40.0f;
        pca = t == 210 ? 1.62f : 4.0f;
    }
    double pna, pk, pc;
    p_partial(gamma, 1.0f, 0.0f, pca, &pna, &pk, &pc);
    double N = m->density * area;
    double fna = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double fca = ghk_di(V, s->intra_ions[2], s->extra_ions[2]);
    double ina = pna * fna * N * po * kAtoPa;
    double ica = pc * fca * N * po * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0;
    s->intra_ions[0] -= (ina * dt) / cpv;
    s->extra_ions[0] += (ina * dt) / (cpv * kEvr);
    s->intra_ions[2] -= (ica * dt) / cpv;
    s->extra_ions[2] += (ica * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    floor_ions(s->intra_ions);
    return -(ina + ica);
}

// TRPV3 (175, hh_biological.cpp trpv3_current): ligand state[4] (EC50 50) drives the gate with
// tau 20 ms; float Na+ and Ca2+ currents.
static double chan_trpv3(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double lig = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    float r = lig / (lig + 50.0f);
    m->state[0] += (r - m->state[0]) * (dt / 20.0f);
    if (m->state[0] < 0) m->state[0] = 0;
    if (m->state[0] > 1) m->state[0] = 1;
    double op = m->state[0];
    double pna, pk, pca;
    p_partial(185.0f, 0.5f, 0.0f, 0.5f, &pna, &pk, &pca);
    double N = m->density * area;
    double fca = ghk_di(V, s->intra_ions[2], s->extra_ions[2]);
    double fna = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    float ica = (float)(pca * fca * N * op * kAtoPa);
    float ina = (float)(pna * fna * N * op * kAtoPa);
    double cpm = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    s->intra_ions[0] -= (ina * dt) / cpm;
    s->extra_ions[2] += (ica * dt) / (2.0f * cpm * kEvr);
    s->extra_ions[0] += (ina * dt) / (cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -(ica + ina);
}

// ENaC (182, hh_biological.cpp enac_current): Po 0.5, amiloride block state[5] (IC50 100),
// 4.5 pS; float current.
static double chan_enac(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float op = 0.5f;
    double am = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    op /= (1.0f + am / 100.0f);
    double P = p_single(4.5f, 1.0f, kCbarNa);
    double flux = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    float I = (float)(P * flux * (m->density * area) * op * kAtoPa);
    double cpv = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[0] -= (I * dt) / cpv;
    s->extra_ions[0] += (I * dt) / (cpv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -I;
}

// Ligand-gated TRP channels with float Na+/Ca2+ currents (hh_biological.cpp trpv1_channel/
// trpm8/trpa1/trpc67): the ligand state[4] sets the gate target (EC50 ec), a Ca-driven
// desensitisation in state[1] (TRPV1 172, TRPM8 173, TRPA1 174) and, for TRPA1, Ca2+
// co-activation; TRPC6/7 (179) has DAG gating only.
struct TrpLig { float ec, tau, tdes, gamma, fna, fca; bool des, ca_act; };
static constexpr TrpLig kTrpv1 = {0.7f, 10.0f, 150.0f, 70.0f, 0.09f, 0.91f, true, false};
static constexpr TrpLig kTrpm8 = {30.0f, 20.0f, 200.0f, 63.0f, 0.5f, 0.5f, true, false};
static constexpr TrpLig kTrpa1 = {30.0f, 30.0f, 300.0f, 90.0f, 0.54f, 0.46f, true, true};
static constexpr TrpLig kTrpc67 = {10.0f, 50.0f, 0.0f, 35.0f, 0.17f, 0.83f, false, false};
static double chan_trp_lig(Modifier* m, Segment* s, float dt, const TrpLig& g) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double lig = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    float r = lig / (lig + g.ec);
    double op;
    if (g.des) {
        float ca = s->intra_ions[2] * kMmToUm;
        float ca_act = ca / (ca + 6.0f);
        float d_inf = ca / (ca + 0.5f);
        m->state[0] += (r - m->state[0]) * (dt / g.tau);
        m->state[1] += (d_inf - m->state[1]) * (dt / g.tdes);
        if (m->state[0] < 0) m->state[0] = 0;
        if (m->state[0] > 1) m->state[0] = 1;
        if (m->state[1] < 0) m->state[1] = 0;
        if (m->state[1] > 1) m->state[1] = 1;
        double act = g.ca_act ? (double)fminf(1.0f, m->state[0] + ca_act) : m->state[0];
        op = act * (1.0f - m->state[1]);
    } else {
        m->state[0] += (r - m->state[0]) * (dt / g.tau);
        if (m->state[0] < 0) m->state[0] = 0;
        if (m->state[0] > 1) m->state[0] = 1;
        op = m->state[0];
    }
    double pna, pk, pca;
    p_partial(g.gamma, g.fna, 0.0f, g.fca, &pna, &pk, &pca);
    double N = m->density * area;
    double fca = ghk_di(V, s->intra_ions[2], s->extra_ions[2]);
    double fna = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    float ica = (float)(pca * fca * N * op * kAtoPa);
    float ina = (float)(pna * fna * N * op * kAtoPa);
    double cpm = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    s->intra_ions[0] -= (ina * dt) / cpm;
    s->extra_ions[2] += (ica * dt) / (2.0f * cpm * kEvr);
    s->extra_ions[0] += (ina * dt) / (cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -(ica + ina);
}

// TRPC5 (163, Gq state[4] x intracellular Ca2+ co-activation, HC-070 block state[5]) and TRPM4
// (164, Ca2+-activated Na+/K+ channel, 9-phenanthrol block state[5]) (hh_biological.cpp
// trpc5/trpm4).
static double chan_trpc5_m4(Modifier* m, Segment* s, float dt, bool m4) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float cu = s->intra_ions[2] * 1000.0f;
    float po;
    if (m4) {
        po = cu / (cu + 0.4f);
        double ph = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        po /= (1.0f + ph / 18.0f);
    } else {
        double gq = m->state[4];
        if (gq < 0.0f) gq = 0.0f;
        if (gq > 1.0f) gq = 1.0f;
        float ca_act = cu / (cu + 0.5f);
        po = gq * ca_act;
        double hc = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
        po /= (1.0f + hc / 9.3f);
    }
    double pna, pk, pca;
    if (m4) p_partial(25.0f, 1.0f, 1.0f, 0.0f, &pna, &pk, &pca);
    else p_partial(50.0f, 1.0f, 0.0f, 5.0f, &pna, &pk, &pca);
    double N = m->density * area;
    double fna = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double ina = pna * fna * N * po * kAtoPa;
    double cpv = kFaraday * volume(s) / 1000.0;
    const int k2 = m4 ? 1 : 2;
    double f2 = m4 ? ghk_mono(V, s->intra_ions[1], s->extra_ions[1]) : ghk_di(V, s->intra_ions[2], s->extra_ions[2]);
    double i2 = (m4 ? pk : pca) * f2 * N * po * kAtoPa;
    s->intra_ions[0] -= (ina * dt) / cpv;
    s->extra_ions[0] += (ina * dt) / (cpv * kEvr);
    s->intra_ions[k2] -= (i2 * dt) / cpv;
    s->extra_ions[k2] += (i2 * dt) / (cpv * kEvr);
    floor_ions(s->extra_ions);
    floor_ions(s->intra_ions);
    return -(ina + i2);
}

// CALHM1/3 (180, hh_biological.cpp calhm13_current): depolarisation-activated gate (V1/2 30.5,
// tau 20 ms), float Na+/K+/Ca2+ currents.
static double chan_calhm(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float mi = 1.0f / (1.0f + exp(-(V - 30.5f) / 8.0f));
    m->state[0] += (mi - m->state[0]) * (dt / 20.0f);
    if (m->state[0] < 0) m->state[0] = 0;
    if (m->state[0] > 1) m->state[0] = 1;
    double op = m->state[0];
    double pna, pk, pca;
    p_partial(25.0f, 1.0f, 1.17f, 10.0f, &pna, &pk, &pca);
    double N = m->density * area;
    double fna = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double fk = ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    double fca = ghk_di(V, s->intra_ions[2], s->extra_ions[2]);
    float ina = (float)(pna * fna * N * op * kAtoPa);
    float ik = (float)(pk * fk * N * op * kAtoPa);
    float ica = (float)(pca * fca * N * op * kAtoPa);
    double cpv = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[0] -= (ina * dt) / cpv;
    s->intra_ions[1] -= (ik * dt) / cpv;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpv);
    s->extra_ions[0] += (ina * dt) / (cpv * kEvr);
    s->extra_ions[1] += (ik * dt) / (cpv * kEvr);
    s->extra_ions[2] += (ica * dt) / (2.0f * cpv * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -(ina + ik + ica);
}

// Kv3.4 (159, hh_biological.cpp kv34_current, modifier_helpers.h:564-576): m h hs with PKC
// relief of the fast inactivation, 4-AP/TEA blocks from states 5 and 6; float K+ in the flux.
static double chan_kv34(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float mi = 1.0f / (1.0f + exp(-(V - 19.0f) / 11.3f));
    float tm = 40.0f - 37.7f / (1.0f + exp(-V / 28.0f));
    m->state[0] += (mi - m->state[0]) * dt / tm;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float hi = 1.0f / (1.0f + exp((V + 15.0f) / 7.4f));
    float th = 12.0f + 165.6f * exp(-V / 11.2f);
    if (th > 2000.0f) th = 2000.0f;
    m->state[1] += (hi - m->state[1]) * dt / th;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    float pkc = first_state_of(s, T_PKC, 0, 0.0f);
    double he = m->state[1] + (1.0f - m->state[1]) * pkc * 0.9f;
    float hsi = 1.0f / (1.0f + exp((V + 30.0f) / 6.0f));
    m->state[2] += (hsi - m->state[2]) * dt / 2000.0f;
    if (m->state[2] < 0.0f) m->state[2] = 0.0f;
    if (m->state[2] > 1.0f) m->state[2] = 1.0f;
    double op = m->state[0] * he * m->state[2];
    double ap4 = (m->state[5] > 0.0f) ? m->state[5] : 0.0f;
    double tea = (m->state[6] > 0.0f) ? m->state[6] : 0.0f;
    op /= (1.0f + ap4 / 500.0f) * (1.0f + tea / 1000.0f);
    float ki = s->intra_ions[1], ko = s->extra_ions[1];
    double P = p_single(20.0f, 1.0f, kCbarK);
    double flux = ghk_mono(V, ki, ko);
    return k_efflux(s, P * flux * (m->density * area) * op * kAtoPa, dt);
}

// Ca2+/Na+ leak (101, hh_biological.cpp ca_leak_current): always open, 50 pS split 1:2 Na:Ca,
// float concentrations in the fluxes, the volume rounded to float, intracellular change only.
static double chan_ca_leak(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    float V = s->V_m;
    float ci = s->intra_ions[2], co = s->extra_ions[2];
    float ni = s->intra_ions[0], no = s->extra_ions[0];
    double pna, pk, pca;
    p_partial(50.0f, 1.0f, 0.0f, 2.0f, &pna, &pk, &pca);
    double N = m->density * area;
    double fna = ghk_mono(V, ni, no);
    double fca = ghk_di(V, ci, co);
    double ina = pna * fna * N * kAtoPa;
    double ica = pca * fca * N * kAtoPa;
    float vol = volume(s);
    double cpv = kFaraday * vol / 1000.0f;
    s->intra_ions[0] -= (ina * dt) / cpv;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpv);
    floor_ions(s->intra_ions);
    return -(ina + ica);
}

// Post-inhibitory-rebound T-type Ca2+ (96, hh_biological.cpp pir_ttype_current): m^3 h in ten
// substeps, serotonin and dopamine reduce it by up to 70%; float Ca in the flux, the
// intracellular Ca changes with a float charge-per-mole factor and is not floored.
static double chan_pir_ttype(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 7.2f;
    double mv = m->state[0], hv = m->state[1];
    float mi = 1.0f / (1.0f + exp(-(V - -48.4f) / 5.0f));
    float hi = 1.0f / (1.0f + exp((V - -75.6f) / 6.5f));
    float th = 15.0f + (150.0f - 15.0f) / (1.0f + exp((V - -65.0f) / 8.0f));
    float sub = dt / 10.0f;
    for (int k = 0; k < 10; k++) {
        mv += (mi - mv) * sub / 3.0f;
        hv += (hi - hv) * sub / th;
    }
    if (mv < 0.0f) mv = 0.0f;
    if (mv > 1.0f) mv = 1.0f;
    if (hv < 0.0f) hv = 0.0f;
    if (hv > 1.0f) hv = 1.0f;
    m->state[0] = mv;
    m->state[1] = hv;
    float ci = s->intra_ions[2], co = s->extra_ions[2];
    double P = p_single(gamma, 2.0f, kCbarCa);
    float ht = s->neurotransmitter[3], da = s->neurotransmitter[2];
    double mod = (1.0 - 0.7 / (1.0 + (ht + da) / 0.01f));
    double flux = ghk_di(V, ci, co);
    double ia = P * flux * (m->density * area) * mod * (mv * mv * mv * hv);
    double ica = ia * kAtoPa;
    float vol = volume(s);
    float cpm = kFaraday * vol / 1000.0;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    return -ica;
}

// L-type Ca2+ Cav1.2 (84, hh_biological.cpp:599-646, modifier_helpers.h:624-626): m^2 times a
// slow voltage-dependent (state[1]) and a Ca-dependent (state[5]) inactivation, PKA (state[1]
// of the first cAMP/PKA modifier) doubles the current at full activity.
static double chan_ca_l(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double vh = (m->state[3] != 0.0f) ? -fabs(m->state[3]) : -11.0f;
    double k = (m->state[4] > 0.0f) ? m->state[4] : 6.9f;
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 25.0f;
    float mi = 1.0f / (1.0f + exp(-(V - vh) / (k + 1e-6f)));
    m->state[0] += (mi - m->state[0]) * dt / 3.0f;
    float vi = 1.0f / (1.0f + exp((V + 37.3f) / 8.0f));
    m->state[1] += (vi - m->state[1]) * dt / 150.0f;
    float ca = s->intra_ions[2] * kMmToUm;
    float ci_ = 1.0f / (1.0f + ca / 1.0f);
    m->state[5] += (ci_ - m->state[5]) * dt / 35.0f;
    for (int q : {0, 1, 5}) {
        if (m->state[q] < 0.0f) m->state[q] = 0.0f;
        if (m->state[q] > 1.0f) m->state[q] = 1.0f;
    }
    float ci = s->intra_ions[2], co = s->extra_ions[2];
    double P = p_single(gamma, 2.0f, kCbarCa);
    double op = m->state[0] * m->state[0] * m->state[1] * m->state[5];
    float boost = 1.0f + 1.0f * first_state_of(s, T_CAMP_PKA, 1, 0.0f);
    double flux = ghk_di(V, ci, co);
    double ia = P * flux * (m->density * area) * op * boost;
    double ica = ia * kAtoPa;
    double cpm = kFaraday * volume(s) / 1000.0;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    s->extra_ions[2] -= (ica * dt) / (2.0f * cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -ica;
}

// Cav1.4 (144, hh_biological.cpp:648-682, modifier_helpers.h:283-313): m^2 without
// inactivation; CaBP4 (state[0] of the first CaBP4 modifier, as it stands when this one runs)
// shifts V1/2 by -10 mV.
static double chan_cav14(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 4.0f;
    float bound = first_state_of(s, 145, 0, 0.0f);
    float vh = -15.0f + -10.0f * bound;
    float mi = 1.0f / (1.0f + exp(-(V - vh) / (7.0f + 1e-6f)));
    m->state[0] += (mi - m->state[0]) * dt / 2.0f;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float ci = s->intra_ions[2], co = s->extra_ions[2];
    double P = p_single(gamma, 2.0f, kCbarCa);
    double op = m->state[0] * m->state[0];
    double flux = ghk_di(V, ci, co);
    double ia = P * flux * (m->density * area) * op;
    double ica = ia * kAtoPa;
    double cpm = kFaraday * volume(s) / 1000.0;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    s->extra_ions[2] -= (ica * dt) / (2.0f * cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -ica;
}

// CaBP4 (145, hh_biological.cpp:684-693): no current; state[0] binds towards 1 with tau 1 ms.
static double chan_cabp4(Modifier* m, Segment*, float dt) {
    m->state[0] += (1.0f - m->state[0]) * dt / 1.0f;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    return 0.0f;
}

// CaCC / ANO2 (146, hh_biological.cpp:695-722, modifier_helpers.h:318-327): Ca-activated Cl-
// channel, Hill 2 with EC50 4 uM (pow(x, 2) of a float is exact in double), GHK anion current;
// only the intracellular Cl changes.
static double chan_cacc(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double gamma = (m->state[2] > 0.0f) ? m->state[2] : 2.0f;
    float cu = s->intra_ions[2] * kMmToUm;
    if (cu < 0.0f) cu = 0.0f;
    float cn = pow(cu, 2.0f);
    float en = pow(4.0f, 2.0f);
    float mi = cn / (cn + en);
    m->state[0] += (mi - m->state[0]) * dt / 2.0f;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    float ci = s->intra_ions[3], co = s->extra_ions[3];
    double P = p_single(gamma, 1.0f, kCbarCl);
    double flux = ghk_an(V, ci, co);
    double ia = P * flux * (m->density * area) * m->state[0];
    double icl = ia * kAtoPa;
    double cpm = kFaraday * volume(s) / 1000.0;
    s->intra_ions[3] += (icl * dt) / cpm;
    floor_ions(s->intra_ions);
    return -icl;
}

// CAN / TRPM4-5 (152, hh_biological.cpp:3643-3682): Ca-activated (Hill 2, Kd 0.5 uM) Na+/K+
// channel with slow Ca-dependent desensitisation in state[1]; float currents.
static double chan_can(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    float ca = s->intra_ions[2];
    float kd = 5e-4f;
    float po = (ca * ca) / (ca * ca + kd * kd);
    float di = ca / (ca + 0.01f);
    m->state[1] += (di - m->state[1]) * dt / 2000.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    po *= (1.0f - m->state[1]);
    double pna, pk, pca;
    p_partial(30.0f, 1.0f, 1.0f, 0.0f, &pna, &pk, &pca);
    double N = m->density * area;
    double fna = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    double fk = ghk_mono(V, s->intra_ions[1], s->extra_ions[1]);
    float ina = (float)(pna * fna * N * po * kAtoPa);
    float ik = (float)(pk * fk * N * po * kAtoPa);
    double cpm = kFaraday * volume(s) / 1000.0;
    s->intra_ions[0] -= (ina * dt) / cpm;
    s->intra_ions[1] -= (ik * dt) / cpm;
    s->extra_ions[0] += (ina * dt) / (cpm * kEvr);
    s->extra_ions[1] += (ik * dt) / (cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -(ina - ik);
}

// TMC1 (190, hh_biological.cpp tmc1_current): hair-cell transduction channel, open fraction
// relaxes to 0.1 + 0.9 stim/(stim + 0.3) (stim = state[4]), fast (state[1]) and slow (state[2])
// Ca-dependent adaptation; Ca2+ and Na+ currents from partial permeabilities, as floats.
static double chan_tmc1(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    float area = eff_area(m, s);
    double stim = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    float pi = 0.1f + 0.9f * (stim / (stim + 0.3f));
    m->state[0] += (pi - m->state[0]) * (dt / 1.0f);
    float ca = s->intra_ions[2] * kMmToUm;
    float ad = ca / (ca + 5.0f);
    m->state[1] += (ad - m->state[1]) * (dt / 2.0f);
    m->state[2] += (ad - m->state[2]) * (dt / 45.0f);
    for (int q = 0; q < 3; q++) {
        if (m->state[q] < 0) m->state[q] = 0;
        if (m->state[q] > 1) m->state[q] = 1;
    }
    double open = m->state[0] * (1.0f - m->state[1]) * (1.0f - m->state[2]);
    double pna, pk, pca;
    p_partial(80.0f, 0.21f, 0.0f, 0.79f, &pna, &pk, &pca);
    double N = m->density * area;
    double fca = ghk_di(V, s->intra_ions[2], s->extra_ions[2]);
    double fna = ghk_mono(V, s->intra_ions[0], s->extra_ions[0]);
    float ica = (float)(pca * fca * N * open * kAtoPa);
    float ina = (float)(pna * fna * N * open * kAtoPa);
    double cpm = kFaraday * volume(s) / 1000.0f;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    s->intra_ions[0] -= (ina * dt) / cpm;
    s->extra_ions[2] += (ica * dt) / (2.0f * cpm * kEvr);
    s->extra_ions[0] += (ina * dt) / (cpm * kEvr);
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return -(ica + ina);
}

// eCB synthesis (86, hh_biological.cpp:1746-1764): 2-AG from Ca (Hill 2, Kd 0.5 uM) scaled by
// DAG (state[1] of the segment's first PLC/IP3/DAG modifier, 136; modifier_helpers.h:1496-1500),
// degraded at state[2]; dt scaled by the density; no current. Step 3 deposits it into clefts.
static float seg_dag(const Segment* s) {
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == 136) return s->modifiers[j].state[1];
    return 0.0f;
}
static double mod_ecb(Modifier* m, Segment* s, float dt) {
    float ca = s->intra_ions[2] * kMmToUm;
    double synth_rate = (m->state[1] > 0.0f) ? m->state[1] : 0.01f;
    double degrad = (m->state[2] > 0.0f) ? m->state[2] : 0.005f;
    float dag = seg_dag(s);
    dt *= m->density;
    float can = ca * ca;
    float synth = synth_rate * can / (can + 0.25f + 1e-9f) * (1.0f + dag);
    m->state[0] += (synth - m->state[0] * degrad) * dt;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 10.0f) m->state[0] = 10.0f;
    return 0.0;
}

// Metabotropic (3, hh_biological.cpp metabotropic_current): a G protein driven by the sum of
// all 15 transmitters (state[0]), an effector state[1] = coupling * state[0], and a GIRK-like
// K+ GHK current (35 pS) through the effector.
static double mod_metabotropic(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    float total = 0.0f;
    for (int n = 0; n < 15; n++) total += s->neurotransmitter[n];
    double coupling = (m->state[2] > 0.0f) ? m->state[2] : 1.0f;
    double tau = (m->state[3] > 0.0f) ? m->state[3] : 100.0f;
    float act = total / (total + 1.0f);
    m->state[0] += (act - m->state[0]) * dt / tau;
    m->state[1] = coupling * m->state[0];
    double cpm = kFaraday * volume(s) / 1000.0;
    double P = p_single(35.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s->V_m, s->intra_ions[1], s->extra_ions[1]);
    double N = m->density * area;
    double ik = P * flux * N * m->state[1] * kAtoPa;
    s->intra_ions[1] -= (ik * dt) / cpm;
    floor_ions(s->intra_ions);
    return -ik;
}
// Synaptic release site (5, hh_biological.cpp synaptic_release_current): a Ca2+ sensor (Hill 2,
// K_d 0.5 uM) feeding facilitation (state[1], tau state[4] or 100 ms) and depression (state[2],
// tau state[5] or 500 ms); an SK-like K+ GHK current (10 pS) opens with facilitation / (1 +
// depression).
static double mod_release_site(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    float ca = s->intra_ions[2];
    float sensor = (ca * ca) / (ca * ca + 0.0005f * 0.0005f);
    double fac_tau = (m->state[4] > 0.0f) ? m->state[4] : 100.0f;
    double dep_tau = (m->state[5] > 0.0f) ? m->state[5] : 500.0f;
    m->state[0] += (sensor - m->state[0]) * dt / 1.0f;
    m->state[1] += (m->state[0] - m->state[1]) * dt / fac_tau;
    m->state[2] += (m->state[0] - m->state[2]) * dt / dep_tau;
    double open = m->state[1] / (1.0f + m->state[2]);
    double cpm = kFaraday * volume(s) / 1000.0;
    double P = p_single(10.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s->V_m, s->intra_ions[1], s->extra_ions[1]);
    double N = m->density * area;
    double ik = P * flux * N * open * kAtoPa;
    s->intra_ions[1] -= (ik * dt) / cpm;
    floor_ions(s->intra_ions);
    return -ik;
}
// MRGPR (188, modifier_helpers.h mrgpr_signal): a Gq receptor of the ligand level in state[4]
// (EC50 28 uM): activation state[0] (tau 70 ms up, 1500 ms down), desensitisation state[1]
// (tau 70 s); no current. PLC reads its state[0] with the other Gq receptors.
static double mod_mrgpr(Modifier* m, Segment*, float dt) {
    float lig = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    float occ = lig / (lig + 28.0f);
    float act = m->state[0];
    float tau = (occ > act) ? 70.0f : 1500.0f;
    act += (occ - act) * dt / tau;
    if (act < 0.0f) act = 0.0f;
    if (act > 1.0f) act = 1.0f;
    m->state[0] = act;
    float dz = m->state[1];
    dz += (occ - dz) * dt / 70000.0f;
    if (dz < 0.0f) dz = 0.0f;
    if (dz > 1.0f) dz = 1.0f;
    m->state[1] = dz;
    return 0.0f;
}
// Somatic ATP synthesis (77, hh_biological.cpp nb_soma_atp_current): NADH (state[0]) relaxing
// to 0.1, ATP made at state[1] (0.01 if not positive) x NADH x (1 - ATP/5) x density up to 5,
// warming the segment by a little (capped at 42 C); no current.
static double mod_soma_atp(Modifier* m, Segment* s, float dt) {
    double nadh = m->state[0];
    float atp = s->ATP_level;
    float inh = 1.0f - atp / 5.0f;
    if (inh < 0.0f) inh = 0.0f;
    nadh += (0.001f - nadh * 0.01f) * dt;
    if (nadh < 0.0f) nadh = 0.0f;
    if (nadh > 1.0f) nadh = 1.0f;
    double rate = (m->state[1] > 0.0f) ? m->state[1] : 0.01f;
    s->ATP_level += rate * nadh * inh * m->density * dt;
    if (s->ATP_level > 5.0f) s->ATP_level = 5.0f;
    s->temperature += 0.01f * nadh * inh * 0.5f * dt * 0.001f;
    if (s->temperature > 42.0f) s->temperature = 42.0f;
    m->state[0] = nadh;
    return 0.0f;
}
// Somatic thermoregulation (78, hh_biological.cpp nb_soma_thermal_current): temperature relaxes
// to 37 C with tau state[1] (5000 ms if not positive), clamped to [30, 42]; heat stress
// (state[0]) builds above 39 C and decays below; no current.
static double mod_soma_thermal(Modifier* m, Segment* s, float dt) {
    double stress = m->state[0];
    float T = s->temperature;
    double tau = (m->state[1] > 0.0f) ? m->state[1] : 5000.0f;
    T += (37.0f - T) * dt / tau;
    if (T < 30.0f) T = 30.0f;
    if (T > 42.0f) T = 42.0f;
    if (T > 39.0f) stress += (T - 39.0f) * dt * 0.0001f;
    else stress -= stress * dt * 0.001f;
    if (stress < 0.0f) stress = 0.0f;
    if (stress > 1.0f) stress = 1.0f;
    s->temperature = T;
    m->state[0] = stress;
    return 0.0f;
}
// CREB transcription (79, hh_biological.cpp nb_creb_transcription_current): Ca2+ (Hill 2, K_d
// state[3] or 0.5) drives pCREB (state[0]) -> late genes (state[1]) -> structural plasticity
// (state[2]); the segment's transcription signal is their weighted sum; no current.
static double mod_creb(Modifier* m, Segment* s, float dt) {
    float ca = s->intra_ions[2];
    double p = m->state[0], late = m->state[1], st = m->state[2];
    double kd = (m->state[3] > 0.0f) ? m->state[3] : 0.5f;
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
    m->state[0] = p;
    m->state[1] = late;
    m->state[2] = st;
    s->transcription_signal = 0.5f * p + 0.3f * late + 0.2f * st;
    return 0.0f;
}

// The last instance's state[k] of a type on the segment, as the reference's loops without a
// break read it (0 if none).
static float last_state_of(const Segment* s, uint8_t type, int k) {
    float v = 0.0f;
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == type) v = s->modifiers[j].state[k];
    return v;
}
// TRPV1 as a transmitter receptor (16, hh_biological.cpp trpv1_current): heat (Boltzmann around
// 43 C, slope 0.5 C) or capsaicin (transmitter 14, EC50 0.5) opens it with slow activation
// (state[0], 200 ms) and desensitisation (state[1], 5 s); a 90 pS Na+/K+/Ca2+ GHK current.
static double mod_trpv1_receptor(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    float V = s->V_m;
    double cpm = kFaraday * volume(s) / 1000.0;
    float ft = 1.0f / (1.0f + exp((43.0f - s->temperature) / 0.5f));
    float cap = s->neurotransmitter[14];
    if (cap < 0.0f) cap = 0.0f;
    float fc = cap / (cap + 0.5f);
    double r = m->state[0];
    float pinf = ft + fc - ft * fc;
    r += (pinf - r) * dt / 200.0f;
    if (r < 0.0f) r = 0.0f;
    if (r > 1.0f) r = 1.0f;
    m->state[0] = r;
    double d = m->state[1];
    d += (pinf - d) * dt / 5000.0f;
    if (d < 0.0f) d = 0.0f;
    if (d > 1.0f) d = 1.0f;
    m->state[1] = d;
    float open = r * (1.0f - d);
    double pna, pk, pca;
    p_partial(90.0f, 0.125f, 0.125f, 0.75f, &pna, &pk, &pca);
    double N = m->density * area;
    double ina = pna * ghk_mono(V, s->intra_ions[0], s->extra_ions[0]) * N * open * kAtoPa;
    float ik = (float)(pk * ghk_mono(V, s->intra_ions[1], s->extra_ions[1]) * N * open * kAtoPa);
    double ica = pca * ghk_di(V, s->intra_ions[2], s->extra_ions[2]) * N * open * kAtoPa;
    s->intra_ions[0] -= (ina * dt) / cpm;
    s->intra_ions[1] -= (ik * dt) / cpm;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    floor_ions(s->intra_ions);
    return -(ina + ik + ica);
}
// Rod phototransduction (192, hh_biological.cpp rod_phototransduction_current): light (state[4])
// -> R* (state[0]) -> transducin (state[1]) -> PDE (state[2]) hydrolysing cGMP (state[3]); CNG
// channels (Hill 3 in cGMP, K 0.3; 25 pS Na+/K+/Ca2+) carry the dark current.
static double mod_rod(Modifier* m, Segment* s, float dt) {
    float V = s->V_m;
    double light = (m->state[4] > 0.0f) ? m->state[4] : 0.0f;
    m->state[0] += (light - m->state[0]) * dt / 1500.0f;
    m->state[1] += (m->state[0] - m->state[1]) * dt / 10.0f;
    m->state[2] += (m->state[1] - m->state[2]) * dt / 200.0f;
    m->state[3] += (1.0f - m->state[3]) * dt / 500.0f - m->state[2] * m->state[3] * dt / 50.0f;
    for (int k = 0; k < 4; k++) {
        if (m->state[k] < 0) m->state[k] = 0;
        if (m->state[k] > 1) m->state[k] = 1;
    }
    double cg = m->state[3];
    float cg3 = cg * cg * cg;
    float kh = 0.3f, kh3 = kh * kh * kh;
    float po = cg3 / (kh3 + cg3);
    float area = eff_area(m, s);
    double cpm = kFaraday * volume(s) / 1000.0;
    double pna, pk, pca;
    p_partial(25.0f, 0.5f, 0.3f, 0.2f, &pna, &pk, &pca);
    double N = m->density * area * po;
    double ina = pna * ghk_mono(V, s->intra_ions[0], s->extra_ions[0]) * N * kAtoPa;
    float ik = (float)(pk * ghk_mono(V, s->intra_ions[1], s->extra_ions[1]) * N * kAtoPa);
    double ica = pca * ghk_di(V, s->intra_ions[2], s->extra_ions[2]) * N * kAtoPa;
    s->intra_ions[0] -= (ina * dt) / cpm;
    s->intra_ions[1] -= (ik * dt) / cpm;
    s->intra_ions[2] -= (ica * dt) / (2.0f * cpm);
    floor_ions(s->intra_ions);
    return -(ina + ik + ica);
}
static inline void clamp01_4(double* st) {
    for (int k = 0; k < 4; k++) {
        if (st[k] < 0.0f) st[k] = 0.0f;
        if (st[k] > 1.0f) st[k] = 1.0f;
    }
}
// CaMKII (82, hh_biological.cpp:545-570): activation by Ca2+ (uM, Hill 4, K_d state[2] or 1
// uM) with time constant state[3] (2000 ms), autophosphorylation state[1] at rate state[5]
// (0.0005) decaying at 1e-4 per ms; time scaled by the density; no current. Connexins, Kv4.2 and
// vesicle recovery read the last CaMKII's state[0] (and state[1]).
static double mod_camkii(Modifier* m, Segment* s, float dt) {
    float ca = s->intra_ions[2];
    float cu = ca * kMmToUm;
    double kd = (m->state[2] > 0.0f) ? m->state[2] : 1.0f;
    double tau = (m->state[3] > 0.0f) ? m->state[3] : 2000.0f;
    double ar = (m->state[5] > 0.0f) ? m->state[5] : 0.0005f;
    dt *= m->density;
    float c4 = cu * cu * cu * cu;
    float inf = c4 / (c4 + kd * kd * kd * kd + 1e-9f);
    m->state[0] += (inf - m->state[0]) * dt / tau;
    m->state[1] += (m->state[0] * ar - m->state[1] * 0.0001f) * dt;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    return 0.0f;
}
// Calcineurin (83, hh_biological.cpp:575-591): Ca2+ (uM, Hill 2, K_d state[1] or 0.05 uM), time
// constant state[2] (300 ms), time scaled by the density; no current. Vesicle recovery reads the
// last calcineurin's state[0].
static double mod_calcineurin(Modifier* m, Segment* s, float dt) {
    float ca = s->intra_ions[2];
    float cu = ca * kMmToUm;
    double kd = (m->state[1] > 0.0f) ? m->state[1] : 0.05f;
    double tau = (m->state[2] > 0.0f) ? m->state[2] : 300.0f;
    dt *= m->density;
    float c2 = cu * cu;
    float inf = c2 / (c2 + kd * kd + 1e-9f);
    m->state[0] += (inf - m->state[0]) * dt / tau;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    return 0.0f;
}
// PP1 / DARPP-32 (113, hh_biological.cpp:3222-3253): PKA activity (state[1] of the segment's
// first cAMP/PKA modifier 135, modifier_helpers.h:1486-1490) phosphorylates DARPP-32 (state[0])
// and inhibitor-1 (state[1]) with time constant state[4] (5000 ms), time scaled by the density;
// PP1 activity state[3] = basal (state[5] or 0.5) / (1 + 10 p-DARPP + 5 p-I1 + 1e-9); state[6]
// holds the PKA activity; no current. The plasticity step reads the last PP1's state[3].
static double mod_pp1(Modifier* m, Segment* s, float dt) {
    double tau = (m->state[4] > 0.0f) ? m->state[4] : 5000.0f;
    double basal = (m->state[5] > 0.0f) ? m->state[5] : 0.5f;
    dt *= m->density;
    float pka = first_state_of(s, T_CAMP_PKA, 1, 0.0f);
    float pd = pka, pi = pka;
    m->state[0] += (pd - m->state[0]) * dt / tau;
    m->state[1] += (pi - m->state[1]) * dt / tau;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    if (m->state[1] > 1.0f) m->state[1] = 1.0f;
    double inh = 10.0f * m->state[0] + 5.0f * m->state[1];
    m->state[3] = basal / (1.0f + inh + 1e-9f);
    m->state[6] = pka;
    return 0.0f;
}
// PKC (137, hh_biological.cpp:3365-3367 -> modifier_helpers.h:1584-1593): state[0] = DAG / (DAG
// + 0.1) * Ca / (Ca + 0.0005 mM), DAG the state[1] of the segment's first PLC/IP3/DAG modifier
// and Ca the intracellular Ca2+ floored at 0; no current. Channels read the first PKC's state[0].
static float seg_dag(const Segment* s);
static double mod_pkc(Modifier* m, Segment* s, float dt) {
    (void)dt;
    float dag = seg_dag(s);
    float ca = s->intra_ions[2];
    if (ca < 0.0f) ca = 0.0f;
    float df = dag / (dag + 0.1f);
    float cf = ca / (ca + 0.0005f);
    m->state[0] = df * cf;
    return 0.0f;
}
// MAPK/ERK (87, hh_biological.cpp mapk_erk_current): Ca2+ (uM, Hill 3, K_d state[4] or 0.5) ->
// Ras (state[2]) -> MEK (state[1]) -> ERK (state[0], squared MEK) -> nuclear ERK (state[3]);
// time scaled by the density; nuclear ERK raises every CREB's pCREB (state[0]) on the segment.
static double mod_mapk(Modifier* m, Segment* s, float dt) {
    float ca = s->intra_ions[2] * 1e3f;
    double kd = (m->state[4] > 0.0f) ? m->state[4] : 0.5f;
    double tau_erk = (m->state[5] > 0.0f) ? m->state[5] : 5000.0f;
    double tau_nuc = (m->state[6] > 0.0f) ? m->state[6] : 30000.0f;
    dt *= m->density;
    float ca3 = ca * ca * ca;
    float kd3 = kd * kd * kd;
    float ras = ca3 / (ca3 + kd3 + 1e-9f);
    m->state[2] += (ras - m->state[2]) * dt / 500.0f;
    m->state[1] += (m->state[2] - m->state[1]) * dt / 1000.0f;
    double erk = m->state[1] * m->state[1];
    m->state[0] += (erk - m->state[0]) * dt / tau_erk;
    m->state[3] += (m->state[0] * 0.8f - m->state[3]) * dt / tau_nuc;
    clamp01_4(m->state);
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == 79) {
            s->modifiers[j].state[0] += m->state[3] * 0.01f * dt / 1000.0f;
            if (s->modifiers[j].state[0] > 1.0f) s->modifiers[j].state[0] = 1.0f;
        }
    return 0.0f;
}
// IKK/NF-kB (88, hh_biological.cpp ikk_nfkb_current): Ca2+ -> IKK (state[1]) -> IkB (state[2],
// floor 0.1) -> free NF-kB (state[0]) -> nuclear NF-kB (state[3]), which adds to the segment's
// transcription signal (capped at 1); time scaled by the density.
static double mod_nfkb(Modifier* m, Segment* s, float dt) {
    float ca = s->intra_ions[2] * 1e3f;
    double kd = (m->state[4] > 0.0f) ? m->state[4] : 0.3f;
    double tau_nfkb = (m->state[5] > 0.0f) ? m->state[5] : 10000.0f;
    double tau_ikb = (m->state[6] > 0.0f) ? m->state[6] : 5000.0f;
    dt *= m->density;
    float ikk = ca / (ca + kd + 1e-9f);
    m->state[1] += (ikk - m->state[1]) * dt / 2000.0f;
    m->state[2] += (1.0f - m->state[1] * 0.9f - m->state[2]) * dt / tau_ikb;
    if (m->state[2] < 0.1f) m->state[2] = 0.1f;
    double free_ = 1.0f - m->state[2];
    m->state[0] += (free_ - m->state[0]) * dt / tau_nfkb;
    m->state[3] += (m->state[0] * 0.7f - m->state[3]) * dt / (tau_nfkb * 2.0f);
    clamp01_4(m->state);
    s->transcription_signal += m->state[3] * 0.1f;
    if (s->transcription_signal > 1.0f) s->transcription_signal = 1.0f;
    return 0.0f;
}
// mTORC1 (89, hh_biological.cpp mtorc1_current): amino acids (state[3]) -> Rheb (state[1]) and
// AKT (state[2]); mTORC1 (state[0]) follows their product scaled by ATP/5; time scaled by the
// density.
static double mod_mtorc1(Modifier* m, Segment* s, float dt) {
    double tau_mtor = (m->state[5] > 0.0f) ? m->state[5] : 15000.0f;
    double tau_akt = (m->state[6] > 0.0f) ? m->state[6] : 8000.0f;
    double kd = (m->state[4] > 0.0f) ? m->state[4] : 0.3f;
    dt *= m->density;
    double aa = m->state[3];
    float rheb = aa / (aa + kd + 1e-9f);
    m->state[1] += (rheb - m->state[1]) * dt / 5000.0f;
    float akt = aa * 0.8f;
    m->state[2] += (akt - m->state[2]) * dt / tau_akt;
    double inf = m->state[1] * m->state[2];
    float an = s->ATP_level / 5.0f;
    if (an > 1.0f) an = 1.0f;
    if (an < 0.0f) an = 0.0f;
    inf *= an;
    m->state[0] += (inf - m->state[0]) * dt / tau_mtor;
    if (m->state[0] < 0.0f) m->state[0] = 0.0f;
    if (m->state[0] > 1.0f) m->state[0] = 1.0f;
    if (m->state[1] < 0.0f) m->state[1] = 0.0f;
    if (m->state[2] < 0.0f) m->state[2] = 0.0f;
    return 0.0f;
}
// Protein synthesis (90, hh_biological.cpp protein_synth_current): the last mTORC1's state[0]
// drives S6K (state[1]) and 4E-BP (state[2]); translation (state[0]) consumes ATP.
static double mod_protein_synth(Modifier* m, Segment* s, float dt) {
    double tau_s6k = (m->state[4] > 0.0f) ? m->state[4] : 10000.0f;
    double tau_4ebp = (m->state[5] > 0.0f) ? m->state[5] : 8000.0f;
    double tau_tr = (m->state[6] > 0.0f) ? m->state[6] : 60000.0f;
    dt *= m->density;
    float mt = last_state_of(s, 89, 0);
    m->state[1] += (mt - m->state[1]) * dt / tau_s6k;
    m->state[2] += (mt - m->state[2]) * dt / tau_4ebp;
    double tr = m->state[1] * m->state[2];
    m->state[0] += (tr - m->state[0]) * dt / tau_tr;
    m->state[3] = m->state[0] * 0.9f;
    clamp01_4(m->state);
    s->ATP_level -= m->state[0] * 0.0001f * dt;
    if (s->ATP_level < 0.0f) s->ATP_level = 0.0f;
    return 0.0f;
}
// Cofilin (91, hh_biological.cpp cofilin_current): the last MAPK's ERK (state[0]) -> LIMK
// (state[1]) -> phospho-cofilin (state[2]); active cofilin (state[0]) severs F-actin (state[3]).
static double mod_cofilin(Modifier* m, Segment* s, float dt) {
    double tau_cof = (m->state[4] > 0.0f) ? m->state[4] : 5000.0f;
    double tau_limk = (m->state[5] > 0.0f) ? m->state[5] : 8000.0f;
    double tau_actin = (m->state[6] > 0.0f) ? m->state[6] : 30000.0f;
    dt *= m->density;
    float erk = last_state_of(s, 87, 0);
    m->state[1] += (erk * 0.8f - m->state[1]) * dt / tau_limk;
    double pc = m->state[1] * 0.9f;
    m->state[2] += (pc - m->state[2]) * dt / tau_cof;
    double ct = 1.0f - m->state[2];
    m->state[0] += (ct - m->state[0]) * dt / tau_cof;
    float poly = 0.3f;
    double sev = m->state[0] * 0.1f;
    m->state[3] += (poly - sev * m->state[3]) * dt / tau_actin;
    clamp01_4(m->state);
    return 0.0f;
}
// MMP-9 (92, hh_biological.cpp mmp9_current): nuclear ERK and NF-kB (the last MAPK's and IKK's
// state[3]) drive MMP-9 (state[0]) against TIMP (state[1]); extracellular matrix (state[2],
// floor 0.1) and integrin activation (state[3]) follow.
static double mod_mmp9(Modifier* m, Segment* s, float dt) {
    double tau_mmp = (m->state[4] > 0.0f) ? m->state[4] : 20000.0f;
    double tau_timp = (m->state[5] > 0.0f) ? m->state[5] : 15000.0f;
    double tau_ecm = (m->state[6] > 0.0f) ? m->state[6] : 100000.0f;
    dt *= m->density;
    float erk = last_state_of(s, 87, 3), nfkb = last_state_of(s, 88, 3);
    float tr = (erk + nfkb) * 0.5f;
    m->state[0] += (tr - m->state[0]) * dt / tau_mmp;
    double timp = 1.0f - m->state[0] * 0.5f;
    m->state[1] += (timp - m->state[1]) * dt / tau_timp;
    double deg = m->state[0] * (1.0f - m->state[1]) * 0.01f;
    float syn = 0.005f;
    m->state[2] += (syn - deg * m->state[2]) * dt / tau_ecm;
    if (m->state[2] < 0.1f) m->state[2] = 0.1f;
    if (m->state[2] > 1.0f) m->state[2] = 1.0f;
    double integ = 1.0f - m->state[2];
    m->state[3] += (integ - m->state[3]) * dt / 10000.0f;
    clamp01_4(m->state);
    return 0.0f;
}

// Sensors (hh_biological.cpp:40-52, 483-540, 2181-2440, 3122-3141). Adaptation: state[k] relaxes
// toward a float activation with a float time constant (skipped below 1 ms), clamped to [0, 1]
// (sensor_habituation_step and sensor_sensitization_step are the same update); the gain
// (1 + k_s S)(1 - k_h H) is evaluated in float.
static inline void sensor_relax(Modifier* m, int k, float act, float dt, float tau) {
    if (tau < 1.0f) return;
    m->state[k] += (act - m->state[k]) * dt / tau;
    if (m->state[k] < 0.0f) m->state[k] = 0.0f;
    if (m->state[k] > 1.0f) m->state[k] = 1.0f;
}
static inline float sensor_gain(float H, float S, float kh, float ks) { return (1.0f + ks * S) * (1.0f - kh * H); }
static inline void clamp_state01(Modifier* m, int k) {
    if (m->state[k] < 0.0f) m->state[k] = 0.0f;
    if (m->state[k] > 1.0f) m->state[k] = 1.0f;
}
// GHK Na+/K+/Ca2+ currents of a channel with partial permeabilities; fca == 0 drops the Ca2+ term.
struct SensorI { double na; float k; double ca; };
static SensorI sensor_ghk(Segment* s, float gamma, float fna, float fk, float fca, double N, float dt) {
    double cpm = kFaraday * volume(s) / 1000.0;
    double pna, pk, pca;
    p_partial(gamma, fna, fk, fca, &pna, &pk, &pca);
    double V = s->V_m;   // the segment's potential in double, as the sensors pass it
    SensorI r;
    r.na = pna * ghk_mono(V, s->intra_ions[0], s->extra_ions[0]) * N * kAtoPa;
    r.k = (float)(pk * ghk_mono(V, s->intra_ions[1], s->extra_ions[1]) * N * kAtoPa);
    r.ca = (fca != 0.0f) ? pca * ghk_di(V, s->intra_ions[2], s->extra_ions[2]) * N * kAtoPa : 0.0;
    s->intra_ions[0] -= (r.na * dt) / cpm;
    s->intra_ions[1] -= (r.k * dt) / cpm;
    if (fca != 0.0f) s->intra_ions[2] -= (r.ca * dt) / (2.0f * cpm);
    floor_ions(s->intra_ions);
    return r;
}
// Physical sensor (6): heat above 37 C and cold below 33 C, plus stretch of segments longer than
// 50 um, scaled by state[1] (default 1), drive state[0] (10 ms); habituation state[2] (tau
// state[6] or 2 s), sensitisation state[3] (30 s); a 50 pS Na+/K+/Ca2+ (0.4/0.3/0.3) current.
static double mod_physical_sensor(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    double sens = (m->state[1] > 0.0f) ? m->state[1] : 1.0f;
    float T = s->temperature;
    float hot = (T > 37.0f) ? (T - 37.0f) / 6.0f : 0.0f;
    float cold = (T < 33.0f) ? (33.0f - T) / 8.0f : 0.0f;
    float stretch = (s->length - 50.0f) / 50.0f;
    if (stretch < 0.0f) stretch = 0.0f;
    float stim = sens * ((hot + cold) + stretch * 0.5f);
    m->state[0] += (stim - m->state[0]) * dt / 10.0f;
    clamp_state01(m, 0);
    double tau_h = (m->state[6] > 0.0f) ? m->state[6] : 2000.0f;
    sensor_relax(m, 2, m->state[0], dt, tau_h);
    sensor_relax(m, 3, m->state[0], dt, 30000.0f);
    double kh = (m->state[4] > 0.0f) ? m->state[4] : 0.6f;
    double ks = (m->state[5] > 0.0f) ? m->state[5] : 0.2f;
    double g = sensor_gain(m->state[2], m->state[3], kh, ks);
    SensorI I = sensor_ghk(s, 50.0f, 0.4f, 0.3f, 0.3f, m->density * area * m->state[0] * g, dt);
    return -(I.na - I.k + I.ca);
}
// Hunger sensor (94): low ATP (normalised to 5 mM, sigmoid at 0.3) drives state[0] (500 ms);
// habituation state[1] (300 s), sensitisation state[2] (600 s); 50 pS Na+/K+ (0.6/0.3).
static double mod_hunger_sensor(Modifier* m, Segment* s, float dt) {
    float a = s->ATP_level / 5.0f;
    if (a > 1.0f) a = 1.0f;
    if (a < 0.0f) a = 0.0f;
    float h = 1.0f / (1.0f + exp(10.0f * (a - 0.3f)));
    m->state[0] += (h - m->state[0]) * dt / 500.0f;
    clamp_state01(m, 0);
    sensor_relax(m, 1, m->state[0], dt, 300000.0f);
    sensor_relax(m, 2, m->state[0], dt, 600000.0f);
    double kh = (m->state[3] > 0.0f) ? m->state[3] : 0.5f;
    double ks = (m->state[4] > 0.0f) ? m->state[4] : 0.4f;
    double g = sensor_gain(m->state[1], m->state[2], kh, ks);
    float area = eff_area(m, s);
    SensorI I = sensor_ghk(s, 50.0f, 0.6f, 0.3f, 0.0f, m->density * area * m->state[0] * g, dt);
    return -(I.na - I.k);
}
// Light sensor (95): intensity state[1]; habituation state[2] rises (2 s) above 0.1 and decays
// (5 s) below; sensitisation state[3] (60 s); adapted intensity -> state[0] (50 ms); the CNG
// channel (25 pS, Na+/K+/Ca2+ 0.5/0.3/0.2) is open in the dark (1 - state[0]).
static double mod_light_sensor(Modifier* m, Segment* s, float dt) {
    double I0 = m->state[1];
    if (I0 > 0.1f) sensor_relax(m, 2, 1.0f, dt, 2000.0f);
    else m->state[2] -= dt * m->state[2] / 5000.0f;
    clamp_state01(m, 2);
    sensor_relax(m, 3, I0, dt, 60000.0f);
    float kh = 0.7f;
    double ks = (m->state[4] > 0.0f) ? m->state[4] : 0.5f;
    double ad = I0 * (1.0f - kh * m->state[2]) * (1.0f + ks * m->state[3]);
    if (ad < 0.0f) ad = 0.0f;
    m->state[0] += (ad - m->state[0]) * dt / 50.0f;
    clamp_state01(m, 0);
    float area = eff_area(m, s);
    double open = 1.0f - m->state[0];
    SensorI I = sensor_ghk(s, 25.0f, 0.5f, 0.3f, 0.2f, m->density * area * open, dt);
    return -(I.na + I.k + I.ca);
}
// Olfactory receptor (108): ligand state[1] (steep sigmoid at 0.01) -> state[0] (50 ms);
// habituation state[2] (2 s), sensitisation state[3] (10 s); the Gs path closes up to 80% of a
// 35 pS K+ conductance: the current is the closed part.
static double mod_olfactory(Modifier* m, Segment* s, float dt) {
    double L = m->state[1];
    if (L < 0.0f) L = 0.0f;
    if (L > 1.0f) L = 1.0f;
    float act = 1.0f / (1.0f + exp(-2.0f * (L - 0.01f) * 100.0f));
    m->state[0] += (act - m->state[0]) * dt / 50.0f;
    clamp_state01(m, 0);
    sensor_relax(m, 2, m->state[0], dt, 2000.0f);
    sensor_relax(m, 3, m->state[0], dt, 10000.0f);
    float kh = 0.7f;
    double ks = (m->state[4] > 0.0f) ? m->state[4] : 0.3f;
    double g = sensor_gain(m->state[2], m->state[3], kh, ks);
    float area = eff_area(m, s);
    double cpm = kFaraday * volume(s) / 1000.0;
    double e = m->state[0] * g;
    double P = p_single(35.0f, 1.0f, kCbarK);
    double flux = ghk_mono(s->V_m, s->intra_ions[1], s->extra_ions[1]);
    double N = m->density * area;
    float basal = (float)(P * flux * N * kAtoPa);
    float modulated = (float)(P * (1.0f - e * 0.8f) * flux * N * kAtoPa);
    float eff = basal - modulated;
    s->intra_ions[1] -= (eff * dt) / cpm;
    floor_ions(s->intra_ions);
    return -eff;
}
// The reference converts doubles to uint32_t by a plain cast; its x86-64 build truncates to a
// 64-bit integer (cvttsd2si) and keeps the low 32 bits, so negative values wrap and values
// outside the int64 range (or NaN) give 0. This states that conversion.
static inline uint32_t u32_trunc(double x) {
    if (!(x > -9223372036854775808.0 && x < 9223372036854775808.0)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)x;
}
// Thermal noise (111): a uniform draw from xorshift32 seeded with the segment id, state[1] and
// the membrane potential, scaled to sigma state[0] (default 0.05 mV) times sqrt(3), returned as
// the equivalent current through the segment's capacitance (1 uF/cm^2, floor 1e-4).
static double mod_thermal_noise(Modifier* m, Segment* s, float dt) {
    double sigma = (m->state[0] > 0.0f) ? m->state[0] : 0.05f;
    uint32_t x = s->id * 7919u + u32_trunc(m->state[1] * 131.0f);
    x ^= u32_trunc(s->V_m * 1000.0f);
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    float r = (float)(x & 0x7FFFFFFFu) / (float)0x7FFFFFFFu;
    float noise = (r * 2.0f - 1.0f) * sigma * 1.732f;
    float cm = eff_area(m, s) * 0.01f;
    if (cm < 1e-4f) cm = 1e-4f;
    return noise * cm / dt;
}
// Vibration sensor (115): intensity state[1] (sigmoid at 0.3) -> state[0] (30 ms); habituation
// state[2] (3 s), sensitisation state[3] (10 s); a 150 pS MET current (Na+/K+ 0.6/0.4).
static double mod_vibration(Modifier* m, Segment* s, float dt) {
    double I0 = m->state[1];
    if (I0 < 0.0f) I0 = 0.0f;
    if (I0 > 1.0f) I0 = 1.0f;
    float act = 1.0f / (1.0f + exp(-3.0f * (I0 - 0.3f) * 10.0f));
    m->state[0] += (act - m->state[0]) * dt / 30.0f;
    clamp_state01(m, 0);
    sensor_relax(m, 2, m->state[0], dt, 3000.0f);
    sensor_relax(m, 3, m->state[0], dt, 10000.0f);
    double ks = (m->state[4] > 0.0f) ? m->state[4] : 0.25f;
    double g = sensor_gain(m->state[2], m->state[3], 0.6f, ks);
    double e = m->state[0] * g;
    float area = eff_area(m, s);
    SensorI I = sensor_ghk(s, 150.0f, 0.6f, 0.4f, 0.0f, m->density * area * e, dt);
    return -(I.na + I.k);
}
// Otolith hair cell (116): shear state[1] along the preferred direction (cosine state[5], unset
// below -99), reduced by Ca2+ adaptation state[2] (30 ms), sets the MET open probability
// state[0] (tonic 0.15, 1 ms); a 150 pS Na+/K+/Ca2+ (0.6/0.3/0.1) current.
static double mod_otolith(Modifier* m, Segment* s, float dt) {
    double sh = m->state[1];
    if (sh < 0.0f) sh = 0.0f;
    if (sh > 1.0f) sh = 1.0f;
    double c = (m->state[5] > -99.0f) ? m->state[5] : 1.0f;
    if (c < -1.0f) c = -1.0f;
    if (c > 1.0f) c = 1.0f;
    float es = sh * c;
    if (es < 0.0f) es = 0.0f;
    double adapt = m->state[2];
    float ad = es * (1.0f - 0.7f * adapt);
    double k = (m->state[4] > 0.0f) ? m->state[4] : 5.0f;
    float target = 0.15f + 0.85f / (1.0f + exp(-k * (ad - 0.15f) * 10.0f));
    m->state[0] += (target - m->state[0]) * dt / 1.0f;
    clamp_state01(m, 0);
    m->state[2] += (m->state[0] - m->state[2]) * dt / 30.0f;
    clamp_state01(m, 2);
    float area = eff_area(m, s);
    SensorI I = sensor_ghk(s, 150.0f, 0.60f, 0.30f, 0.10f, m->density * area * m->state[0], dt);
    float tot = I.na - I.k + I.ca;
    return -tot;
}
// Semicircular-canal cupula (117): a damped torsion pendulum driven by angular acceleration
// state[1] (theta state[0], omega state[3]; omega0^2 state[4] or 1, 2 zeta omega0 state[5] or
// 0.9) with an 8 s elastic return; |theta| opens the hair cell (state[2] is its 10 ms Ca2+
// adaptation); a 150 pS Na+/K+/Ca2+ (0.55/0.35/0.10) current.
static double mod_cupula(Modifier* m, Segment* s, float dt) {
    double a = m->state[1];
    if (a < -1.0f) a = -1.0f;
    if (a > 1.0f) a = 1.0f;
    double w2 = (m->state[4] > 0.0f) ? m->state[4] : 1.0f;
    double zw = (m->state[5] > 0.0f) ? m->state[5] : 0.9f;
    float gam = 1.0f;
    double om = m->state[3], th = m->state[0];
    float dom = a * gam - zw * om - w2 * th;
    om += dom * dt / 1000.0f;
    th += om * dt / 1000.0f;
    th *= exp(-dt / 8000.0f);
    om *= exp(-dt / 8000.0f);
    if (th < -1.0f) th = -1.0f;
    if (th > 1.0f) th = 1.0f;
    m->state[0] = th;
    m->state[3] = om;
    float defl = fabsf(th);
    double hs = (m->state[4] > 0.0f) ? m->state[4] * 0.4f : 2.5f;
    float target = 1.0f / (1.0f + exp(-hs * (defl - 0.1f) * 15.0f));
    m->state[2] += (target - m->state[2]) * dt / 10.0f;
    clamp_state01(m, 2);
    double open = target * (1.0f - 0.6f * m->state[2]);
    float area = eff_area(m, s);
    SensorI I = sensor_ghk(s, 150.0f, 0.55f, 0.35f, 0.10f, m->density * area * open, dt);
    float tot = I.na - I.k + I.ca;
    return -tot;
}
// Cryptochrome magnetic sensor (120): radical-pair yield light x sin^2(theta) x wavelength x
// sensitivity (states 1, 2, 5, 4) -> state[0] (50 ms); it multiplies the first light sensor's
// state[0] on the segment by 1 + 0.3 state[0] (capped at 1). No current.
static double mod_cryptochrome(Modifier* m, Segment* s, float dt) {
    double L = m->state[1];
    if (L < 0.0f) L = 0.0f;
    if (L > 1.0f) L = 1.0f;
    double s2 = m->state[2];
    if (s2 < 0.0f) s2 = 0.0f;
    if (s2 > 1.0f) s2 = 1.0f;
    double wf = (m->state[5] > 0.0f) ? m->state[5] : 1.0f;
    double sens = (m->state[4] > 0.0f) ? m->state[4] : 1.0f;
    float rp = L * s2 * wf * sens;
    m->state[0] += (rp - m->state[0]) * dt / 50.0f;
    clamp_state01(m, 0);
    for (int i = 0; i < s->modifier_count; i++)
        if (s->modifiers[i].type == 95) {
            float kc = 0.3f;
            double boost = 1.0f + kc * m->state[0];
            s->modifiers[i].state[0] *= boost;
            if (s->modifiers[i].state[0] > 1.0f) s->modifiers[i].state[0] = 1.0f;
            break;
        }
    return 0.0f;
}

// ER Ca2+ store (hh_biological.cpp:3473-3591): SERCA (126) pumps cytosolic Ca2+ into the
// segment's ER pool Segment.ER_Ca2, IP3R (127) and RyR (128) release it back; none of them
// carries a membrane current. Constants from units.h:130-156.
static float seg_ip3(const Segment* s) {   // modifier_helpers.h:1491-1495: the first PLC's state[0]
    for (uint8_t j = 0; j < s->modifier_count; j++)
        if (s->modifiers[j].type == 136) return s->modifiers[j].state[0];
    return 0.0f;
}
static double mod_serca(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:3473-3506
    float area = eff_area(m, s);
    double km = (m->state[1] > 0.0f) ? m->state[1] : 0.00027f;
    float ca = s->intra_ions[2];
    if (ca < 1e-9f) ca = 1e-9f;
    float fca = (ca * ca) / (ca * ca + km * km);   // Hill 2
    float atp = s->ATP_level;
    if (atp < 0.0f) atp = 0.0f;
    float fatp = atp / (atp + 0.1f);
    double N = m->density * area;
    float flux = N * 100.0f * fca * fatp * 2.0f * dt * 1e-3f / 6.02214076e23f;   // mol of Ca2+
    float dconc = flux / (volume(s) * 1e-18f);
    float avail = s->intra_ions[2] - 1e-9f;
    if (avail < 0.0f) avail = 0.0f;
    float moved = (dconc < avail) ? dconc : avail;
    s->intra_ions[2] -= moved;
    s->ER_Ca2 += moved;
    floor_ions(s->intra_ions);
    if (s->ER_Ca2 < 0.0f) s->ER_Ca2 = 0.0f;
    s->ATP_level -= flux * 1e6f;   // one ATP per two Ca2+, in the pool's units
    if (s->ATP_level < 0.0f) s->ATP_level = 0.0f;
    return 0.0f;
}
static double mod_ip3r(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:3508-3548
    float ip3 = seg_ip3(s);
    double h = m->state[1];
    float ca = s->intra_ions[2];
    if (ca < 1e-9f) ca = 1e-9f;
    float er = s->ER_Ca2;
    if (er < 0.0f) er = 0.0f;
    float minf = ip3 / (ip3 + 0.00013f);   // Li-Rinzel: m_inf^3 n_inf^3 h^3
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
    double dconc = 5.0f * m->density * popen * drive * dt * 1e-3f;
    float avail = er - 1e-9f;
    if (avail < 0.0f) avail = 0.0f;
    float moved = (dconc < avail) ? dconc : avail;
    s->ER_Ca2 -= moved;
    s->intra_ions[2] += moved;
    floor_ions(s->intra_ions);
    if (s->ER_Ca2 < 0.0f) s->ER_Ca2 = 0.0f;
    m->state[1] = h;
    return 0.0f;
}
static double mod_ryr(Modifier* m, Segment* s, float dt) {   // hh_biological.cpp:3550-3591
    double w = m->state[0];
    float ca = s->intra_ions[2];
    if (ca < 1e-9f) ca = 1e-9f;
    float er = s->ER_Ca2;
    if (er < 0.0f) er = 0.0f;
    float ckb = ca / 0.000636f;   // Keizer-Levine: (c/K_b)^3, (K_a/c)^4
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
    double dconc = 20.0f * m->density * popen * drive * dt * 1e-3f;
    float avail = er - 1e-9f;
    if (avail < 0.0f) avail = 0.0f;
    float moved = (dconc < avail) ? dconc : avail;
    s->ER_Ca2 -= moved;
    s->intra_ions[2] += moved;
    floor_ions(s->intra_ions);
    if (s->ER_Ca2 < 0.0f) s->ER_Ca2 = 0.0f;
    m->state[0] = w;
    return 0.0f;
}
// PLC (136, modifier_helpers.h:1553-1580): the segment's Gq-receptor activations (states[0] of
// the types is_gq_receptor_type lists, modifier_helpers.h:1426-1432: 52-62 and 188) drive IP3
// (state[0]), DAG (state[1]) and a desensitisation (state[2]); constants units.h:342-349.
static double mod_plc(Modifier* m, Segment* s, float dt) {
    float gq = 0.0f;
    for (uint8_t j = 0; j < s->modifier_count; j++) {
        const uint8_t t = s->modifiers[j].type;
        if ((t >= 52 && t <= 62) || t == 188) gq += s->modifiers[j].state[0];
    }
    if (gq > 1.0f) gq = 1.0f;
    float vip3 = (m->state[3] > 0.0f) ? m->state[3] : 0.0005f;
    float vdag = (m->state[4] > 0.0f) ? m->state[4] : 0.5f;
    float drive = 0.02f + gq;
    float ip3 = m->state[0];
    ip3 += (vip3 * m->density * drive) * dt * 0.001f - ip3 * dt / 1000.0f;
    if (ip3 < 0.0f) ip3 = 0.0f;
    m->state[0] = ip3;
    float dag = m->state[1];
    dag += (vdag * m->density * drive) * dt * 0.001f - dag * dt / 2000.0f;
    if (dag < 0.0f) dag = 0.0f;
    m->state[1] = dag;
    float dz = m->state[2];
    dz += (gq - dz) * dt / 70000.0f;
    if (dz < 0.0f) dz = 0.0f;
    if (dz > 1.0f) dz = 1.0f;
    m->state[2] = dz;
    return 0.0f;
}

// Presynaptic release modulators (receptors_biological.cpp:427-491), the CPU step's dispatch
// with no plugin registered (kernel_dispatch.h:575-604): Gi/o types lower release by
// 0.7 * state[0] * density * area (state[0] relaxing to the cleft activation nt / (nt + 0.5)
// with tau 50 ms), Gs types raise it by 0.5 * ... (tau 30 ms), the nicotinic one by
// 0.3 * activation * density * area with no state. area is state[7] (the caller stores the
// synapse's pre_area there) read as float, 1 if not positive. Other IDs: factor 1.
static float release_modulator(Modifier* m, const abi_float* cleft, float dt) {
    int nt, kind;   // kind 0: Gi/o, 1: Gs, 2: ionotropic
    switch (m->type) {
        case 18: case 19: nt = 3; kind = 0; break;   // 5-HT1A, 5-HT1B: serotonin
        case 22: nt = 2; kind = 0; break;            // D2: dopamine
        case 27: nt = 5; kind = 0; break;            // A1: adenosine
        case 46: nt = 5; kind = 1; break;            // A2A
        case 44: nt = 2; kind = 1; break;            // D1
        case 29: case 30: nt = 0; kind = 0; break;   // mGluR2/3: glutamate
        case 36: nt = 9; kind = 0; break;            // CB1: endocannabinoid
        case 38: nt = 8; kind = 0; break;            // H3: histamine
        case 13: nt = 4; kind = 2; break;            // nicotinic: acetylcholine
        case 35: nt = 6; kind = 0; break;            // alpha2: noradrenaline
        default: return 1.0f;
    }
    float c = cleft[nt];
    float act = c / (c + 0.5f);
    float area = (m->state[7] > 0.0f) ? (float)m->state[7] : 1.0f;
    double scale = m->density * area;
    if (kind == 2) return 1.0f + 0.3f * act * scale;
    m->state[0] += (act - m->state[0]) * dt / (kind == 0 ? 50.0f : 30.0f);
    if (kind == 0) return 1.0f - 0.7f * m->state[0] * scale;
    return 1.0f + 0.5f * m->state[0] * scale;
}

// AQP4 (186, hh_biological.cpp:4114-4129, modifier_helpers.h:395-396): water follows the
// osmotic difference of the summed (float) concentrations and dilutes every intracellular ion
// by a factor clamped to [0.95, 1]; no current, no floor.
static double chan_aqp4(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    float si = 0.0f, se = 0.0f;
    for (int i = 0; i < 7; i++) {
        si += s->intra_ions[i];
        se += s->extra_ions[i];
    }
    float dpi = si - se;
    double J = m->density * 24e-14f * area * dpi * 1e-6f;
    float f = 1.0f / (1.0f + J * dt);
    if (f < 0.95f) f = 0.95f;
    if (f > 1.0f) f = 1.0f;
    for (int i = 0; i < 7; i++) s->intra_ions[i] *= f;
    return 0.0f;
}

// PMCA (98, hh_biological.cpp:2574-2614): Hill-n Ca2+ extrusion, ATP-dependent. The Hill powers
// are glibc pow calls; with the default n = 2 both are exact squares (a float squared, and the
// default K_m squared, fit a double), which is how they are computed here.
static double chan_pmca(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    double km = (m->state[1] > 0.0f) ? m->state[1] : 0.0003f;
    double hill = (m->state[2] > 0.0f) ? m->state[2] : 2.0f;
    if (hill != 2.0 || (m->state[1] > 0.0f && (double)(float)m->state[1] != m->state[1])) g_unsupported++;
    float atp = s->ATP_level;
    if (atp < 0.0f) atp = 0.0f;
    float f_atp = atp / (atp + 0.2f);
    float ca = s->intra_ions[2];
    if (ca < 1e-9f) ca = 1e-9f;
    float ca_pow = (double)ca * (double)ca;
    float f_ca = ca_pow / (ca_pow + km * km);
    double N = m->density * area;
    float ip = N * 100.0f * f_ca * f_atp * 1.602176634e-19f * 2.0f * kAtoPa;
    float vol = volume(s);
    double cpv = kFaraday * vol / 1000.0;
    float tr = ip * dt / (2.0f * cpv);
    float ca_av = s->intra_ions[2] - 1e-9f;
    if (ca_av < 0.0f) ca_av = 0.0f;
    if (tr > ca_av) tr = ca_av;
    s->intra_ions[2] -= tr;
    s->extra_ions[2] += tr / kEvr;
    floor_ions(s->intra_ions);
    s->ATP_level -= ip * dt * 1e-8f;
    if (s->ATP_level < 0.0f) s->ATP_level = 0.0f;
    return -ip;
}

// NCX (99, hh_biological.cpp:2616-2646, modifier_helpers.h:1181-1201): 3 Na+ in / 1 Ca2+ out,
// voltage-dependent, allosteric Ca activation, either direction.
static double chan_ncx(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    float na_i = s->intra_ions[0]; if (na_i < 1e-9f) na_i = 1e-9f;
    float na_o = s->extra_ions[0]; if (na_o < 1e-9f) na_o = 1e-9f;
    float ca_i = s->intra_ions[2]; if (ca_i < 1e-9f) ca_i = 1e-9f;
    float ca_o = s->extra_ions[2]; if (ca_o < 1e-9f) ca_o = 1e-9f;
    float vv = (float)s->V_m * 1e-3f;
    float frt = kFaraday / (kGas * kTemp);
    float e_eta = exp(0.35f * frt * vv);
    float e_eta1 = exp((0.35f - 1.0f) * frt * vv);
    float nai3 = na_i * na_i * na_i;
    float nao3 = na_o * na_o * na_o;
    float fwd = e_eta1 * nao3 * ca_i;
    float rev = e_eta * nai3 * ca_o;
    float den = 1.0f + 0.27f * e_eta1;
    float cr = 0.000125f / ca_i;
    float allo = 1.0f / (1.0f + cr * cr);
    float fn = allo * (fwd - rev) / (den * 530.2f);
    double N = m->density * area;
    float I = N * 500.0f * fn * 1.602176634e-19f * kAtoPa;
    float vol = volume(s);
    double cpv = kFaraday * vol / 1000.0;
    float tr = I * dt / cpv;
    if (tr > ca_i) tr = ca_i;
    if (tr < -na_i / 3.0f) tr = -na_i / 3.0f;
    s->intra_ions[2] -= tr;
    s->extra_ions[2] += tr / kEvr;
    s->intra_ions[0] += 3.0f * tr;
    s->extra_ions[0] -= 3.0f * tr / kEvr;
    floor_ions(s->intra_ions);
    return I;
}

// KCC2 (100, hh_biological.cpp:2648-2703): electroneutral K+-Cl- cotransport down the product
// gradient, limited by the ions available on the losing side.
static double chan_kcc2(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    double km_k = (m->state[1] > 0.0f) ? m->state[1] : 6.5f;
    double km_cl = (m->state[2] > 0.0f) ? m->state[2] : 6.5f;
    float k_i = s->intra_ions[1]; if (k_i < 1e-9f) k_i = 1e-9f;
    float cl_i = s->intra_ions[3]; if (cl_i < 1e-9f) cl_i = 1e-9f;
    float k_o = s->extra_ions[1]; if (k_o < 1e-9f) k_o = 1e-9f;
    float cl_o = s->extra_ions[3]; if (cl_o < 1e-9f) cl_o = 1e-9f;
    float f_k = k_i / (k_i + km_k);
    float f_cl = cl_i / (cl_i + km_cl);
    float drv = 1.0f - (k_o * cl_o) / (k_i * cl_i);
    if (drv > 1.0f) drv = 1.0f;
    if (drv < -1.0f) drv = -1.0f;
    float flux = 5.0f * 1e-20f * m->density * area * f_k * f_cl * drv * dt;
    float dc = flux / (volume(s) * 1e-18f);
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
    s->intra_ions[1] -= dc;
    s->intra_ions[3] -= dc;
    s->extra_ions[1] += dc / kEvr;
    s->extra_ions[3] += dc / kEvr;
    floor_ions(s->intra_ions);
    return 0.0f;
}

// NKCC1 (122, hh_biological.cpp:2705-2760): Na+-K+-2Cl- uptake with trans-stimulation by
// external Na+ and trans-inhibition by external K+.
static double chan_nkcc1(Modifier* m, Segment* s, float dt) {
    float area = eff_area(m, s);
    double km_nai = (m->state[1] > 0.0f) ? m->state[1] : 10.0f;
    double km_ki = (m->state[2] > 0.0f) ? m->state[2] : 15.0f;
    double km_cli = (m->state[3] > 0.0f) ? m->state[3] : 50.0f;
    double km_nao = (m->state[4] > 0.0f) ? m->state[4] : 0.5f;
    double km_ko = (m->state[5] > 0.0f) ? m->state[5] : 0.2f;
    float na_i = s->intra_ions[0], k_i = s->intra_ions[1], cl_i = s->intra_ions[3];
    float na_o = s->extra_ions[0], k_o = s->extra_ions[1];
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
    float flux = 7.0f * 1e-20f * m->density * area * fna * fk * fcl * stim * inhib * dt;
    float dc = flux / (volume(s) * 1e-18f);
    float naa = (s->extra_ions[0] - 1e-9f) * kEvr;
    float ka = (s->extra_ions[1] - 1e-9f) * kEvr;
    float cla = (s->extra_ions[3] - 1e-9f) * kEvr / 2.0f;
    float mn = (naa < ka) ? naa : ka;
    if (cla < mn) mn = cla;
    if (mn < 0.0f) mn = 0.0f;
    if (dc > mn) dc = mn;
    s->intra_ions[0] += dc;
    s->intra_ions[1] += dc;
    s->intra_ions[3] += 2.0f * dc;
    s->extra_ions[0] -= dc / kEvr;
    s->extra_ions[1] -= dc / kEvr;
    s->extra_ions[3] -= 2.0f * dc / kEvr;
    floor_ions(s->intra_ions);
    return 0.0f;
}

// Ca2+ buffer (102, hh_biological.cpp:2825-2865): first-order binding to a pool of capacity
// state[0] (0.15 mM) x density; state[1] holds the bound Ca (initialised at equilibrium).
static double chan_ca_buffer(Modifier* m, Segment* s, float dt) {
    double bref = (m->state[0] > 0.0f) ? m->state[0] : 0.15f;
    double btot = bref * m->density;
    double kd = (m->state[2] > 0.0f) ? m->state[2] : 0.0005f;
    double kon = (m->state[3] > 0.0f) ? m->state[3] : 0.1f;
    double koff = (m->state[4] > 0.0f) ? m->state[4] : (kd * kon);
    double bound = m->state[1];
    if (bound <= 0.0f && btot > 0.0f) {
        float c0 = s->intra_ions[2];
        bound = btot * c0 / (c0 + kd);
        m->state[1] = bound;
    }
    float cf = s->intra_ions[2];
    if (cf < 1e-9f) cf = 1e-9f;
    float bf = btot - bound;
    if (bf < 0.0f) bf = 0.0f;
    float db = (kon * cf * bf - koff * bound) * dt;
    bound += db;
    if (bound < 0.0f) bound = 0.0f;
    if (bound > btot) bound = btot;
    s->intra_ions[2] -= db;
    if (s->intra_ions[2] < 1e-9f) {
        float def = 1e-9f - s->intra_ions[2];
        s->intra_ions[2] = 1e-9f;
        bound -= def;
        if (bound < 0.0f) bound = 0.0f;
    }
    m->state[1] = bound;
    return 0.0f;
}

// Na+-coupled transmitter transporters (modifier_helpers.h:418-455, hh_biological.cpp:725-780):
// Michaelis-Menten uptake of one transmitter from the segment's extracellular pool, limited by
// the transmitter, extracellular Na+ and (K+-countertransport) intracellular K+ available, then
// both ion arrays floored at 1 nM. Returns the (depolarising) transporter current in pA.
struct Transporter { int nt; float km, imax, rel_per_mm; int na_in, cl_in, k_out, net; };
constexpr Transporter kGlast = {0, 10.0f, 2.0f, 667.0f, 3, 0, 1, 2};    // 147
constexpr Transporter kGlyt1 = {7, 30.0f, 2.0f, 667.0f, 2, 1, 0, 1};    // 148
constexpr Transporter kGat3 = {1, 0.5f, 2.0f, 667.0f, 2, 1, 0, 1};      // 149
constexpr Transporter kEaat3 = {0, 20.0f, 2.0f, 667.0f, 3, 0, 1, 2};    // 183
constexpr Transporter kGlyt2 = {7, 55.7f, 2.0f, 667.0f, 3, 1, 0, 2};    // 184
constexpr Transporter kGat1 = {1, 4.7f, 2.0f, 667.0f, 3, 1, 0, 2};      // 185
static double chan_transporter(Modifier* m, Segment* s, float dt, const Transporter& t) {
    float area = eff_area(m, s);
    float nt = s->neurotransmitter[t.nt];
    if (nt < 0.0f) nt = 0.0f;
    float frac = nt / (nt + t.km);
    float I = t.imax * m->density * area * (1e-8f * 1e6f) * frac;
    float vol = (float)volume(s);
    double cpm = kFaraday * vol / 1000.0;
    float tr = I * dt / ((float)t.net * cpm);
    float nt_av = nt / t.rel_per_mm;
    float na_av = s->extra_ions[0] * kEvr / (float)t.na_in;
    if (tr > nt_av) tr = nt_av;
    if (tr > na_av) tr = na_av;
    if (t.k_out && tr > s->intra_ions[1]) tr = s->intra_ions[1];
    if (tr < 0.0f) tr = 0.0f;
    s->intra_ions[0] += (float)t.na_in * tr;
    s->extra_ions[0] -= (float)t.na_in * tr / kEvr;
    if (t.cl_in) {
        s->intra_ions[3] += (float)t.cl_in * tr;
        s->extra_ions[3] -= (float)t.cl_in * tr / kEvr;
    }
    if (t.k_out) {
        s->intra_ions[1] -= (float)t.k_out * tr;
        s->extra_ions[1] += (float)t.k_out * tr / kEvr;
    }
    s->neurotransmitter[t.nt] -= tr * t.rel_per_mm;
    if (s->neurotransmitter[t.nt] < 0.0f) s->neurotransmitter[t.nt] = 0.0f;
    floor_ions(s->intra_ions);
    floor_ions(s->extra_ions);
    return I;
}

static double dispatch(uint8_t type, Modifier* m, Segment* s, float dt) {
    switch (type) {
        case T_MYELIN: return 0.0;   // hh_biological.cpp:450-453
        case T_PUMP: return chan_pump(m, s, dt);
        case T_AMPA: return iono_current(m, s, dt, kAmpa);
        case T_NMDA: return rec_nmda(m, s, dt);
        case T_GABAA: return iono_current(m, s, dt, kGabaa);
        case T_SK: return chan_sk(m, s, dt);
        case T_BK: return chan_bk(m, s, dt);
        case T_K2P: return chan_k2p(m, s, dt);
        case T_KV7: return chan_kv7(m, s, dt);
        case T_KVDR: return chan_kvdr(m, s, dt);
        case T_HCN1: return chan_hcn1(m, s, dt);
        case T_NALCN: return chan_nalcn(m, s, dt);
        case T_CX36: case T_CX43: return chan_connexin(m, s, dt);
        case T_CAV21: return chan_cav21(m, s, dt);
        case T_NAV16: return chan_nav(m, s, dt, kNav16, 20.7f, 5.0f, 0.2f);
        case T_NAV12: return chan_nav(m, s, dt, kNav12, 20.8f, 10.0f, 0.7f);
        case T_KV31: return chan_kv31(m, s, dt);
        case T_KV11: return chan_kv11(m, s, dt);
        case T_HCN2: return chan_hcn1(m, s, dt, -95.0f, 557.0f, 1.54f);
        case 10: return iono_current(m, s, dt, kKainate);
        case T_GLYCINE: return iono_current(m, s, dt, kGlycine);
        case T_NICOTINIC: return iono_current(m, s, dt, kNicotinic);
        case T_5HT3: return iono_current(m, s, dt, k5ht3);
        case 97: return iono_current(m, s, dt, kAlpha7);
        case 194: return rec_nmda(m, s, dt, kNmda2b);
        case 15: return iono_current(m, s, dt, kP2x);
        case 134: return iono_current(m, s, dt, kP2x2);
        case 64: case 71: return iono_current(m, s, dt, kAmpa);
        case 73: return iono_current(m, s, dt, kGabaa);
        case 67: return iono_current(m, s, dt, kAlpha7);
        case 68: return iono_current(m, s, dt, kNicotinic);
        case 74: return chan_gio(m, s, dt, 1);
        case 63: case 72: return rec_nmda_mg(m, s, dt);
        case 75: return rec_nmda_mg(m, s, dt, msn_scale(s, 65));
        case 76: return iono_current(m, s, dt, kAmpa, msn_scale(s, 66));
        case 131: return chan_hcn1(m, s, dt, -77.0f, 1408.0f, 0.54f);
        case 132: return chan_hcn1(m, s, dt, -100.0f, 3586.0f, 0.51f);
        case T_CAV13: return chan_cav13(m, s, dt);
        case T_CL: return chan_cl(m, s, dt);
        case T_KV42: return chan_kv42(m, s, dt);
        case T_KV21: return chan_kv21(m, s, dt);
        case T_NAP: return chan_nap(m, s, dt, false);
        case T_NAP_TAD: return chan_nap(m, s, dt, true);
        case T_KIR: return chan_kir(m, s, dt);
        case T_NAV12DEV: return chan_nav12dev(m, s, dt);
        case 177: return chan_nav_simple(m, s, dt, kNav13);
        case 169: return chan_nav_simple(m, s, dt, kNav17);
        case 170: return chan_nav_simple(m, s, dt, kNav18);
        case 171: return chan_nav_simple(m, s, dt, kNav19);
        case 199: return chan_nav11(m, s, dt);
        case 157: return chan_kv1x(m, s, dt, true);
        case 154: return chan_kv1x(m, s, dt, false);
        case 206: return chan_kv14(m, s, dt);
        case 207: return chan_kv16_22(m, s, dt, false);
        case 205: return chan_kv16_22(m, s, dt, true);
        case 203: return chan_kvgen(m, s, dt, kKv32);
        case 204: return chan_kvgen(m, s, dt, kKv33);
        case 202: return chan_kvgen(m, s, dt, kKv43);
        case 160: case 161: case 166: case 167: case 208: return chan_kleak(m, s, dt, type);
        case 178: return chan_kir71_kcnq4(m, s, dt, false);
        case 176: return chan_kir71_kcnq4(m, s, dt, true);
        case 168: return chan_cav22(m, s, dt);
        case 200: return chan_cavt(m, s, dt, kCav31);
        case 201: return chan_cavt(m, s, dt, kCav32);
        case 191: return chan_cavt(m, s, dt, kCav33);
        case 210: case 162: case 211: return chan_trp_naca(m, s, dt, type);
        case 175: return chan_trpv3(m, s, dt);
        case 182: return chan_enac(m, s, dt);
        case 172: return chan_trp_lig(m, s, dt, kTrpv1);
        case 173: return chan_trp_lig(m, s, dt, kTrpm8);
        case 174: return chan_trp_lig(m, s, dt, kTrpa1);
        case 179: return chan_trp_lig(m, s, dt, kTrpc67);
        case 163: return chan_trpc5_m4(m, s, dt, false);
        case 164: return chan_trpc5_m4(m, s, dt, true);
        case 180: return chan_calhm(m, s, dt);
        case 159: return chan_kv34(m, s, dt);
        case 101: return chan_ca_leak(m, s, dt);
        case 96: return chan_pir_ttype(m, s, dt);
        case 84: return chan_ca_l(m, s, dt);
        case 144: return chan_cav14(m, s, dt);
        case 145: return chan_cabp4(m, s, dt);
        case 146: return chan_cacc(m, s, dt);
        case 152: return chan_can(m, s, dt);
        case 186: return chan_aqp4(m, s, dt);
        case 190: return chan_tmc1(m, s, dt);
        case T_PUMP_A3: return chan_pump(m, s, dt, PUMP_A3);
        case T_PUMP_GLIA: return chan_pump(m, s, dt, PUMP_GLIA);
        case T_PMCA: return chan_pmca(m, s, dt);
        case T_NCX: return chan_ncx(m, s, dt);
        case T_KCC2: return chan_kcc2(m, s, dt);
        case T_NKCC1: return chan_nkcc1(m, s, dt);
        case T_CA_BUFFER: return chan_ca_buffer(m, s, dt);
        case T_GLAST: return chan_transporter(m, s, dt, kGlast);
        case T_GLYT1: return chan_transporter(m, s, dt, kGlyt1);
        case T_GAT3: return chan_transporter(m, s, dt, kGat3);
        case T_EAAT3: return chan_transporter(m, s, dt, kEaat3);
        case T_GLYT2: return chan_transporter(m, s, dt, kGlyt2);
        case T_GAT1: return chan_transporter(m, s, dt, kGat1);
        case T_ECB: return mod_ecb(m, s, dt);
        case 16: return mod_trpv1_receptor(m, s, dt);
        case 192: return mod_rod(m, s, dt);
        case 87: return mod_mapk(m, s, dt);
        case 88: return mod_nfkb(m, s, dt);
        case 89: return mod_mtorc1(m, s, dt);
        case 90: return mod_protein_synth(m, s, dt);
        case 91: return mod_cofilin(m, s, dt);
        case 92: return mod_mmp9(m, s, dt);
        case 6: return mod_physical_sensor(m, s, dt);
        case 94: return mod_hunger_sensor(m, s, dt);
        case 95: return mod_light_sensor(m, s, dt);
        case 108: return mod_olfactory(m, s, dt);
        case 111: return mod_thermal_noise(m, s, dt);
        case 115: return mod_vibration(m, s, dt);
        case 116: return mod_otolith(m, s, dt);
        case 117: return mod_cupula(m, s, dt);
        case 120: return mod_cryptochrome(m, s, dt);
        case 3: return mod_metabotropic(m, s, dt);
        case 5: return mod_release_site(m, s, dt);
        case 188: return mod_mrgpr(m, s, dt);
        case 77: return mod_soma_atp(m, s, dt);
        case 78: return mod_soma_thermal(m, s, dt);
        case 79: return mod_creb(m, s, dt);
        case 126: return mod_serca(m, s, dt);
        case 127: return mod_ip3r(m, s, dt);
        case 128: return mod_ryr(m, s, dt);
        case 136: return mod_plc(m, s, dt);
        case 82: return mod_camkii(m, s, dt);
        case 83: return mod_calcineurin(m, s, dt);
        case 113: return mod_pp1(m, s, dt);
        case 137: return mod_pkc(m, s, dt);
        // IDs without code in the reference (the reserved 80, 81, 93, 138, 139, 143; the plugin
        // stubs 124 and 125; 212 and above): no current, no state change
        // (kernel_dispatch.h:565 -> dispatch_generated.cpp:26-31 with no plugin loaded)
        case 0: case 80: case 81: case 93: case 124: case 125: case 138: case 139: case 143: return 0.0;
        default:
            if (type >= 212) return 0.0;
            if (type >= 17 && type <= 40) return chan_gio(m, s, dt, gio_nt(type));
            if (type >= 41 && type <= 51) return chan_gs(m, s, dt, gs_nt(type));
            // Gq-coupled (52-62 and the muscarinic 70; modifier_helpers.h:1600-1621): the Gs
            // signal's form and constants. 5-HT2A/B/C serotonin, alpha1A/B noradrenaline,
            // M1/M3/M5 and 70 acetylcholine, mGluR1/5 glutamate, H1 histamine.
            if ((type >= 52 && type <= 62) || type == 70)
                return chan_gs(m, s, dt, type <= 54 ? 3 : type <= 56 ? 6 : type <= 59 || type == 70 ? 4 : type <= 61 ? 0 : 8);
            g_unsupported++;
            return 0.0;
    }
}

// ---- step 2.5: electrodiffusion (diffusion_kernel.h:41-364) ----
static inline float bernoulli(float x) {
    if (fabsf(x) < 1e-3f) return 1.0f - 0.5f * x;
    if (x > 50.0f) return 0.0f;
    if (x < -50.0f) return -x;
    return x / (exp(x) - 1.0f);
}

struct Ctx {
    std::vector<double> phi, sigma, diag, phi_new, goff_d;
    std::vector<double> dacc, dext;
    std::vector<float> gdiag, goff, inpd, vold, vnew, off;
    std::vector<float> dnt_in, dnt, datp, dtemp, dtxn;
    std::vector<uint8_t> delay;   // MAX_DELAY_SLOTS(4) x 16-byte DelayedRelease per synapse
    uint32_t delay_syn = UINT32_MAX;
    std::vector<uint32_t> spikes;
    int substeps_override = 0;
    // Residual-terminated Jacobi (an extension, not in the reference, which always sweeps 50
    // times): each connected cell stops sweeping after the first sweep whose largest update is
    // <= jac_tol, or after jac_max sweeps.
    bool jac_on = false;
    double jac_tol = 0.0;
    int jac_max = 50;
    std::vector<uint32_t> comp;   // cell (component over all connections) of each segment
    uint32_t ncomp = 0;
    // Cable mode (an extension): ion concentrations, ATP, temperature and transmitters are held
    // (every channel sees the stored concentrations, ion injections are ignored), phi_e = 0, and
    // step 2.5 only moves V. scale: per-connection factor on axial conductance and flux (empty: 1).
    bool cable = false;
    std::vector<float> scale;
    // Protein -> release-parameter mapping (simulation_host.cpp:407-414), active while the
    // reference's consolidation runs: nucleus k's protein levels set the vesicle parameters of
    // every chemical synapse whose presynaptic segment belongs to neuron k.
    std::vector<Nucleus> nuclei;
};

// kernel_dispatch.h:651-663: Munc13-1 (gene 42) sets RRP_max and the recovery rate, which
// complexin (44) lowers; synaptotagmin-1 (40) the release probability.
static void protein_map(Synapse* y, const abi_float* protein) {
    float munc13 = protein[42];
    y->RRP_max = 10.0f + munc13 * 40.0f;
    float syt1 = protein[40];
    y->release_probability = 0.3f + syt1 * 0.7f;
    float cplx1 = protein[44];
    y->RRP_recovery = 0.02f + munc13 * 0.08f - cplx1 * 0.01f;
    if (y->RRP_recovery < 0.005f) y->RRP_recovery = 0.005f;
    if (y->RRP_recovery > 0.2f) y->RRP_recovery = 0.2f;
}

static uint32_t uf_root(std::vector<uint32_t>& p, uint32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}
static void cells(Ctx& c, uint32_t n, const IntraConnection* cn, uint32_t nc) {
    std::vector<uint32_t> p(n);
    for (uint32_t i = 0; i < n; i++) p[i] = i;
    for (uint32_t i = 0; i < nc; i++) {
        uint32_t a = cn[i].segment_a, b = cn[i].segment_b;
        if (a >= n || b >= n) continue;
        a = uf_root(p, a);
        b = uf_root(p, b);
        if (a != b) p[a > b ? a : b] = a < b ? a : b;
    }
    c.comp.assign(n, 0);
    std::vector<uint32_t> id(n, UINT32_MAX);
    c.ncomp = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t r = uf_root(p, i);
        if (id[r] == UINT32_MAX) id[r] = c.ncomp++;
        c.comp[i] = id[r];
    }
}
// After a sweep: cells whose largest update was <= tol are done. Returns true when all are.
static bool cells_settle(Ctx& c, std::vector<double>& res, std::vector<uint8_t>& done) {
    bool all = true;
    for (uint32_t k = 0; k < c.ncomp; k++) {
        if (!done[k] && res[k] <= c.jac_tol) done[k] = 1;
        if (!done[k]) all = false;
        res[k] = 0.0;
    }
    return all;
}

static void solve_phi(Ctx& c, Segment* sg, uint32_t n, IntraConnection* cn, uint32_t nc) {
    c.phi.assign(n, 0.0);
    c.sigma.assign(n, 0.0);
    c.diag.assign(n, 0.0);
    c.phi_new.assign(n, 0.0);
    c.goff_d.assign(nc, 0.0);
    for (uint32_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int ion = 0; ion < 7; ion++) s += kZ[ion] * kZ[ion] * kDiff[ion] * sg[i].extra_ions[ion];
        c.sigma[i] = kSigmaE * s;
        double gb = c.sigma[i] * membrane_area(&sg[i]) / 1.0f;
        c.diag[i] = gb;
    }
    for (uint32_t i = 0; i < nc; i++) {
        uint32_t a = cn[i].segment_a, b = cn[i].segment_b;
        if (a >= n || b >= n) continue;
        float rmin = sg[a].radius < sg[b].radius ? sg[a].radius : sg[b].radius;
        float ac = kPi * rmin * rmin;
        if (ac < 1e-6f) ac = 1e-6f;
        float dist = (sg[a].length + sg[b].length) * 0.5f;
        if (dist < 0.1f) dist = 0.1f;
        double g = 0.5 * (c.sigma[a] + c.sigma[b]) * ac / dist;
        c.goff_d[i] = g;
        c.diag[a] += g;
        c.diag[b] += g;
    }
    std::vector<double> res(c.jac_on ? c.ncomp : 0, 0.0);
    std::vector<uint8_t> done(c.jac_on ? c.ncomp : 0, 0);
    for (int it = 0; it < (c.jac_on ? c.jac_max : 50); it++) {
        c.phi_new.assign(n, 0.0);
        for (uint32_t i = 0; i < nc; i++) {
            uint32_t a = cn[i].segment_a, b = cn[i].segment_b;
            if (a >= n || b >= n) continue;
            double g = c.goff_d[i];
            if (g <= 0.0) continue;
            c.phi_new[a] += g * c.phi[b];
            c.phi_new[b] += g * c.phi[a];
        }
        for (uint32_t i = 0; i < n; i++) {
            if (c.jac_on && done[c.comp[i]]) continue;
            double v = (c.diag[i] > 1e-9) ? (c.phi_new[i] - sg[i].I_buffer) / c.diag[i] : 0.0;
            if (v > 10.0) v = 10.0;
            if (v < -10.0) v = -10.0;
            if (c.jac_on) res[c.comp[i]] = std::max(res[c.comp[i]], fabs(v - c.phi[i]));
            c.phi[i] = v;
        }
        if (c.jac_on && cells_settle(c, res, done)) break;
    }
}

struct ConnGeom { float ac, dist, va, vb; };
static inline ConnGeom conn_geom(const Segment* a, const Segment* b) {
    ConnGeom g;
    float rmin = a->radius < b->radius ? a->radius : b->radius;
    g.ac = kPi * rmin * rmin;
    if (g.ac < 1e-6f) g.ac = 1e-6f;
    g.dist = (a->length + b->length) * 0.5f;
    if (g.dist < 0.1f) g.dist = 0.1f;
    g.va = volume(a);
    g.vb = volume(b);
    if (g.va < 1e-6f) g.va = 1e-6f;
    if (g.vb < 1e-6f) g.vb = 1e-6f;
    return g;
}

static void diffusion(Ctx& c, Segment* sg, uint32_t n, IntraConnection* cn, uint32_t nc, float dt) {
    if (nc == 0) return;
    c.dext.assign((size_t)n * 7, 0.0);
    c.dacc.assign((size_t)n * 7, 0.0);
    c.dnt_in.assign((size_t)n * 15, 0.0f);
    c.dnt.assign((size_t)n * 15, 0.0f);
    c.datp.assign(n, 0.0f);
    c.dtemp.assign(n, 0.0f);
    c.dtxn.assign(n, 0.0f);
    c.gdiag.assign(n, 0.0f);
    c.goff.assign(nc, 0.0f);
    c.inpd.assign(n, 0.0f);
    c.vold.assign(n, 0.0f);
    c.vnew.assign(n, 0.0f);
    c.off.assign(n, 0.0f);
    if (c.jac_on) cells(c, n, cn, nc);
    if (c.cable) c.phi.assign(n, 0.0);
    else solve_phi(c, sg, n, cn, nc);
    const std::vector<double>& phi = c.phi;
    const bool scaled = c.scale.size() >= nc;

    const int kSub = 5;
    const float dts = dt / kSub;
    for (int sub = 0; sub < kSub; sub++) {
        std::fill(c.dacc.begin(), c.dacc.end(), 0.0);
        std::fill(c.gdiag.begin(), c.gdiag.end(), 0.0f);
        std::fill(c.inpd.begin(), c.inpd.end(), 0.0f);
        for (uint32_t i = 0; i < nc; i++) {
            uint32_t a = cn[i].segment_a, b = cn[i].segment_b;
            if (a >= n || b >= n) continue;
            if (cn[i].is_intracellular == 0) continue;
            ConnGeom g = conn_geom(&sg[a], &sg[b]);
            float sig = 0.0f;
            for (int ion = 0; ion < 7; ion++) {
                float cbar = 0.5f * (sg[a].intra_ions[ion] + sg[b].intra_ions[ion]);
                sig += kZ[ion] * kZ[ion] * kDiff[ion] * cbar;
            }
            float gax = kSigmaE * sig * g.ac / g.dist;
            if (scaled) gax *= c.scale[i];
            c.goff[i] = gax;
            c.gdiag[a] += gax;
            c.gdiag[b] += gax;
            float dV = (sg[a].V_m + phi[a]) - (sg[b].V_m + phi[b]);
            float inp = 0.0f;
            for (int ion = 0; ion < 7; ion++) {
                float ca = sg[a].intra_ions[ion], cb = sg[b].intra_ions[ion];
                float x = kZ[ion] * kFrtMv * dV;
                float fl = kDiff[ion] * g.ac * dts / g.dist * (ca * bernoulli(-x) - cb * bernoulli(x));
                if (scaled) fl *= c.scale[i];
                c.dacc[(size_t)a * 7 + ion] -= (double)fl / g.va;
                c.dacc[(size_t)b * 7 + ion] += (double)fl / g.vb;
                inp += kZ[ion] * kChargeToPa * fl / dts;
            }
            c.inpd[a] += inp - gax * dV;
            c.inpd[b] -= inp - gax * dV;
        }
        for (uint32_t i = 0; i < n; i++) { c.vold[i] = sg[i].V_m; c.vnew[i] = sg[i].V_m; }
        std::vector<double> res(c.jac_on ? c.ncomp : 0, 0.0);
        std::vector<uint8_t> done(c.jac_on ? c.ncomp : 0, 0);
        for (int it = 0; it < (c.jac_on ? c.jac_max : 50); it++) {
            std::fill(c.off.begin(), c.off.end(), 0.0f);
            for (uint32_t i = 0; i < nc; i++) {
                uint32_t a = cn[i].segment_a, b = cn[i].segment_b;
                if (a >= n || b >= n) continue;
                if (cn[i].is_intracellular == 0) continue;
                c.off[a] += c.goff[i] * c.vnew[b];
                c.off[b] += c.goff[i] * c.vnew[a];
            }
            for (uint32_t i = 0; i < n; i++) {
                if (c.jac_on && done[c.comp[i]]) continue;
                float cdt = capacitance(&sg[i]) / dts;
                if (cdt < 1e-9f) cdt = 1e-9f;
                float v = (cdt * c.vold[i] + c.inpd[i] + c.off[i]) / (cdt + c.gdiag[i]);
                if (c.jac_on) res[c.comp[i]] = std::max(res[c.comp[i]], (double)fabsf(v - c.vnew[i]));
                c.vnew[i] = v;
            }
            if (c.jac_on && cells_settle(c, res, done)) break;
        }
        for (uint32_t i = 0; i < n; i++) sg[i].V_m = c.vnew[i];
        if (!c.cable)
            for (uint32_t i = 0; i < n; i++)
                for (int ion = 0; ion < 7; ion++) sg[i].intra_ions[ion] += (float)c.dacc[(size_t)i * 7 + ion];
    }
    if (c.cable) return;

    for (int sub = 0; sub < kSub; sub++) {
        std::fill(c.dext.begin(), c.dext.end(), 0.0);
        for (uint32_t i = 0; i < nc; i++) {
            uint32_t a = cn[i].segment_a, b = cn[i].segment_b;
            if (a >= n || b >= n) continue;
            ConnGeom g = conn_geom(&sg[a], &sg[b]);
            for (int ion = 0; ion < 7; ion++) {
                double ca = sg[a].extra_ions[ion], cb = sg[b].extra_ions[ion];
                double dphi = phi[a] - phi[b];
                float x = kZ[ion] * kFrtMv * dphi;
                double fl = kDiff[ion] * g.ac * dts / g.dist * (ca * bernoulli(-x) - cb * bernoulli(x));
                c.dext[(size_t)a * 7 + ion] -= fl / (g.va * kEvr);
                c.dext[(size_t)b * 7 + ion] += fl / (g.vb * kEvr);
            }
        }
        for (uint32_t i = 0; i < n; i++)
            for (int ion = 0; ion < 7; ion++) sg[i].extra_ions[ion] += c.dext[(size_t)i * 7 + ion];
    }

    for (uint32_t i = 0; i < nc; i++) {
        uint32_t a = cn[i].segment_a, b = cn[i].segment_b;
        if (a >= n || b >= n) continue;
        ConnGeom g = conn_geom(&sg[a], &sg[b]);
        bool intra = cn[i].is_intracellular != 0;
        if (intra) {
            for (int nt = 0; nt < 15; nt++) {
                float dC = sg[a].intra_nt[nt] - sg[b].intra_nt[nt];
                if (dC < 1e-12f && dC > -1e-12f) continue;
                float fl = nt_diffusion(nt) * g.ac * dC * dt / g.dist;
                c.dnt_in[(size_t)a * 15 + nt] -= fl / g.va;
                c.dnt_in[(size_t)b * 15 + nt] += fl / g.vb;
            }
        }
        for (int nt = 0; nt < 15; nt++) {
            float dC = sg[a].neurotransmitter[nt] - sg[b].neurotransmitter[nt];
            if (dC < 1e-12f && dC > -1e-12f) continue;
            float fl = nt_diffusion(nt) * g.ac * dC * dt / g.dist;
            c.dnt[(size_t)a * 15 + nt] -= fl / g.va;
            c.dnt[(size_t)b * 15 + nt] += fl / g.vb;
        }
        if (intra) {
            float d1 = sg[a].ATP_level - sg[b].ATP_level;
            if (d1 > 1e-12f || d1 < -1e-12f) {
                float fl = 0.14f * g.ac * d1 * dt / g.dist;
                c.datp[a] -= fl / g.va;
                c.datp[b] += fl / g.vb;
            }
            float d2 = sg[a].temperature - sg[b].temperature;
            if (d2 > 0.001f || d2 < -0.001f) {
                float fl = 0.0014f * g.ac * d2 * dt / g.dist;
                c.dtemp[a] -= fl / g.va;
                c.dtemp[b] += fl / g.vb;
            }
            float d3 = sg[a].transcription_signal - sg[b].transcription_signal;
            if (d3 > 1e-9f || d3 < -1e-9f) {
                float fl = 0.01f * g.ac * d3 * dt / g.dist;
                c.dtxn[a] -= fl / g.va;
                c.dtxn[b] += fl / g.vb;
            }
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        for (int ion = 0; ion < 7; ion++) {
            sg[i].extra_ions[ion] += (kBath[ion] - sg[i].extra_ions[ion]) * dt / 1.0f;
            if (sg[i].extra_ions[ion] < 1e-9) sg[i].extra_ions[ion] = 1e-9;
            if (sg[i].extra_ions[ion] > 1000.0) sg[i].extra_ions[ion] = 1000.0;
        }
        for (int nt = 0; nt < 15; nt++) {
            sg[i].intra_nt[nt] += c.dnt_in[(size_t)i * 15 + nt];
            sg[i].neurotransmitter[nt] += c.dnt[(size_t)i * 15 + nt];
        }
        sg[i].ATP_level += c.datp[i];
        sg[i].temperature += c.dtemp[i];
        sg[i].transcription_signal += c.dtxn[i];
    }
}

// ---- step 3 helpers (delay_buffer.h, kernel_dispatch.h:612-663, modifier_helpers.h:764-836) ----
#pragma pack(push, 1)
struct Delayed { uint64_t arrival; abi_float amount; uint8_t nt; uint8_t pad[3]; };
#pragma pack(pop)
static_assert(sizeof(Delayed) == 16, "delay slot");

static inline float xorshift_draw(uint32_t seed) {
    uint32_t x = seed;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return (float)(x & 0x7FFFFFFFu) / (float)0x7FFFFFFFu;
}
static inline uint8_t nt_of_receptor(uint8_t rt) {
    if (rt == T_GABAA || rt == T_GABAB) return 1;
    if (rt == T_GLYCINE) return 7;
    if (rt == T_5HT3) return 3;
    if (rt == T_NICOTINIC) return 4;
    if (rt == T_D1 || rt == T_D2) return 2;
    return 0;
}
static void rrp_recover(Synapse* y, const Segment* pre, float dt) {
    if (y->RRP_max <= 0.0f || y->RRP >= y->RRP_max) return;
    float ca = pre->intra_ions[2];
    float caf = ca / (0.0005f + ca);
    float camkii = 0, calcn = 0;
    for (int j = 0; j < pre->modifier_count; j++) {
        if (pre->modifiers[j].type == T_CAMKII) camkii = pre->modifiers[j].state[0];
        if (pre->modifiers[j].type == T_CALCINEURIN) calcn = pre->modifiers[j].state[0];
    }
    float prim = y->RRP_recovery + camkii * 0.02f + calcn * 0.01f;
    prim *= caf;
    float rec = (y->RRP_max - y->RRP) * prim * dt;
    y->RRP += rec;
    if (y->RRP > y->RRP_max) y->RRP = y->RRP_max;
}
static float rrp_release(Synapse* y, float amount, float p) {
    if (y->RRP_max <= 0.0f) return amount * p;
    float av = amount;
    if (av > y->RRP) av = y->RRP;
    float rel = av * p;
    y->RRP -= rel;
    if (y->RRP < 0) y->RRP = 0;
    return rel;
}
static void bouton_step(const Segment* pre, Synapse* y, float dt) {
    float vb = pre->V_m;
    float dens = 0.0f, mg = 0.0f, hg = 0.0f;
    uint8_t mc = pre->modifier_count;
    if (mc > 32) mc = 0;
    for (uint8_t j = 0; j < mc; j++)
        if (pre->modifiers[j].type == T_CAV21) {
            dens = pre->modifiers[j].density;
            float mi = cav21_minf(vb);
            float hi = cav21_hinf(vb, y->bouton_Ca * kMmToUm);
            mg = mi;
            hg = hi + (pre->modifiers[j].state[1] - hi) * exp(-dt / 300.0f);
        }
    float open = mg * mg * hg;
    double P = p_single(15.0f, 2.0f, kCbarCa);
    double flux = ghk_di(vb, y->bouton_Ca, pre->extra_ions[2]);
    double ia = P * flux * (dens * y->pre_area) * open;
    float ica = (float)(ia * kAtoPa);
    y->bouton_Ca -= ica * dt * 1000.0f / (2.0f * kFaraday * 0.1f);
    y->bouton_Ca += (0.0001f - y->bouton_Ca) * dt / 5.0f;
    if (y->bouton_Ca < 0.0f) y->bouton_Ca = 0.0f;
}
static void spine_ca(Synapse* y, float isyn, float dt) {
    float ia = isyn < 0.0f ? -isyn : isyn;
    float ica = ia * 0.1f;
    float influx = ica * dt * 1000.0f / (2.0f * kFaraday * 0.05f * 100.0f);
    y->spine_Ca += influx;
    y->spine_Ca += (0.0001f - y->spine_Ca) * dt / 50.0f;
    if (y->spine_Ca < 0.0f) y->spine_Ca = 0.0f;
}
static void spine_plast(Synapse* y, float dt) {
    float cu = y->spine_Ca * kMmToUm;
    float c4 = cu * cu * cu * cu;
    float ck_inf = c4 / (c4 + 1.0f + 1e-9f);
    y->spine_camkii += (ck_inf - y->spine_camkii) * dt / 2000.0f;
    if (y->spine_camkii < 0.0f) y->spine_camkii = 0.0f;
    if (y->spine_camkii > 1.0f) y->spine_camkii = 1.0f;
    float c2 = cu * cu;
    float cn_inf = c2 / (c2 + 0.05f * 0.05f + 1e-9f);
    y->spine_calcineurin += (cn_inf - y->spine_calcineurin) * dt / 300.0f;
    if (y->spine_calcineurin < 0.0f) y->spine_calcineurin = 0.0f;
    if (y->spine_calcineurin > 1.0f) y->spine_calcineurin = 1.0f;
}

// Gap-junction conductance (run_tick_common.h:79-110) and the slow Ca/CaM gate
// (run_tick_common.h:112-122), with the junction's first-order filter when it has a delay.
static float gap_g_eff(const Segment* pre, const Segment* post, float m_slow, float vj) {
    uint8_t pm = pre->modifier_count, qm = post->modifier_count;
    if (pm > 32) pm = 0;
    if (qm > 32) qm = 0;
    float g = 0.0f;
    for (int cx = 0; cx < 2; cx++) {
        const uint8_t ty = cx == 0 ? T_CX36 : T_CX43;
        const float gamma = cx == 0 ? 14.0f : 100.0f;
        float dens = 0.0f, f = 1.0f;
        for (uint8_t j = 0; j < pm; j++)
            if (pre->modifiers[j].type == ty) {
                dens = pre->modifiers[j].density;
                f *= pre->modifiers[j].state[0] * pre->modifiers[j].state[4];
            }
        for (uint8_t j = 0; j < qm; j++)
            if (post->modifiers[j].type == ty) f *= post->modifiers[j].state[0] * post->modifiers[j].state[4];
        if (cx == 1) {
            float a = vj < 0.0f ? -vj : vj;
            float fv = 1.0f / (1.0f + exp((a - 40.0f) / 5.0f));
            f *= fv;
        }
        if (f < 0.01f) f = 0.01f;
        g += dens * f * gamma;
    }
    return g * 1e-3f * (1.0f - m_slow);
}
static void gap_junction(Synapse* y, Segment* pre, Segment* post, float dt) {
    float vj = pre->V_m - post->V_m;
    float ca_pre = pre->intra_ions[2], ca_post = post->intra_ions[2];
    float m = y->depression;
    float avg = 0.5f * (ca_pre + ca_post);
    float cam = (avg * avg) / (avg * avg + 0.0003f * 0.0003f);
    m += (cam - m) * dt / 100.0f;
    if (m < 0.0f) m = 0.0f;
    if (m > 1.0f) m = 1.0f;
    y->depression = m;
    float raw = gap_g_eff(pre, post, m, vj) * vj;
    if (y->delay > 0.001f) {
        float flt = y->cleft_neurotransmitter[0];
        flt += (raw - flt) * dt / y->delay;
        y->cleft_neurotransmitter[0] = flt;
        pre->I_buffer -= flt;
        post->I_buffer += flt;
    } else {
        pre->I_buffer -= raw;
        post->I_buffer += raw;
    }
}

static void synapses_step(Ctx& c, Segment* sg, uint32_t n, Synapse* sy, uint32_t ns, uint64_t tick, float dt) {
    if (ns != c.delay_syn) {
        c.delay.assign((size_t)ns * 4 * 16, 0);
        c.delay_syn = ns;
    }
    if (ns > 0)
        for (uint32_t i = 0; i < n; i++)
            for (int k = 0; k < 15; k++) sg[i].neurotransmitter[k] = 0.0f;
    for (uint32_t i = 0; i < n; i++)
        for (uint8_t j = 0; j < sg[i].modifier_count; j++)
            if (sg[i].modifiers[j].type == T_NMDA) sg[i].neurotransmitter[0] = 1.0f;
    for (uint32_t i = 0; i < n; i++)
        for (uint8_t j = 0; j < sg[i].modifier_count; j++)
            if (sg[i].modifiers[j].type == T_5HT7) sg[i].neurotransmitter[3] = 0.0f;
    // glial segments (GLAST carriers), the first 256 in segment order (simulation_host.cpp:350-354)
    uint32_t glia[256], nglia = 0;
    for (uint32_t g = 0; g < n && nglia < 256; g++)
        for (uint8_t j = 0; j < sg[g].modifier_count; j++)
            if (sg[g].modifiers[j].type == T_GLAST) { glia[nglia++] = g; break; }

    for (uint32_t i = 0; i < ns; i++) {
        Synapse* y = &sy[i];
        if (y->pre_segment >= n || y->post_segment >= n) continue;
        Segment* pre = &sg[y->pre_segment];
        Segment* post = &sg[y->post_segment];
        Delayed* slots = reinterpret_cast<Delayed*>(&c.delay[(size_t)i * 64]);
        if (y->delay > 0.0f && !c.delay.empty()) {
            for (int s = 0; s < 4; s++) {
                if (slots[s].arrival == 0) continue;
                if (slots[s].arrival <= tick) {
                    y->cleft_neurotransmitter[slots[s].nt] += slots[s].amount;
                    slots[s].arrival = 0;
                }
            }
        }
        if (y->type == 1 || y->type == 2) {   // gap junction (simulation_host.cpp:372-400)
            gap_junction(y, pre, post, dt);
            continue;
        }
        if (!c.nuclei.empty() && pre->neuron_id < c.nuclei.size())
            protein_map(y, c.nuclei[pre->neuron_id].protein_level);
        rrp_recover(y, pre, dt);
        bouton_step(pre, y, dt);
        if (y->refractory == 0) {
            float cu = y->bouton_Ca * kMmToUm;
            float act = pow(cu, 4.0f) / (pow(cu, 4.0f) + pow(10.0f, 4.0f));
            float p = y->release_probability * act;
            for (int j = 0; j < y->pre_autoreceptor_count; j++) {   // simulation_host.cpp:430-436
                Modifier* ar = &y->pre_autoreceptors[j];
                ar->state[7] = y->pre_area;
                float f = release_modulator(ar, y->cleft_neurotransmitter, dt);
                p *= f;
            }
            uint32_t seed = y->id ^ ((uint32_t)tick) ^ (uint32_t)(int64_t)(pre->V_m * 1000.0f);
            if (xorshift_draw(seed) < p) {
                uint8_t nt = (y->post_receptor_count > 0) ? nt_of_receptor(y->post_receptors[0].type) : 0;
                float amt = rrp_release(y, y->release_amount, p);
                uint64_t dly = (y->delay <= 0.0f || dt <= 0.0f) ? 0 : (uint64_t)((y->delay / dt) + 0.5f);
                if (dly > 0) {
                    for (int s = 0; s < 4; s++)
                        if (slots[s].arrival == 0) {
                            slots[s].arrival = tick + dly;
                            slots[s].amount = amt;
                            slots[s].nt = nt;
                            break;
                        }
                } else {
                    y->cleft_neurotransmitter[nt] += amt;
                }
                y->refractory = (uint32_t)(20.0f / dt + 0.5f);
            }
        }
        for (int k = 0; k < 15; k++) post->neurotransmitter[k] += y->cleft_neurotransmitter[k];
        if (nglia > 0)   // glial capture of glutamate, glycine, GABA (simulation_host.cpp:472-490)
            for (int q = 0; q < 3; q++) {
                const int nt = q == 0 ? 0 : (q == 1 ? 7 : 1);
                if (y->cleft_neurotransmitter[nt] <= 0.0f) continue;
                float best = 1e30f;
                uint32_t g = glia[0];
                for (uint32_t k = 0; k < nglia; k++) {
                    float dz = fabsf(sg[glia[k]].position.z - post->position.z);
                    if (dz < best) { best = dz; g = glia[k]; }
                }
                float tau = (y->tau_clearance > 0.0f) ? y->tau_clearance : 5.0f;
                float cap = y->cleft_neurotransmitter[nt] * (1.0f - exp(-dt / tau));
                sg[g].neurotransmitter[nt] += cap;
                y->cleft_neurotransmitter[nt] -= cap;
            }
        bool my = has_myelin(post);
        float isyn = 0.0f;
        for (int j = 0; j < y->post_receptor_count; j++) {
            uint8_t rt = y->post_receptors[j].type;
            if (my && is_ionotropic(rt)) continue;
            y->post_receptors[j].state[7] = y->post_area;
            isyn += dispatch(rt, &y->post_receptors[j], post, dt);
        }
        post->I_buffer += isyn;
        spine_ca(y, isyn, dt);
        for (int k = 0; k < 15; k++) {
            if ((k == 0 || k == 7 || k == 1) && nglia > 0) continue;   // the glia took them up
            y->cleft_neurotransmitter[k] *= exp(-dt / y->tau_clearance);
            if (y->cleft_neurotransmitter[k] < 0.0f) y->cleft_neurotransmitter[k] = 0.0f;
        }
        float cu = y->spine_Ca * 1000.0f;
        float csq = cu * cu;
        float tau_theta = 300000.0f;
        y->facilitation += (csq - y->facilitation) * dt / tau_theta;
        if (y->facilitation < 0.0f) y->facilitation = 0.0f;
        y->depression = y->facilitation + 0.01f;
        spine_plast(y, dt);
        float ck = y->spine_camkii, cn = y->spine_calcineurin;
        bool inhib = false;
        for (int j = 0; j < y->post_receptor_count; j++) {
            uint8_t rt = y->post_receptors[j].type;
            if (rt == T_GABAA || rt == T_GABAB || rt == T_GLYCINE) { inhib = true; break; }
        }
        float dw = 0.0f;
        if (ck > 0.0f) dw += 0.0003f * ck * dt;
        float pp1 = 0.5f;
        for (uint8_t j = 0; j < post->modifier_count; j++)
            if (post->modifiers[j].type == T_PP1) pp1 = post->modifiers[j].state[3];
        if (cn > 0.0f) dw -= 0.0001f * cn * pp1 * dt;
        if (inhib) dw = -dw;
        y->weight += dw;
        if (std::isnan(y->weight)) y->weight = 0.5f;
        if (std::isnan(y->facilitation)) y->facilitation = 0.0f;
        if (std::isnan(y->depression)) y->depression = 0.01f;
        if (y->weight < 0.01f) y->weight = 0.01f;
        if (y->weight > 10.0f) y->weight = 10.0f;
        for (uint8_t j = 0; j < post->modifier_count; j++) {
            if (post->modifiers[j].type == T_ECB) {
                float ecb = post->modifiers[j].state[0];
                if (ecb > 0.0f) {
                    y->cleft_neurotransmitter[9] += ecb * 0.1f * dt;
                    post->modifiers[j].state[0] *= 0.9f;
                }
                break;
            }
        }
        if (y->refractory > 0) y->refractory--;
    }
}

struct Stim {
    CurrentStimulus* cur; uint32_t* ncur;
    VoltageClamp* vc; uint32_t* nvc;
    RampStimulus* rp; uint32_t* nrp;
    IonInjection* ii; uint32_t* nii;
    LigandInjection* li; uint32_t* nli;
};

static void run_tick(Ctx& c, Segment* sg, uint32_t n, IntraConnection* cn, uint32_t nc, Synapse* sy,
                     uint32_t ns, uint64_t* tick_count, float dt, Stim st) {
    uint64_t tick = (*tick_count)++;
    if (st.cur && st.ncur) {
        uint32_t w = 0;
        for (uint32_t r = 0; r < *st.ncur; r++)
            if (!(tick >= st.cur[r].start_tick + (uint64_t)(st.cur[r].duration_ms / dt))) st.cur[w++] = st.cur[r];
        *st.ncur = w;
    }
    if (st.vc && st.nvc) {
        uint32_t w = 0;
        for (uint32_t r = 0; r < *st.nvc; r++)
            if (!(tick >= st.vc[r].end_tick)) st.vc[w++] = st.vc[r];
        *st.nvc = w;
    }
    if (st.rp && st.nrp) {
        uint32_t w = 0;
        for (uint32_t r = 0; r < *st.nrp; r++)
            if (!(tick >= st.rp[r].end_tick)) st.rp[w++] = st.rp[r];
        *st.nrp = w;
    }
    // step 1 (simulation_host.cpp:233-307)
    for (uint32_t i = 0; i < n; i++) {
        Segment* s = &sg[i];
        double itot = 0.0;
        bool my = has_myelin(s);
        for (uint8_t j = 0; j < s->modifier_count; j++) {
            uint8_t t = s->modifiers[j].type;
            if (my && t != T_MYELIN && (is_ionotropic(t) || t == T_K2P || t == T_NALCN || t == T_KV31 ||
                                         t == T_KV11 || t == T_KV7 || t == T_KVDR || t == T_CAV21 ||
                                         t == T_HCN1 || t == T_SK || t == T_BK || t == T_KV42 ||
                                         t == T_NAP || t == T_NAP_TAD || t == T_HCN2 || t == 131 ||
                                         t == 132 || t == 154 || t == 157 || t == 205 || t == 206 ||
                                         t == 207 || t == 202 || t == 203 || t == 204 || t == 168 ||
                                         t == 200 || t == 201 || t == 191 || t == 159 || t == 96 ||
                                         t == 84 || t == 152 || t == 151 || t == 153 || t == 158 ||
                                         t == 121 || t == 209))
                continue;
            s->modifiers[j].state[7] = is_sphere(s) ? 4.0f * kPi * s->radius * s->radius
                                                    : 2.0f * kPi * s->radius * s->length;
            if (c.cable) {   // held concentrations: every channel sees the stored ones
                double in[BC_ION_MAX], ex[BC_ION_MAX];
                abi_float nt[15];
                const float atp = s->ATP_level;
                memcpy(in, s->intra_ions, sizeof in);
                memcpy(ex, s->extra_ions, sizeof ex);
                memcpy(nt, s->neurotransmitter, sizeof nt);
                itot += dispatch(t, &s->modifiers[j], s, dt);
                memcpy(s->intra_ions, in, sizeof in);
                memcpy(s->extra_ions, ex, sizeof ex);
                memcpy(s->neurotransmitter, nt, sizeof nt);
                s->ATP_level = atp;
            } else {
                itot += dispatch(t, &s->modifiers[j], s, dt);
            }
        }
        if (st.cur && st.ncur)
            for (uint32_t k = 0; k < *st.ncur; k++) {
                const CurrentStimulus& q = st.cur[k];
                if (q.segment_id == s->id) {
                    uint64_t dur = (uint64_t)(q.duration_ms / dt);
                    if (tick >= q.start_tick && tick < q.start_tick + dur) itot += q.amplitude_pA;
                }
            }
        if (st.rp && st.nrp)
            for (uint32_t k = 0; k < *st.nrp; k++) {
                const RampStimulus& q = st.rp[k];
                if (q.segment_id == s->id && tick >= q.start_tick && tick < q.end_tick) {
                    float frac = (q.end_tick > q.start_tick)
                                     ? (float)(tick - q.start_tick) / (float)(q.end_tick - q.start_tick)
                                     : 1.0f;
                    float amp = q.start_pA + (q.end_pA - q.start_pA) * frac;
                    itot += amp;
                }
            }
        if (st.ii && st.nii)
            for (uint32_t k = 0; k < *st.nii; k++) {
                const IonInjection& q = st.ii[k];
                if (q.segment_id == s->id && q.ion_idx >= 0 && q.ion_idx < 7 && !c.cable)
                    (q.extracellular ? s->extra_ions : s->intra_ions)[q.ion_idx] += q.amount_mM;
            }
        if (st.li && st.nli)
            for (uint32_t k = 0; k < *st.nli; k++) {
                const LigandInjection& q = st.li[k];
                if (q.segment_id == s->id && q.nt_idx >= 0 && q.nt_idx < 15) {
                    if (q.intracellular) s->intra_nt[q.nt_idx] += q.amount;
                    else s->neurotransmitter[q.nt_idx] += q.amount;
                }
            }
        s->I_buffer = itot;
    }
    if (st.ii && st.nii) *st.nii = 0;
    if (st.li && st.nli) *st.nli = 0;

    diffusion(c, sg, n, cn, nc, dt);
    synapses_step(c, sg, n, sy, ns, tick, dt);

    // step 4 (simulation_host.cpp:596-649)
    if (c.spikes.size() < n) c.spikes.resize(n, 0);
    for (uint32_t i = 0; i < n; i++) {
        Segment* s = &sg[i];
        if (std::isnan(s->I_buffer)) s->I_buffer = 0.0f;
        if (s->I_buffer > 10000.0f) s->I_buffer = 10000.0f;
        if (s->I_buffer < -10000.0f) s->I_buffer = -10000.0f;
        float cm = capacitance(s);
        float prev = s->V_m;
        float dV = s->I_buffer / cm * dt;
        if (dV > 50.0f) dV = 50.0f;
        if (dV < -50.0f) dV = -50.0f;
        s->V_m += dV;
        if (prev <= -40.0f && s->V_m > -40.0f) c.spikes[i]++;
        s->I_buffer = 0.0f;
        if (std::isnan(s->position.x)) s->position.x = 0.0f;
        if (std::isnan(s->position.y)) s->position.y = 50.0f;
        if (std::isnan(s->position.z)) s->position.z = 0.0f;
        if (std::isnan(s->radius)) s->radius = 5.0f;
        if (std::isnan(s->length)) s->length = 10.0f;
        if (st.vc && st.nvc)
            for (uint32_t v = 0; v < *st.nvc; v++) {
                const VoltageClamp& q = st.vc[v];
                if (q.segment_id == s->id && tick >= q.start_tick && tick < q.end_tick) s->V_m = q.target_mV;
            }
        if (s->V_m != s->V_m) s->V_m = -65.0f;
        else {
            if (s->V_m < -120.0f) s->V_m = -120.0f;
            if (s->V_m > 150.0f) s->V_m = 150.0f;
        }
    }
}

}  // namespace orc

// ---- FP64 restatement of the tick of passive networks (no modifiers, synapses or stimuli):
// step 2.5 (include/diffusion_kernel.h:41-364) and step 4 (src/simulation_host.cpp:596-649)
// with every quantity in double, for error reports of the reference's mixed precision. With no
// modifiers I_buffer is 0, so step 4 only clamps V_m; transmitters stay 0.
namespace orc {
struct F64 {
    uint32_t n = 0, nc = 0;
    std::vector<double> V, in, ex, atp, temp, txn, area, vol, cap, rad, len;
    std::vector<uint32_t> ca, cb;
    std::vector<uint8_t> intra;
    Ctx cells;   // residual rule and cell ids
};
static inline double bern64(double x) {
    if (fabs(x) < 1e-3) return 1.0 - 0.5 * x;
    if (x > 50.0) return 0.0;
    if (x < -50.0) return -x;
    return x / (exp(x) - 1.0);
}
static void f64_tick(F64& s, double dt) {
    const uint32_t n = s.n, nc = s.nc;
    Ctx& c = s.cells;
    const double pi = (double)kPi, evr = (double)kEvr;
    if (nc > 0) {
        std::vector<double> sigma(n), diag(n), phi(n, 0.0), acc(n), g(nc), ac(nc), dist(nc);
        for (uint32_t i = 0; i < n; i++) {
            double sm = 0.0;
            for (int k = 0; k < 7; k++) sm += (double)kZ[k] * kZ[k] * kDiff[k] * s.ex[(size_t)i * 7 + k];
            sigma[i] = (double)kSigmaE * sm;
            diag[i] = sigma[i] * s.area[i];
        }
        for (uint32_t e = 0; e < nc; e++) {
            uint32_t a = s.ca[e], b = s.cb[e];
            double rmin = std::min(s.rad[a], s.rad[b]);
            ac[e] = std::max(pi * rmin * rmin, 1e-6);
            dist[e] = std::max((s.len[a] + s.len[b]) * 0.5, 0.1);
            g[e] = 0.5 * (sigma[a] + sigma[b]) * ac[e] / dist[e];
            diag[a] += g[e];
            diag[b] += g[e];
        }
        std::vector<double> res(c.jac_on ? c.ncomp : 0, 0.0);
        std::vector<uint8_t> done(c.jac_on ? c.ncomp : 0, 0);
        for (int it = 0; it < (c.jac_on ? c.jac_max : 50); it++) {
            std::fill(acc.begin(), acc.end(), 0.0);
            for (uint32_t e = 0; e < nc; e++) {
                if (g[e] <= 0.0) continue;
                acc[s.ca[e]] += g[e] * phi[s.cb[e]];
                acc[s.cb[e]] += g[e] * phi[s.ca[e]];
            }
            for (uint32_t i = 0; i < n; i++) {
                if (c.jac_on && done[c.comp[i]]) continue;
                double v = diag[i] > 1e-9 ? acc[i] / diag[i] : 0.0;
                v = std::min(10.0, std::max(-10.0, v));
                if (c.jac_on) res[c.comp[i]] = std::max(res[c.comp[i]], fabs(v - phi[i]));
                phi[i] = v;
            }
            if (c.jac_on && cells_settle(c, res, done)) break;
        }
        const double dts = dt / 5.0;
        std::vector<double> dacc((size_t)n * 7), gdiag(n), inpd(n), gax(nc), vold(n), vnew(n), off(n);
        for (int sub = 0; sub < 5; sub++) {
            std::fill(dacc.begin(), dacc.end(), 0.0);
            std::fill(gdiag.begin(), gdiag.end(), 0.0);
            std::fill(inpd.begin(), inpd.end(), 0.0);
            for (uint32_t e = 0; e < nc; e++) {
                if (!s.intra[e]) continue;
                uint32_t a = s.ca[e], b = s.cb[e];
                double sg = 0.0;
                for (int k = 0; k < 7; k++)
                    sg += (double)kZ[k] * kZ[k] * kDiff[k] * 0.5 * (s.in[(size_t)a * 7 + k] + s.in[(size_t)b * 7 + k]);
                gax[e] = (double)kSigmaE * sg * ac[e] / dist[e];
                gdiag[a] += gax[e];
                gdiag[b] += gax[e];
                double dV = (s.V[a] + phi[a]) - (s.V[b] + phi[b]);
                double inp = 0.0;
                for (int k = 0; k < 7; k++) {
                    double x = (double)kZ[k] * kFrtMv * dV;
                    double fl = (double)kDiff[k] * ac[e] * dts / dist[e] *
                                (s.in[(size_t)a * 7 + k] * bern64(-x) - s.in[(size_t)b * 7 + k] * bern64(x));
                    dacc[(size_t)a * 7 + k] -= fl / std::max(s.vol[a], 1e-6);
                    dacc[(size_t)b * 7 + k] += fl / std::max(s.vol[b], 1e-6);
                    inp += (double)kZ[k] * kChargeToPa * fl / dts;
                }
                inpd[a] += inp - gax[e] * dV;
                inpd[b] -= inp - gax[e] * dV;
            }
            vold = s.V;
            vnew = s.V;
            std::fill(res.begin(), res.end(), 0.0);
            std::fill(done.begin(), done.end(), 0);
            for (int it = 0; it < (c.jac_on ? c.jac_max : 50); it++) {
                std::fill(off.begin(), off.end(), 0.0);
                for (uint32_t e = 0; e < nc; e++) {
                    if (!s.intra[e]) continue;
                    off[s.ca[e]] += gax[e] * vnew[s.cb[e]];
                    off[s.cb[e]] += gax[e] * vnew[s.ca[e]];
                }
                for (uint32_t i = 0; i < n; i++) {
                    if (c.jac_on && done[c.comp[i]]) continue;
                    double cdt = std::max(s.cap[i] / dts, 1e-9);
                    double v = (cdt * vold[i] + inpd[i] + off[i]) / (cdt + gdiag[i]);
                    if (c.jac_on) res[c.comp[i]] = std::max(res[c.comp[i]], fabs(v - vnew[i]));
                    vnew[i] = v;
                }
                if (c.jac_on && cells_settle(c, res, done)) break;
            }
            s.V = vnew;
            for (size_t q = 0; q < dacc.size(); q++) s.in[q] += dacc[q];
        }
        std::vector<double> dext((size_t)n * 7);
        for (int sub = 0; sub < 5; sub++) {
            std::fill(dext.begin(), dext.end(), 0.0);
            for (uint32_t e = 0; e < nc; e++) {
                uint32_t a = s.ca[e], b = s.cb[e];
                for (int k = 0; k < 7; k++) {
                    double x = (double)kZ[k] * kFrtMv * (phi[a] - phi[b]);
                    double fl = (double)kDiff[k] * ac[e] * dts / dist[e] *
                                (s.ex[(size_t)a * 7 + k] * bern64(-x) - s.ex[(size_t)b * 7 + k] * bern64(x));
                    dext[(size_t)a * 7 + k] -= fl / (std::max(s.vol[a], 1e-6) * evr);
                    dext[(size_t)b * 7 + k] += fl / (std::max(s.vol[b], 1e-6) * evr);
                }
            }
            for (size_t q = 0; q < dext.size(); q++) s.ex[q] += dext[q];
        }
        std::vector<double> datp(n, 0.0), dtemp(n, 0.0), dtxn(n, 0.0);
        for (uint32_t e = 0; e < nc; e++) {
            if (!s.intra[e]) continue;
            uint32_t a = s.ca[e], b = s.cb[e];
            double va = std::max(s.vol[a], 1e-6), vb = std::max(s.vol[b], 1e-6);
            double d1 = s.atp[a] - s.atp[b], d2 = s.temp[a] - s.temp[b], d3 = s.txn[a] - s.txn[b];
            if (fabs(d1) > 1e-12) { double fl = 0.14 * ac[e] * d1 * dt / dist[e]; datp[a] -= fl / va; datp[b] += fl / vb; }
            if (fabs(d2) > 0.001) { double fl = 0.0014 * ac[e] * d2 * dt / dist[e]; dtemp[a] -= fl / va; dtemp[b] += fl / vb; }
            if (fabs(d3) > 1e-9) { double fl = 0.01 * ac[e] * d3 * dt / dist[e]; dtxn[a] -= fl / va; dtxn[b] += fl / vb; }
        }
        for (uint32_t i = 0; i < n; i++) {
            for (int k = 0; k < 7; k++) {
                double& x = s.ex[(size_t)i * 7 + k];
                x += ((double)kBath[k] - x) * dt;
                x = std::min(1000.0, std::max(1e-9, x));
            }
            s.atp[i] += datp[i];
            s.temp[i] += dtemp[i];
            s.txn[i] += dtxn[i];
        }
    }
    for (uint32_t i = 0; i < n; i++) s.V[i] = std::min(150.0, std::max(-120.0, s.V[i]));
}
}  // namespace orc

// ---- FP64 restatement of the cable-mode tick (an extension of the backend with no reference
// counterpart): held ion concentrations and ATP, phi_e = 0. Step 1 with Nav1.2, Kv-DR, K2P,
// NALCN and HCN1 (hh_biological.cpp:193-277, 2763-2798, 3063-3085, 3609-3641;
// modifier_helpers.h:1442-1484) and current pulses/ramps (simulation_host.cpp:247-300); step 2.5
// as five substeps of axial Scharfetter-Gummel flux + chord term and 50 Jacobi sweeps of V
// (diffusion_kernel.h:151-238) with each connection's conductance and flux times its scale;
// step 4 (simulation_host.cpp:596-649). Every quantity is double. Its central differences are
// the reference gradient of the backend's adjoint.
namespace orc {
struct Cab64 {
    struct Ch { uint8_t type; double dens, st[7]; };
    uint32_t n = 0, nc = 0;
    uint64_t tick = 0;
    std::vector<double> V, cap, area, vol, rad, len, in, ex, pkc, pka, camp, scale;
    std::vector<uint32_t> ca, cb, ids, ch_off;
    std::vector<uint8_t> intra;
    std::vector<Ch> ch;
    std::vector<CurrentStimulus> cur;
    std::vector<RampStimulus> ramp;
    std::vector<VoltageClamp> vc;
};
static inline double ghk64(double V, double ci, double co, bool di) {   // run_tick_common.h:274-312
    const double z = di ? 2.0 : 1.0, rt = (double)kGas * kTemp;
    const double vn = z * kFaraday * V * 0.001 / rt, k = z * z * kFaraday / rt;
    if (fabs(vn) < 1e-6) return z * kFaraday * (ci - co);
    if (vn > 20.0) return k * kFaraday * V * 0.001 * ci;
    if (vn < -20.0) return k * kFaraday * V * 0.001 * co;
    const double e = exp(-vn);
    return k * kFaraday * V * 0.001 * (ci - co * e) / (1.0 - e);
}
static inline double vtrap64(double x, double y) {
    if (fabs(x / y) < 1e-6) return y * (1.0 - x / (2.0 * y));
    return x / (1.0 - exp(-x / y));
}
static inline double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }
// One channel of segment i: updates its gates and returns its current (pA).
static double cab_channel(Cab64& s, uint32_t i, Cab64::Ch& c, double dt) {
    const double V = s.V[i], A = s.area[i], N = c.dens * A;
    const double* ci = &s.in[(size_t)i * 7];
    const double* co = &s.ex[(size_t)i * 7];
    double* st = c.st;
    switch (c.type) {
        case T_NAV12: {
            const double am = 0.182 * vtrap64(V + 23.0, 6.0), bm = 0.124 * vtrap64(-(V + 23.0), 6.0);
            const double ah = 0.024 * vtrap64(V + 41.5, 5.0), bh = 0.0091 * vtrap64(-(V + 66.5), 5.0);
            const double minf = am / (am + bm), hinf = 1.0 / (1.0 + exp((V + 63.5) / 6.2));
            st[0] = clamp01(minf + (st[0] - minf) * exp(-dt * (am + bm)));
            st[1] = clamp01(hinf + (st[1] - hinf) * exp(-dt * (ah + bh)));
            double op = st[0] * st[0] * st[0] * st[1];
            op /= 1.0 + (st[5] > 0.0 ? st[5] : 0.0) / 10.0;
            op *= 1.0 - 0.7 * (s.pkc[i] + s.pka[i]) * 0.5;
            if (op < 0.0) op = 0.0;
            return -(p_single(20.8f, 1.0f, kCbarNa) * ghk64(V, ci[0], co[0], false) * N * op * kAtoPa);
        }
        case T_KVDR: {
            const double vh = st[1] != 0.0 ? st[1] : -30.0, kk = st[2] > 0.0 ? st[2] : 6.0;
            double tau = st[4];
            if (!(tau > 0.0)) tau = std::min(200.0, std::max(0.3, 14.6 * exp(-(V - vh) / 30.1)));
            const double mi = 1.0 / (1.0 + exp(-(V - vh) / (kk + 1e-6)));
            st[0] = clamp01(st[0] + (mi - st[0]) * dt / (tau + 1e-6));
            return -(p_single(12.0f, 1.0f, kCbarK) * ghk64(V, ci[1], co[1], false) * N * st[0] * kAtoPa);
        }
        case T_K2P: {
            const double P = p_single(st[0] > 0.0 ? (float)st[0] : 100.0f, 1.0f, kCbarK);
            const double po = st[1] > 0.0 ? st[1] : 0.012;
            return -(P * ghk64(V, ci[1], co[1], false) * N * po * kAtoPa);
        }
        case T_NALCN: {
            double pna, pk, pca;
            p_nalcn(st[0] > 0.0 ? (float)st[0] : 27.0f, &pna, &pk, &pca);
            const double ina = pna * ghk64(V, ci[0], co[0], false) * N * kAtoPa;
            const double ik = pk * ghk64(V, ci[1], co[1], false) * N * kAtoPa;
            const double ica = pca * ghk64(V, ci[2], co[2], true) * N * kAtoPa;
            return -(ina - ik + ica);
        }
        case T_HCN1: {
            const double vh = st[1] != 0.0 ? st[1] : -69.0, kk = st[2] > 0.0 ? st[2] : 6.0;
            const double to = st[4] > 0.0 ? st[4] : 84.0, tc = st[5] > 0.0 ? st[5] : 300.0;
            const double ve = vh + 15.0 * s.camp[i] / (s.camp[i] + 0.5);
            const double mi = 1.0 / (1.0 + exp((V - ve) / kk));
            const double tau = V < ve ? to * exp((V - ve) / 15.0) : tc;
            st[0] = clamp01(st[0] + (mi - st[0]) * dt / tau);
            double pna, pk, pca;
            p_partial(st[3] > 0.0 ? (float)st[3] : 0.84f, 0.2f, 0.8f, 0.0f, &pna, &pk, &pca);
            return -(pna * ghk64(V, ci[0], co[0], false) * N * st[0] * kAtoPa +
                     pk * ghk64(V, ci[1], co[1], false) * N * st[0] * kAtoPa);
        }
        default: return 0.0;
    }
}
static void cab_tick(Cab64& s, double dt) {
    const uint32_t n = s.n, nc = s.nc;
    const uint64_t tick = s.tick;
    std::vector<double> ib(n, 0.0);
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t j = s.ch_off[i]; j < s.ch_off[i + 1]; j++) ib[i] += cab_channel(s, i, s.ch[j], dt);
        for (const CurrentStimulus& q : s.cur)
            if (q.segment_id == s.ids[i] && tick >= q.start_tick &&
                tick < q.start_tick + (uint64_t)(q.duration_ms / (float)dt))
                ib[i] += q.amplitude_pA;
        for (const RampStimulus& q : s.ramp)
            if (q.segment_id == s.ids[i] && tick >= q.start_tick && tick < q.end_tick) {
                const double f = q.end_tick > q.start_tick
                                     ? (double)(tick - q.start_tick) / (double)(q.end_tick - q.start_tick) : 1.0;
                ib[i] += q.start_pA + (q.end_pA - q.start_pA) * f;
            }
    }
    const double dts = dt / 5.0, pi = (double)kPi;
    std::vector<double> G(nc), inpd(n), gdiag(n), vold(n), off(n);
    for (uint32_t e = 0; e < nc; e++) {
        const uint32_t a = s.ca[e], b = s.cb[e];
        double sg = 0.0;
        for (int k = 0; k < 7; k++)
            sg += (double)kZ[k] * kZ[k] * kDiff[k] * 0.5 * (s.in[(size_t)a * 7 + k] + s.in[(size_t)b * 7 + k]);
        const double rmin = std::min(s.rad[a], s.rad[b]);
        const double ac = std::max(pi * rmin * rmin, 1e-6), dist = std::max((s.len[a] + s.len[b]) * 0.5, 0.1);
        G[e] = s.intra[e] ? (double)kSigmaE * sg * ac / dist * s.scale[e] : 0.0;
    }
    std::fill(gdiag.begin(), gdiag.end(), 0.0);
    for (uint32_t e = 0; e < nc; e++) {
        gdiag[s.ca[e]] += G[e];
        gdiag[s.cb[e]] += G[e];
    }
    for (int sub = 0; sub < 5; sub++) {
        std::fill(inpd.begin(), inpd.end(), 0.0);
        for (uint32_t e = 0; e < nc; e++) {
            if (!s.intra[e]) continue;
            const uint32_t a = s.ca[e], b = s.cb[e];
            const double rmin = std::min(s.rad[a], s.rad[b]);
            const double ac = std::max(pi * rmin * rmin, 1e-6), dist = std::max((s.len[a] + s.len[b]) * 0.5, 0.1);
            const double dV = s.V[a] - s.V[b];
            double inp = 0.0;
            for (int k = 0; k < 7; k++) {
                const double x = (double)kZ[k] * kFrtMv * dV;
                const double fl = (double)kDiff[k] * ac * dts / dist *
                                  (s.in[(size_t)a * 7 + k] * bern64(-x) - s.in[(size_t)b * 7 + k] * bern64(x));
                inp += (double)kZ[k] * kChargeToPa * fl * s.scale[e] / dts;
            }
            inpd[a] += inp - G[e] * dV;
            inpd[b] -= inp - G[e] * dV;
        }
        vold = s.V;
        for (int it = 0; it < 50; it++) {
            std::fill(off.begin(), off.end(), 0.0);
            for (uint32_t e = 0; e < nc; e++) {
                off[s.ca[e]] += G[e] * s.V[s.cb[e]];
                off[s.cb[e]] += G[e] * s.V[s.ca[e]];
            }
            for (uint32_t i = 0; i < n; i++) {
                const double cdt = std::max(s.cap[i] / dts, 1e-9);
                s.V[i] = (cdt * vold[i] + inpd[i] + off[i]) / (cdt + gdiag[i]);
            }
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        double x = ib[i];
        if (x != x) x = 0.0;
        x = std::min(10000.0, std::max(-10000.0, x));
        double dV = x / s.cap[i] * dt;
        dV = std::min(50.0, std::max(-50.0, dV));
        double v = s.V[i] + dV;
        for (const VoltageClamp& q : s.vc)
            if (q.segment_id == s.ids[i] && tick >= q.start_tick && tick < q.end_tick) v = q.target_mV;
        if (v != v) v = -65.0;
        s.V[i] = std::min(150.0, std::max(-120.0, v));
    }
    s.tick++;
}
}  // namespace orc

extern "C" {

void* oracle_create(void) { return new orc::Ctx(); }
void oracle_destroy(void* c) { delete static_cast<orc::Ctx*>(c); }

// Same argument meaning as core_run_tick (src/simulation_host.cpp:660); the delay ring,
// spike counters and scratch live in ctx instead of function-local statics.
void oracle_run_tick(void* ctx, Segment* segments, uint32_t num_segments, IntraConnection* connections,
                     uint32_t num_connections, Synapse* synapses, uint32_t num_synapses, uint64_t* tick_count,
                     float dt, CurrentStimulus* stimuli, uint32_t* stimulus_count, VoltageClamp* vclamps,
                     uint32_t* vclamp_count, RampStimulus* ramps, uint32_t* ramp_count, IonInjection* ion_inj,
                     uint32_t* ion_inj_count, LigandInjection* lig_inj, uint32_t* lig_inj_count) {
    orc::Stim st{stimuli, stimulus_count, vclamps, vclamp_count, ramps, ramp_count,
                 ion_inj, ion_inj_count, lig_inj, lig_inj_count};
    orc::run_tick(*static_cast<orc::Ctx*>(ctx), segments, num_segments, connections, num_connections,
                  synapses, num_synapses, tick_count, dt, st);
}

// Spike counter of step 4 (prev <= -40 mV < V_m), as core_enable_spike_count.
void oracle_spike_counts(void* ctx, uint32_t* out, uint32_t n) {
    orc::Ctx* c = static_cast<orc::Ctx*>(ctx);
    for (uint32_t i = 0; i < n; i++) out[i] = i < c->spikes.size() ? c->spikes[i] : 0;
}

// The 4 x 16-byte delay slots of every synapse (layout of DelayedRelease).
uint64_t oracle_delay_bytes(void* ctx, uint8_t* out, uint64_t cap) {
    orc::Ctx* c = static_cast<orc::Ctx*>(ctx);
    uint64_t n = c->delay.size() < cap ? c->delay.size() : cap;
    if (out && n) memcpy(out, c->delay.data(), n);
    return c->delay.size();
}

uint64_t oracle_unsupported_count(void) { return orc::g_unsupported; }
void oracle_reset_unsupported(void) { orc::g_unsupported = 0; }

// Single-call entry for per-modifier known-answer tests: dispatches one modifier.
double oracle_modifier_current(uint8_t type, Modifier* m, Segment* s, float dt) {
    return orc::dispatch(type, m, s, dt);
}

float oracle_xorshift(uint32_t seed) { return orc::xorshift_draw(seed); }

// Residual-terminated Jacobi for both potential solves (tol <= 0: the reference's fixed 50).
void oracle_set_jacobi(void* ctx, double tol, int max_sweeps) {
    orc::Ctx* c = static_cast<orc::Ctx*>(ctx);
    c->jac_on = tol > 0.0;
    c->jac_tol = tol;
    c->jac_max = max_sweeps > 0 ? max_sweeps : 50;
}

// Cable mode and per-connection coupling scale (see orc::Ctx::cable); scale may be NULL.
void oracle_set_protein_mapping(void* ctx, const Nucleus* nuclei, uint32_t count) {
    orc::Ctx* c = static_cast<orc::Ctx*>(ctx);
    if (nuclei && count) c->nuclei.assign(nuclei, nuclei + count);
    else c->nuclei.clear();
}

void oracle_set_cable(void* ctx, int on, const float* scale, uint32_t nc) {
    orc::Ctx* c = static_cast<orc::Ctx*>(ctx);
    c->cable = on != 0;
    if (scale) c->scale.assign(scale, scale + nc);
    else c->scale.clear();
}

// FP64 restatement for passive networks (see orc::f64_tick).
void* oracle_fp64_create(const Segment* sg, uint32_t n, const IntraConnection* cn, uint32_t nc) {
    orc::F64* s = new orc::F64();
    s->n = n;
    s->nc = nc;
    s->V.resize(n); s->in.resize((size_t)n * 7); s->ex.resize((size_t)n * 7);
    s->atp.resize(n); s->temp.resize(n); s->txn.resize(n); s->area.resize(n); s->vol.resize(n);
    s->cap.resize(n); s->rad.resize(n); s->len.resize(n);
    const double pi = (double)orc::kPi;
    for (uint32_t i = 0; i < n; i++) {
        const Segment& x = sg[i];
        const bool sph = orc::is_sphere(&x);
        const double r = x.radius, l = x.length;
        s->V[i] = x.V_m;
        for (int k = 0; k < 7; k++) { s->in[(size_t)i * 7 + k] = x.intra_ions[k]; s->ex[(size_t)i * 7 + k] = x.extra_ions[k]; }
        s->atp[i] = x.ATP_level; s->temp[i] = x.temperature; s->txn[i] = x.transcription_signal;
        s->rad[i] = r; s->len[i] = l;
        s->area[i] = sph ? 4.0 * pi * r * r : 2.0 * pi * r * l;
        s->vol[i] = sph ? (4.0 / 3.0) * pi * r * r * r : pi * r * r * l;
        s->cap[i] = std::max(s->area[i] * 0.01, 1e-4);
    }
    for (uint32_t e = 0; e < nc; e++) {
        if (cn[e].segment_a >= n || cn[e].segment_b >= n) continue;
        s->ca.push_back(cn[e].segment_a);
        s->cb.push_back(cn[e].segment_b);
        s->intra.push_back(cn[e].is_intracellular != 0);
    }
    s->nc = (uint32_t)s->ca.size();
    orc::cells(s->cells, n, cn, nc);
    return s;
}
void oracle_fp64_set_jacobi(void* h, double tol, int max_sweeps) {
    oracle_set_jacobi(&static_cast<orc::F64*>(h)->cells, tol, max_sweeps);
}
void oracle_fp64_run_ticks(void* h, float dt, uint32_t ticks) {
    for (uint32_t t = 0; t < ticks; t++) orc::f64_tick(*static_cast<orc::F64*>(h), dt);
}
// V (n), intra (n x 7), extra (n x 7), ATP, temperature (n each).
void oracle_fp64_read(void* h, double* V, double* in, double* ex, double* atp, double* temp) {
    const orc::F64& s = *static_cast<orc::F64*>(h);
    memcpy(V, s.V.data(), s.n * 8);
    memcpy(in, s.in.data(), s.in.size() * 8);
    memcpy(ex, s.ex.data(), s.ex.size() * 8);
    memcpy(atp, s.atp.data(), s.n * 8);
    memcpy(temp, s.temp.data(), s.n * 8);
}
void oracle_fp64_destroy(void* h) { delete static_cast<orc::F64*>(h); }

// FP64 cable-mode restatement (see orc::cab_tick). scale: per connection (NULL: all 1).
void* oracle_cable64_create(const Segment* sg, uint32_t n, const IntraConnection* cn, uint32_t nc,
                            const float* scale) {
    orc::Cab64* s = new orc::Cab64();
    const double pi = (double)orc::kPi;
    s->n = n;
    s->ch_off.push_back(0);
    for (uint32_t i = 0; i < n; i++) {
        const Segment& x = sg[i];
        const bool sph = orc::is_sphere(&x);
        const double r = x.radius, l = x.length;
        s->ids.push_back(x.id);
        s->V.push_back(x.V_m);
        s->rad.push_back(r);
        s->len.push_back(l);
        s->area.push_back(sph ? 4.0 * pi * r * r : 2.0 * pi * r * l);
        s->vol.push_back(sph ? (4.0 / 3.0) * pi * r * r * r : pi * r * r * l);
        s->cap.push_back(std::max(s->area.back() * 0.01, 1e-4));
        for (int k = 0; k < 7; k++) { s->in.push_back(x.intra_ions[k]); s->ex.push_back(x.extra_ions[k]); }
        s->pkc.push_back(orc::first_state_of(&x, orc::T_PKC, 0, 0.0f));
        s->pka.push_back(orc::first_state_of(&x, orc::T_CAMP_PKA, 1, 0.0f));
        s->camp.push_back(orc::first_state_of(&x, orc::T_CAMP_PKA, 0, 0.1f));
        for (uint32_t j = 0; j < x.modifier_count; j++) {
            orc::Cab64::Ch c;
            c.type = x.modifiers[j].type;
            c.dens = x.modifiers[j].density;
            for (int q = 0; q < 7; q++) c.st[q] = x.modifiers[j].state[q];
            s->ch.push_back(c);
        }
        s->ch_off.push_back((uint32_t)s->ch.size());
    }
    for (uint32_t e = 0; e < nc; e++) {
        if (cn[e].segment_a >= n || cn[e].segment_b >= n) continue;
        s->ca.push_back(cn[e].segment_a);
        s->cb.push_back(cn[e].segment_b);
        s->intra.push_back(cn[e].is_intracellular != 0);
        s->scale.push_back(scale ? scale[e] : 1.0);
    }
    s->nc = (uint32_t)s->ca.size();
    return s;
}
void oracle_cable64_set_stimuli(void* h, const CurrentStimulus* cur, uint32_t ncur, const RampStimulus* ramp,
                                uint32_t nramp, const VoltageClamp* vc, uint32_t nvc) {
    orc::Cab64& s = *static_cast<orc::Cab64*>(h);
    s.cur.assign(cur, cur + ncur);
    s.ramp.assign(ramp, ramp + nramp);
    s.vc.assign(vc, vc + nvc);
}
// Runs `ticks` ticks; returns mean over ticks and cells of (V(soma[c]) - target[t][c])^2 and,
// when trace is not NULL, V(soma[c]) after each tick (ticks x ncells).
double oracle_cable64_loss(void* h, float dt, uint32_t ticks, const uint32_t* soma, uint32_t ncells,
                           const float* target, double* trace) {
    orc::Cab64& s = *static_cast<orc::Cab64*>(h);
    double acc = 0.0;
    for (uint32_t t = 0; t < ticks; t++) {
        orc::cab_tick(s, dt);
        for (uint32_t c = 0; c < ncells; c++) {
            const double v = s.V[soma[c]];
            if (trace) trace[(size_t)t * ncells + c] = v;
            const double e = v - (double)target[(size_t)t * ncells + c];
            acc += e * e;
        }
    }
    return acc / ((double)ticks * ncells);
}
void oracle_cable64_destroy(void* h) { delete static_cast<orc::Cab64*>(h); }
float oracle_bernoulli(float x) { return orc::bernoulli(x); }
// Array forms for checking device evaluations: host glibc ::exp (the function the reference
// calls), and the Bernoulli function of diffusion_kernel.h:108-113.
void oracle_exp_n(const double* x, double* y, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) y[i] = ::exp(x[i]);
}
void oracle_bernoulli_n(const double* x, double* y, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) y[i] = orc::bernoulli((float)x[i]);
}

}  // extern "C"
