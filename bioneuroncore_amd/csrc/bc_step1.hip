// bc_step1.hip — step 1: k_modifiers, one lane per segment over a range of the device order
// (modifier slots in order, then the segment's stimuli).
#include "bc_tick.h"

namespace bc {

// ------------------------------------------------------------------ step 1
// kNt: some segment carries a modifier that reads or takes up transmitters (SF_NTREAD); kExt:
// some segment carries a modifier outside the core set. The variants without keep those
// registers free.
template <bool kNt, bool kExt>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_modifiers(Net d, float dt,
                                                                                          uint32_t first, uint32_t last) {
    exp_table_stage();
    const uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= last) return;
    const uint64_t tick = *d.tick;
    const size_t P = d.npad;
    SegRegs s;
    s.V = d.V[i];
#pragma unroll
    for (int k = 0; k < 7; k++) {
        s.in[k] = d.in[k * P + i];
        s.ex[k] = d.ex[k * P + i];
    }
    s.atp = d.atp[i];
    s.area = d.area[i];
    s.vol = d.vol[i];
    s.pkc = d.pkc[i];
    s.pka = d.pka[i];
    s.camp = d.camp[i];
    s.camkii = d.camkii[i];
    s.camkii01 = d.camkii01[i];
    seg_consts_init(s);
    ghk_cache_init(s);
    const uint8_t sfl = d.sflags[i];
    if (kNt && (sfl & SF_NTREAD)) {   // receptors read transmitters 0-11, transporters 0, 1 and 7
#pragma unroll
        for (int k = 0; k < 15; k++) s.nt[k] = d.nt[k * P + i];
    }
    if (kExt && (sfl & SF_ER)) s.er = d.er[i];
    if (kExt && (sfl & SF_TT)) {
        s.temp = d.temp[i];
        s.txn = d.txn[i];
    }
    const bool my = sfl & SF_MYELIN;
    const uint32_t cnt = d.mcount[i];
    double itot = 0.0;
    // Slot j + 1's type, active-vars index and density load while slot j runs (none of them
    // changes during step 1), so each slot waits on one round trip to memory, for its states.
    uint8_t tn = 0, avn = 0;
    double dn = 0.0;
    if (cnt > 0) {
        tn = d.mtype[i];
        avn = d.mav[i];
        dn = d.mdens[i];
    }
    for (uint32_t j = 0; j < cnt; j++) {
        const uint8_t t = tn;
        ModSlot m;
        m.dens = dn;
        m.av = avn;
        if (j + 1 < cnt) {
            tn = d.mtype[(j + 1) * P + i];
            avn = d.mav[(j + 1) * P + i];
            dn = d.mdens[(j + 1) * P + i];
        }
        if (my && t != 4 && myelin_blocked(t)) continue;
        m.st = d.mstate + (size_t)j * 7 * P + i;
        m.stride = P;
        m.a7 = s.area;
        if (kExt && t == 144) {   // Cav1.4 reads the first CaBP4's state[0] as it stands now
            s.cabp4 = 0.0f;
            for (uint32_t q = 0; q < cnt; q++)
                if (d.mtype[q * P + i] == 145) {
                    s.cabp4 = d.mstate[(size_t)q * 7 * P + i];
                    break;
                }
        }
        if (kExt && t == 127) {   // IP3R reads the first PLC/IP3/DAG modifier's state[0] as it stands now
            s.ip3 = 0.0f;
            for (uint32_t q = 0; q < cnt; q++)
                if (d.mtype[q * P + i] == 136) {
                    s.ip3 = d.mstate[(size_t)q * 7 * P + i];
                    break;
                }
        }
        if (kExt && t == 136) {   // PLC sums the Gq receptors' state[0] (52-62, 188) as they stand now
            float gq = 0.0f;
            for (uint32_t q = 0; q < cnt; q++) {
                const uint8_t tq = d.mtype[q * P + i];
                if ((tq >= 52 && tq <= 62) || tq == 188) gq += d.mstate[(size_t)q * 7 * P + i];
            }
            s.gq = gq;
        }
        if (kExt && (t == 90 || t == 91 || t == 92)) {   // the last mTORC1 / MAPK / NF-kB state as it stands now
            float xa = 0.0f, xb = 0.0f;
            for (uint32_t q = 0; q < cnt; q++) {
                const uint8_t tq = d.mtype[q * P + i];
                if (t == 90 && tq == 89) xa = d.mstate[(size_t)q * 7 * P + i];
                if (t == 91 && tq == 87) xa = d.mstate[(size_t)q * 7 * P + i];
                if (t == 92 && tq == 87) xa = d.mstate[((size_t)q * 7 + 3) * P + i];
                if (t == 92 && tq == 88) xb = d.mstate[((size_t)q * 7 + 3) * P + i];
            }
            s.xa = xa;
            s.xb = xb;
        }
        if (kExt && (t == 86 || t == 137)) {   // eCB synthesis and PKC read the first PLC/IP3/DAG modifier's state[1]
            s.dag = 0.0f;
            for (uint32_t q = 0; q < cnt; q++)
                if (d.mtype[q * P + i] == 136) {
                    s.dag = d.mstate[((size_t)q * 7 + 1) * P + i];
                    break;
                }
        }
        if (kExt && (t == 75 || t == 76)) s.xa = msn_scale(d.mtype, d.mstate, P, i, cnt, t);
        if (kExt && t == 6) s.len = d.length[i];
        if (kExt && t == 111) s.sid = d.segid[i];
        itot += dispatch<kNt, kExt>(t, m, s, dt);
        if (kExt && t == 120) {   // cryptochrome boosts the first light sensor's state[0] (hh_biological.cpp cryptochrome_magnetic_current)
            const double rp = m.get(0);
            for (uint32_t q = 0; q < cnt; q++)
                if (d.mtype[q * P + i] == 95) {
                    double* c = d.mstate + (size_t)q * 7 * P + i;
                    double v = *c * (1.0f + 0.3f * rp);
                    if (v > 1.0f) v = 1.0f;
                    *c = v;
                    break;
                }
        }
        // The cascades other modifiers and step 3 read: CaMKII (the last one's state[0], and
        // state[0] + state[1]), calcineurin (the last one's state[0]), PP1 (the last one's
        // state[3]) and PKC (the first one's state[0]). The per-segment planes hold those values
        // as they stand, for the readers later in this loop and in step 3 and for the next tick.
        if (kExt && (t == 82 || t == 83 || t == 113)) {
            bool last = true;
            for (uint32_t q = j + 1; q < cnt; q++)
                if (d.mtype[q * P + i] == t) last = false;
            if (last) {
                if (t == 82) {
                    const double a = m.get(0), b = m.get(1);
                    s.camkii = a;
                    s.camkii01 = a + b;
                    d.camkii[i] = s.camkii;
                    d.camkii01[i] = s.camkii01;
                } else if (t == 83) {
                    d.calcn[i] = m.get(0);
                } else {
                    d.pp1[i] = m.get(3);
                }
            }
        }
        if (kExt && t == 137) {
            bool first = true;
            for (uint32_t q = 0; q < j; q++)
                if (d.mtype[q * P + i] == 137) first = false;
            if (first) {
                s.pkc = m.get(0);
                d.pkc[i] = s.pkc;
            }
        }
        if (kExt && t == 87) {   // nuclear ERK raises every CREB's pCREB on the segment (hh_biological.cpp mapk_erk_current)
            float dts = dt;
            dts *= m.dens;
            const double nuc = m.get(3);
            for (uint32_t q = 0; q < cnt; q++)
                if (d.mtype[q * P + i] == 79) {
                    double* c = d.mstate + (size_t)q * 7 * P + i;
                    double v = *c + nuc * 0.01f * dts / 1000.0f;
                    if (v > 1.0f) v = 1.0f;
                    *c = v;
                }
        }
        if (d.cable) {   // concentrations held: every channel sees the stored ones
#pragma unroll
            for (int k = 0; k < 7; k++) {
                s.in[k] = d.in[k * P + i];
                s.ex[k] = d.ex[k * P + i];
            }
            s.atp = d.atp[i];
            if (kNt && (sfl & SF_NTREAD)) {
                s.nt[0] = d.nt[0 * P + i];
                s.nt[1] = d.nt[1 * P + i];
                s.nt[7] = d.nt[7 * P + i];
            }
        }
    }
    if (kNt && (sfl & SF_NTREAD) && !d.cable) {   // the transporters' uptake (before injections land)
        d.nt[0 * P + i] = s.nt[0];
        d.nt[1 * P + i] = s.nt[1];
        d.nt[7 * P + i] = s.nt[7];
    }
    // stimuli (simulation_host.cpp:247-300): this segment's list, in the reference's order
    const uint32_t sh = d.stim.any ? d.stim.head[i] : ~0u;
    const uint64_t wtick = tick + d.stim.tick_shift;   // time windows; injections use the tick
    if (sh != ~0u) {
        const uint32_t cnt = d.stim.ent[sh];
        for (uint32_t e = 1; e <= cnt; e++) {
            const uint32_t en = d.stim.ent[sh + e];
            const uint32_t kind = en >> 28, x = en & 0x0FFFFFFFu;
            if (kind == SK_CUR) {
                const CurrentStimulus q = d.stim.cur[x];
                uint64_t dur = (uint64_t)(q.duration_ms / dt);
                if (wtick >= q.start_tick && wtick < q.start_tick + dur) itot += q.amplitude_pA;
            } else if (kind == SK_RAMP) {
                const RampStimulus q = d.stim.ramp[x];
                if (wtick >= q.start_tick && wtick < q.end_tick) {
                    float frac = (q.end_tick > q.start_tick)
                                     ? (float)(wtick - q.start_tick) / (float)(q.end_tick - q.start_tick)
                                     : 1.0f;
                    float amp = q.start_pA + (q.end_pA - q.start_pA) * frac;
                    itot += amp;
                }
            } else if (kind == SK_ION) {
                const IonInjection q = d.stim.ion[x];
                if (tick == d.stim.inj_tick && q.ion_idx >= 0 && q.ion_idx < 7) {
                    // register arrays: index with compile-time constants only
#pragma unroll
                    for (int k = 0; k < 7; k++)
                        if (k == q.ion_idx) {
                            if (q.extracellular) s.ex[k] += q.amount_mM;
                            else s.in[k] += q.amount_mM;
                        }
                }
            } else if (kind == SK_LIG) {
                const LigandInjection q = d.stim.lig[x];
                if (tick == d.stim.inj_tick && q.nt_idx >= 0 && q.nt_idx < 15) {
                    if (q.intracellular) d.nt_in[q.nt_idx * P + i] += q.amount;
                    else d.nt[q.nt_idx * P + i] += q.amount;
                }
            }
        }
    }
    d.V[i] = s.V;
    if (!d.cable) {
#pragma unroll
        for (int k = 0; k < 7; k++) {
            d.in[k * P + i] = s.in[k];
            d.ex[k * P + i] = s.ex[k];
        }
        d.atp[i] = s.atp;
        if (kExt && (sfl & SF_ER)) d.er[i] = s.er;
        if (kExt && (sfl & SF_TT)) {
            d.temp[i] = s.temp;
            d.txn[i] = s.txn;
        }
    }
    d.ibuf[i] = itot;
}

template __global__ void k_modifiers<false, false>(Net, float, uint32_t, uint32_t);
template __global__ void k_modifiers<true, false>(Net, float, uint32_t, uint32_t);
template __global__ void k_modifiers<false, true>(Net, float, uint32_t, uint32_t);
template __global__ void k_modifiers<true, true>(Net, float, uint32_t, uint32_t);

}  // namespace bc
