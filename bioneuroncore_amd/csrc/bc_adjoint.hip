// bc_adjoint.hip — reverse-mode derivative of the cable-mode tick (an extension; the reference
// has no adjoint). The differentiated tick: step 1 with the gating ODEs and GHK currents of
// Nav1.2 (142), Kv-DR (123), K2P (107), NALCN (133) and HCN1 (129) at held ion
// concentrations; step 2.5 in cable mode (phi_e = 0, five substeps of axial SG flux + chord
// term and 50 Jacobi sweeps of V); step 4. Parameters: each modifier's density (g_max) and each
// connection's conductance scale. Local partial derivatives of the channels come from
// restating them over dual numbers; the tick-level reverse pass (transposed Jacobi sweeps,
// flux and chord terms, step-4 clamps) is written out below.
#pragma once

namespace bc {
namespace adj {

// ---- dual numbers: value + N partial derivatives (double) ----
template <int N>
struct Dn {
    double v, d[N];
};
template <int N>
BC_HD Dn<N> cst(double v) {
    Dn<N> r;
    r.v = v;
    for (int k = 0; k < N; k++) r.d[k] = 0.0;
    return r;
}
template <int N>
BC_HD Dn<N> var(double v, int slot) {
    Dn<N> r = cst<N>(v);
    r.d[slot] = 1.0;
    return r;
}
template <int N> BC_HD Dn<N> operator+(Dn<N> a, Dn<N> b) { for (int k = 0; k < N; k++) a.d[k] += b.d[k]; a.v += b.v; return a; }
template <int N> BC_HD Dn<N> operator-(Dn<N> a, Dn<N> b) { for (int k = 0; k < N; k++) a.d[k] -= b.d[k]; a.v -= b.v; return a; }
template <int N> BC_HD Dn<N> operator-(Dn<N> a) { for (int k = 0; k < N; k++) a.d[k] = -a.d[k]; a.v = -a.v; return a; }
template <int N> BC_HD Dn<N> operator+(Dn<N> a, double b) { a.v += b; return a; }
template <int N> BC_HD Dn<N> operator+(double b, Dn<N> a) { a.v += b; return a; }
template <int N> BC_HD Dn<N> operator-(Dn<N> a, double b) { a.v -= b; return a; }
template <int N> BC_HD Dn<N> operator-(double b, Dn<N> a) { return cst<N>(b) - a; }
template <int N> BC_HD Dn<N> operator*(Dn<N> a, Dn<N> b) {
    Dn<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N> BC_HD Dn<N> operator*(Dn<N> a, double b) { a.v *= b; for (int k = 0; k < N; k++) a.d[k] *= b; return a; }
template <int N> BC_HD Dn<N> operator*(double b, Dn<N> a) { return a * b; }
template <int N> BC_HD Dn<N> operator/(Dn<N> a, Dn<N> b) {
    Dn<N> r;
    r.v = a.v / b.v;
    for (int k = 0; k < N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
template <int N> BC_HD Dn<N> operator/(Dn<N> a, double b) { return a * (1.0 / b); }
template <int N> BC_HD Dn<N> operator/(double a, Dn<N> b) { return cst<N>(a) / b; }
template <int N> BC_HD Dn<N> dexp(Dn<N> a) {
    Dn<N> r;
    r.v = bc::exp_t(a.v, ExpTabGlobal{});
    for (int k = 0; k < N; k++) r.d[k] = r.v * a.d[k];
    return r;
}
template <int N> BC_HD Dn<N> clamp01(Dn<N> a) {   // the gates' clamp; derivative 0 where it binds
    if (a.v < 0.0) return cst<N>(0.0);
    if (a.v > 1.0) return cst<N>(1.0);
    return a;
}

// GHK flux (run_tick_common.h:274-312) of a monovalent / divalent ion at held concentrations.
template <int N>
BC_HD Dn<N> ghk(Dn<N> V, double ci, double co, bool di) {
    const double rt = (double)kGas * kTemp;
    Dn<N> vn = (di ? 2.0 : 1.0) * kFaraday * V * 0.001 / rt;
    const double k = (di ? 4.0 : 1.0) * kFaraday / rt;
    if (fabs(vn.v) < 1e-6) return cst<N>((di ? 2.0 : 1.0) * kFaraday * (ci - co));
    if (vn.v > 20.0) return k * kFaraday * V * 0.001 * ci;
    if (vn.v < -20.0) return k * kFaraday * V * 0.001 * co;
    Dn<N> e = dexp(-vn);
    return k * kFaraday * V * 0.001 * (ci - co * e) / (1.0 - e);
}
template <int N>
BC_HD Dn<N> vtrapd(Dn<N> x, double y) {
    if (fabs(x.v / y) < 1e-6) return y * (1.0 - x / (2.0 * y));
    return x / (1.0 - dexp(-x / y));
}

// One channel at one segment: value and partials used by the reverse pass.
struct ChanAd {
    double r;          // returned current (the tick adds it to I_buffer)
    double dr[3];      // d r / d (V, gate0', gate1')
    double r1;         // d r / d density
    double dg[2][2];   // d gate_q' / d V, d gate_q' / d gate_q
    int ng;            // gates
};
struct SegAd {         // held segment quantities the channels read
    double in[3], ex[3];
    double area, pkc, pka, camp;
};
BC_HD bool adjoint_supported(uint8_t t) {
    return t == M_NAV12 || t == M_KVDR || t == M_K2P || t == M_NALCN || t == M_HCN1;
}

// The current of each type as a function of (V, new gates), per unit density (x dens outside).
// Each follows the corresponding forward function of bc_device.h (same formulas and branches).
BC_HD Dn<3> current1(uint8_t t, Dn<3> V, Dn<3> g0, Dn<3> g1, const double* st, const SegAd& s) {
    if (t == M_NAV12) {   // ch_nav with the Nav1.2 constants
        Dn<3> op = g0 * g0 * g0 * g1;
        double ttx = st[5] > 0.0 ? st[5] : 0.0;
        op = op / (1.0 + ttx / 10.0);
        op = op * (1.0 - 0.7 * (s.pkc + s.pka) * 0.5);
        if (op.v < 0.0) op = cst<3>(0.0);
        const double P = p_single(20.8f, 1.0f, kCbarNa);
        return -(P * ghk(V, s.in[0], s.ex[0], false) * s.area * op * (double)kAtoPa);
    }
    if (t == M_KVDR) {
        const double P = p_single(12.0f, 1.0f, kCbarK);
        return -(P * ghk(V, s.in[1], s.ex[1], false) * s.area * g0 * (double)kAtoPa);
    }
    if (t == M_K2P) {
        const double gamma = st[0] > 0.0 ? st[0] : 100.0;
        const double po = st[1] > 0.0 ? st[1] : 0.012;
        const double P = st[0] > 0.0 ? p_single((float)gamma, 1.0f, kCbarK) : p_single(100.0f, 1.0f, kCbarK);
        return -(P * ghk(V, s.in[1], s.ex[1], false) * s.area * po * (double)kAtoPa);
    }
    if (t == M_NALCN) {   // -(ina - ik + ica), as the forward function
        const double gamma = st[0] > 0.0 ? st[0] : 27.0;
        const P3 pp = st[0] > 0.0 ? p_nalcn3((float)gamma) : p_nalcn3(27.0f);
        Dn<3> ina = pp.na * ghk(V, s.in[0], s.ex[0], false) * s.area * (double)kAtoPa;
        Dn<3> ik = pp.k * ghk(V, s.in[1], s.ex[1], false) * s.area * (double)kAtoPa;
        Dn<3> ica = pp.ca * ghk(V, s.in[2], s.ex[2], true) * s.area * (double)kAtoPa;
        return -(ina - ik + ica);
    }
    if (t == M_HCN1) {
        const float gamma = st[3] > 0.0 ? (float)st[3] : 0.84f;
        const P3 pp = st[3] > 0.0 ? p_partial3(gamma, 0.2f, 0.8f, 0.0f) : p_partial3(0.84f, 0.2f, 0.8f, 0.0f);
        Dn<3> in = pp.na * ghk(V, s.in[0], s.ex[0], false) * s.area * g0 * (double)kAtoPa;
        Dn<3> ik = pp.k * ghk(V, s.in[1], s.ex[1], false) * s.area * g0 * (double)kAtoPa;
        return -(in + ik);
    }
    return cst<3>(0.0);
}

// Gate updates of each type: gate_q' as a function of (V, gate_q) (slots 0: V, 1: gate).
BC_HD int gates_of(uint8_t t) { return t == M_NAV12 ? 2 : (t == M_KVDR || t == M_HCN1) ? 1 : 0; }
BC_HD Dn<2> gate_update(uint8_t t, int q, double Vv, double gv, const double* st, const SegAd& s, float dt) {
    Dn<2> V = var<2>(Vv, 0), g = var<2>(gv, 1);
    if (t == M_NAV12) {
        const double am_off = 23.0, am_k = 6.0, ah_off = 41.5, bh_off = 66.5, h_off = 63.5;
        if (q == 0) {
            Dn<2> am = 0.182 * vtrapd(V + am_off, am_k), bm = 0.124 * vtrapd(-(V + am_off), am_k);
            Dn<2> minf = am / (am + bm), mtau = 1.0 / (am + bm);
            return clamp01(minf + (g - minf) * dexp(-(double)dt / mtau));
        }
        Dn<2> hinf = 1.0 / (1.0 + dexp((V + h_off) / 6.2));
        Dn<2> ah = 0.024 * vtrapd(V + ah_off, 5.0), bh = 0.0091 * vtrapd(-(V + bh_off), 5.0);
        Dn<2> htau = 1.0 / (ah + bh);
        return clamp01(hinf + (g - hinf) * dexp(-(double)dt / htau));
    }
    if (t == M_KVDR) {
        const double vh = st[1] != 0.0 ? st[1] : -30.0, kk = st[2] > 0.0 ? st[2] : 6.0;
        Dn<2> tau;
        if (st[4] > 0.0) tau = cst<2>(st[4]);
        else {
            tau = 14.6 * dexp(-(V - vh) / 30.1);
            if (tau.v < 0.3) tau = cst<2>(0.3);
            if (tau.v > 200.0) tau = cst<2>(200.0);
        }
        Dn<2> mi = 1.0 / (1.0 + dexp(-(V - vh) / (kk + 1e-6)));
        return clamp01(g + (mi - g) * (double)dt / (tau + 1e-6));
    }
    if (t == M_HCN1) {
        const double vh = st[1] != 0.0 ? st[1] : -69.0, k = st[2] > 0.0 ? st[2] : 6.0;
        const double t_open = st[4] > 0.0 ? st[4] : 84.0, t_close = st[5] > 0.0 ? st[5] : 300.0;
        const double ve = vh + 15.0 * s.camp / (s.camp + 0.5);
        Dn<2> mi = 1.0 / (1.0 + dexp((V - ve) / k));
        Dn<2> tau = V.v < ve ? t_open * dexp((V - ve) / 15.0) : cst<2>(t_close);
        return clamp01(g + (mi - g) * (double)dt / tau);
    }
    return cst<2>(gv);
}

BC_HD ChanAd channel_ad(uint8_t t, double V, const double* st, const SegAd& s, double dens, float dt) {
    ChanAd c;
    c.ng = gates_of(t);
    double gnew[2] = {st[0], st[1]};
    for (int q = 0; q < 2; q++) {
        c.dg[q][0] = 0.0;
        c.dg[q][1] = 1.0;
    }
    for (int q = 0; q < c.ng; q++) {
        Dn<2> gq = gate_update(t, q, V, st[q], st, s, dt);
        gnew[q] = gq.v;
        c.dg[q][0] = gq.d[0];
        c.dg[q][1] = gq.d[1];
    }
    Dn<3> r1 = current1(t, var<3>(V, 0), var<3>(gnew[0], 1), var<3>(gnew[1], 2), st, s);
    c.r1 = r1.v;
    c.r = dens * r1.v;
    for (int k = 0; k < 3; k++) c.dr[k] = dens * r1.d[k];
    return c;
}

BC_DEV double bern_d(float x) {   // dB/dx of the SG Bernoulli function (the forward's branches)
    if (fabsf(x) < 1e-3f) return -0.5;
    if (x > 50.0f) return 0.0;
    if (x < -50.0f) return -1.0;
    double e = bc::exp_t((double)x, ExpTabGlobal{});
    double em1 = e - 1.0;
    return (em1 - x * e) / (em1 * em1);
}

}  // namespace adj

// Reverse-mode state of the adjoint pass.
struct AdjNet {
    const double* recV;      // recorded pre-tick V of the tick being reversed [n]
    const double* recG;      // recorded pre-tick gate states [slot][2][n]
    const double* recIb;     // recorded I_buffer of the tick being reversed [n]
    double* lamIb;           // d loss / d I_buffer of the tick being reversed [n]
    double* lamV;            // adjoint of V after the tick in, before the tick out [n]
    double* lamG;            // adjoint of the gate states [slot][2][n]
    double* gdens;           // d loss / d density [slot][n] (accumulated)
    double* gscale;          // d loss / d connection scale [device connection] (accumulated)
    const float* xrec;       // Jacobi iterates the forward recorded for this tick [5][51][npad],
                             // or NULL: the kernel runs the substeps again
};

// One tick of the reverse pass through steps 4 and 2.5 for one tile of cells (lane = segment =
// connection index).
template <int TS>
__global__ __launch_bounds__(TS) void k_adjoint_tick(Net d, TileSet ts, AdjNet a, uint64_t tick, float dt) {
    __shared__ float xs[51 * TS];      // Jacobi iterates x^0 .. x^50 of one substep
    __shared__ double sV0[TS];         // V before the tick (start of the first substep)
    __shared__ float sVf[4][TS];       // V at the start of substeps 1-4
    __shared__ double sLam[TS];        // adjoint exchanged between lanes
    __shared__ double sMu2[2 * TS];    // lam / D of the transposed sweeps, two planes
    __shared__ float muh[50 * TS];     // lam / D of every sweep of the substep
    __shared__ double sDinp[TS];
    __shared__ float sG[TS], sCt[TS], sDV[TS], sInp1[TS];
    __shared__ uint32_t sAB[TS];
    __shared__ uint8_t sIntra[TS];
    __shared__ uint16_t sEnt[2 * TS];
    const uint32_t tile = blockIdx.x;
    const uint32_t s0 = ts.seg0[tile], nseg = ts.nseg[tile], c0 = ts.conn0[tile], nconn = ts.nconn[tile];
    const uint32_t t = threadIdx.x;
    const bool act = t < nseg, cact = t < nconn;
    const uint32_t g = s0 + t;
    const size_t P = d.npad;
    const uint32_t E0 = d.csr_off[s0];
    const uint32_t nent = d.csr_off[s0 + nseg] - E0;
    uint32_t e0 = 0, e1 = 0;
    if (act) { e0 = d.csr_off[g] - E0; e1 = d.csr_off[g + 1] - E0; }

    // ---- step 4 inputs: pre-tick V, this tick's I_buffer (recorded), clamps ----
    double V = act ? a.recV[g] : 0.0;
    const double ibuf = act ? a.recIb[g] : 0.0;
    const float cap = act ? d.cap[g] : 1.0f;
    bool vclamped = false;
    const uint32_t sh = (act && d.stim.any) ? d.stim.head[g] : ~0u;
    if (sh != ~0u) {   // an active voltage clamp fixes V after step 4
        const uint32_t sc = d.stim.ent[sh];
        for (uint32_t e = 1; e <= sc; e++) {
            const uint32_t en = d.stim.ent[sh + e];
            if ((en >> 28) != SK_VCLAMP) continue;
            const VoltageClamp q = d.stim.vc[en & 0x0FFFFFFFu];
            if (tick + d.stim.tick_shift >= q.start_tick && tick + d.stim.tick_shift < q.end_tick) vclamped = true;
        }
    }

    // ---- step 2.5 in cable mode: axial terms per connection, V per substep ----
    for (uint32_t e = t; e < nent; e += TS) {
        uint32_t en = d.csr_ent[E0 + e];
        uint32_t c = en >> 1;
        sEnt[e] = (uint16_t)((c << 2) | ((d.conn_intra[c0 + c] != 0) << 1) | (en & 1));
    }
    uint32_t ca_ = 0, cb_ = 0;
    float G = 0.0f, csc = 1.0f, kin[7];
    bool cintra = false;
    const float dts = dt / 5.0f;
    if (cact) {
        const uint32_t ab = d.conn_ab[c0 + t];
        sAB[t] = ab;
        ca_ = ab & 0xFFFF;
        cb_ = ab >> 16;
        cintra = d.conn_intra[c0 + t] != 0;
        float ra = d.radius[s0 + ca_], rb = d.radius[s0 + cb_];
        float rmin = ra < rb ? ra : rb;
        float cac = kPi * rmin * rmin;
        if (cac < 1e-6f) cac = 1e-6f;
        float cdist = (d.length[s0 + ca_] + d.length[s0 + cb_]) * 0.5f;
        if (cdist < 0.1f) cdist = 0.1f;
        float sgi = 0.0f;
        for (int k = 0; k < 7; k++) {
            float cbar = 0.5f * (d.in[k * P + s0 + ca_] + d.in[k * P + s0 + cb_]);
            sgi += kZ[k] * kZ[k] * kDiff[k] * cbar;
        }
        G = kSigmaE * sgi * cac / cdist;
        csc = d.conn_scale ? d.conn_scale[c0 + t] : 1.0f;
        for (int k = 0; k < 7; k++) kin[k] = kDiff[k] * cac * dts / cdist;
    }
    const float gax = G * csc;
    float cdt = cap / dts;
    if (cdt < 1e-9f) cdt = 1e-9f;
    if (cact) sG[t] = cintra ? gax : 0.0f;
    if (act) sV0[t] = V;
    __syncthreads();
    float gdiag = 0.0f;
    if (act)
        for (uint32_t e = e0; e < e1; e++) {
            const uint32_t en = sEnt[e];
            if ((en >> 1) & 1) gdiag += sG[en >> 2];
        }
    const float den = cdt + gdiag;
    // V of segment i at the start of substep sub, as the forward's dV reads it (the pre-tick V
    // in double for the first substep, the float iterate after)
    auto vstart = [&](int sub, uint32_t i) -> double { return sub == 0 ? sV0[i] : (double)sVf[sub - 1][i]; };
    // axial terms of substep sub (sCt for the forward, sDinp/sDV/sInp1 for the reverse)
    auto conn_terms = [&](int sub) {
        if (cact) {
            // inp as the forward sums it (each flux times the scale), inp1 unscaled
            float inp = 0.0f, inp1 = 0.0f, dV = 0.0f;
            double dinp = 0.0;   // d inp1 / d dV in double: inp - G dV nearly cancels
            if (cintra) {
                dV = vstart(sub, ca_) - vstart(sub, cb_);
                for (int k = 0; k < 7; k++) {
                    float ca = d.in[k * P + s0 + ca_], cb = d.in[k * P + s0 + cb_];
                    float x = kZ[k] * kFrtMv * dV;
                    float fl = kin[k] * (ca * bernoulli(-x, ExpTabGlobal{}) - cb * bernoulli(x, ExpTabGlobal{}));
                    const double dfl = (double)kin[k] * (-ca * adj::bern_d(-x) - cb * adj::bern_d(x)) * kZ[k] * kFrtMv;
                    inp1 += kZ[k] * kChargeToPa * fl / dts;
                    inp += kZ[k] * kChargeToPa * (fl * csc) / dts;
                    dinp += (double)kZ[k] * kChargeToPa * dfl / dts;
                }
            }
            sInp1[t] = inp1;
            sDinp[t] = dinp * csc;
            sDV[t] = dV;
            sCt[t] = inp - gax * dV;
        }
    };
    // the forward's 50 sweeps of substep sub into xs (x^0 .. x^50); ends with a barrier
    auto sweeps = [&](int sub) {
        conn_terms(sub);
        __syncthreads();
        float inpd = 0.0f, vold = 0.0f;
        if (act) {
            for (uint32_t e = e0; e < e1; e++) {
                const uint32_t en = sEnt[e];
                if (!((en >> 1) & 1)) continue;
                if (en & 1) inpd -= sCt[en >> 2]; else inpd += sCt[en >> 2];
            }
            vold = (float)vstart(sub, t);
            xs[t] = vold;
        }
        __syncthreads();
        const float base = cdt * vold + inpd;
        for (int it = 0; it < 50; it++) {
            if (act) {
                float off = 0.0f;
                for (uint32_t e = e0; e < e1; e++) {
                    const uint32_t en = sEnt[e];
                    if (!((en >> 1) & 1)) continue;
                    const uint32_t c = en >> 2, ab = sAB[c];
                    const uint32_t o = (en & 1) ? (ab & 0xFFFF) : (ab >> 16);
                    off += sG[c] * xs[it * TS + o];
                }
                xs[(it + 1) * TS + t] = (base + off) / den;
            }
            __syncthreads();
        }
    };
    // the recorded iterates of substep sub into xs; ends with a barrier
    auto load_iterates = [&](int sub) {   // all 51 loads in flight before the first store
        float r[51];
#pragma unroll
        for (int it = 0; it <= 50; it++) r[it] = act ? a.xrec[((size_t)sub * 51 + it) * P + g] : 0.0f;
#pragma unroll
        for (int it = 0; it <= 50; it++) xs[it * TS + t] = r[it];
        __syncthreads();
    };
    if (a.xrec) {   // substep start values from the recording
        for (int sub = 1; sub < 5; sub++)
            if (act) sVf[sub - 1][t] = a.xrec[((size_t)(sub - 1) * 51 + 50) * P + g];
        if (act) V = a.xrec[((size_t)4 * 51 + 50) * P + g];
        __syncthreads();
    } else {        // ... or from running the substeps again
        for (int sub = 0; sub < 5; sub++) {
            sweeps(sub);
            if (sub < 4) {
                if (act) sVf[sub][t] = xs[50 * TS + t];
            } else if (act) {
                V = xs[50 * TS + t];   // V after step 2.5
            }
            __syncthreads();
        }
    }

    // ---- reverse: step 4 ----
    double lam = act ? a.lamV[g] : 0.0;
    double lam_ibuf = 0.0;
    if (act) {
        double ib = ibuf;
        bool ib_c = false;
        if (ib != ib) { ib = 0.0; ib_c = true; }
        if (ib > 10000.0f) { ib = 10000.0f; ib_c = true; }
        if (ib < -10000.0f) { ib = -10000.0f; ib_c = true; }
        float dVs = ib / cap * dt;
        bool dv_c = dVs > 50.0f || dVs < -50.0f;
        if (dv_c) dVs = dVs > 0 ? 50.0f : -50.0f;
        double Vn = V + dVs;
        bool v_c = Vn < -120.0f || Vn > 150.0f || Vn != Vn || vclamped;
        if (v_c) lam = 0.0;
        lam_ibuf = (ib_c || dv_c) ? 0.0 : lam * (double)dt / (double)cap;
    }

    // ---- reverse: step 2.5, substeps in reverse ----
    // The first kRegNb intracellular neighbours of each lane stay in registers for the
    // transposed sweeps (g and lane; padding reads the lane itself with g = 0), the rest is
    // walked from the incidence list.
    constexpr int kRegNb = 4;
    double nbg[kRegNb];
    uint32_t nbo[kRegNb];
#pragma unroll
    for (int q = 0; q < kRegNb; q++) { nbg[q] = 0.0; nbo[q] = t; }
    uint32_t e_rest = e1;
    {
        int nn = 0;
        if (act)
            for (uint32_t e = e0; e < e1; e++) {
                const uint32_t en = sEnt[e];
                if (!((en >> 1) & 1)) continue;
                if (nn == kRegNb) { e_rest = e; break; }
                const uint32_t c = en >> 2, ab = sAB[c];
                const double gq = sG[c];
                const uint32_t oq = (en & 1) ? (ab & 0xFFFF) : (ab >> 16);
#pragma unroll
                for (int q = 0; q < kRegNb; q++)
                    if (q == nn) { nbg[q] = gq; nbo[q] = oq; }
                nn++;
            }
    }
    const double rden = 1.0 / (double)den;
    double lam_s = 0.0;   // this connection lane's scale gradient
    for (int sub = 4; sub >= 0; sub--) {
        if (a.xrec) {
            load_iterates(sub);
            conn_terms(sub);
        } else {
            sweeps(sub);
        }
        // transposed sweeps: lam holds d loss / d x^{k+1}; mu = lam / D alternates between two
        // LDS planes, so one barrier per sweep suffices, and is kept (as float) for the
        // conductance terms summed after the sweeps
        double lb = 0.0, lg = 0.0;
        for (int it = 49, pl = 0; it >= 0; it--, pl ^= 1) {
            double* mus = sMu2 + pl * TS;
            const double mu = act ? lam * rden : 0.0;
            mus[t] = mu;
            muh[it * TS + t] = (float)mu;
            __syncthreads();
            double nl = 0.0;
#pragma unroll
            for (int q = 0; q < kRegNb; q++) nl += nbg[q] * mus[nbo[q]];
            for (uint32_t e = e_rest; e < e1; e++) {
                const uint32_t en = sEnt[e];
                if (!((en >> 1) & 1)) continue;
                const uint32_t c = en >> 2, ab = sAB[c];
                nl += (double)sG[c] * mus[(en & 1) ? (ab & 0xFFFF) : (ab >> 16)];
            }
            lb += mu;
            lam = nl;
        }
        __syncthreads();   // the last sweep's reads of the mu planes are done
        if (cact && cintra)   // d x^{k+1}_i / d g = (x^k_other - x^{k+1}_i) / D_i at both ends
            for (int it = 0; it < 50; it++) {
                const float* xk = xs + it * TS;
                const float* xk1 = xs + (it + 1) * TS;
                lg += (double)muh[it * TS + ca_] * ((double)xk[cb_] - xk1[ca_]) +
                      (double)muh[it * TS + cb_] * ((double)xk[ca_] - xk1[cb_]);
            }
        // x^0 = V(sub start) and base = cdt * V + inpd, so d loss / d inpd = lb
        if (act) sLam[t] = lb;
        __syncthreads();
        double lam_dv = 0.0;
        if (cact && cintra) {
            const double lct = sLam[ca_] - sLam[cb_];   // inpd_a += cterm, inpd_b -= cterm
            lam_dv = lct * (sDinp[t] - (double)gax);     // cterm = s * inp(dV) - s * G * dV
            lg += -lct * sDV[t];
            lam_s += lg * G + lct * sInp1[t];
        }
        __syncthreads();   // every read of d loss / d inpd is done
        sLam[t] = lam_dv;
        __syncthreads();
        if (act) {
            lam = lam + lb * cdt;
            for (uint32_t e = e0; e < e1; e++) {
                const uint32_t en = sEnt[e];
                if (!((en >> 1) & 1)) continue;
                const double ldv = sLam[en >> 2];
                lam += (en & 1) ? -ldv : ldv;   // dV = V_a - V_b
            }
        }
        __syncthreads();
    }

    if (act) {   // d loss / d V before the tick through step 2.5; the channels' part follows
        a.lamV[g] = lam;
        a.lamIb[g] = lam_ibuf;
    }
    if (cact) a.gscale[c0 + t] += lam_s;
}

// One tick of the reverse pass through step 1 (the channels), per segment: the gating ODEs and
// GHK currents at the tick's pre-tick V and gates, their partials from dual numbers.
__global__ __launch_bounds__(256) void k_adjoint_channels(Net d, AdjNet a, float dt) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= d.n) return;
    const size_t P = d.npad;
    adj::SegAd sa;
    for (int k = 0; k < 3; k++) { sa.in[k] = d.in[k * P + g]; sa.ex[k] = d.ex[k * P + g]; }
    sa.area = d.area[g];
    sa.pkc = d.pkc[g]; sa.pka = d.pka[g]; sa.camp = d.camp[g];
    const double V0 = a.recV[g];
    const double lr = a.lamIb[g];
    double lam = a.lamV[g];
    const uint32_t cnt = d.mcount[g];
    const bool my = d.sflags[g] & SF_MYELIN;
    for (uint32_t j = 0; j < cnt; j++) {
        const uint8_t ty = d.mtype[j * P + g];
        if (my && ty != 4 && myelin_blocked(ty)) continue;
        double st[7];
        for (int q = 0; q < 7; q++) st[q] = d.mstate[((size_t)j * 7 + q) * P + g];
        st[0] = a.recG[((size_t)j * 2 + 0) * P + g];
        st[1] = a.recG[((size_t)j * 2 + 1) * P + g];
        const adj::ChanAd c = adj::channel_ad(ty, V0, st, sa, d.mdens[j * P + g], dt);
        lam += lr * c.dr[0];
        a.gdens[(size_t)j * P + g] += lr * c.r1;
        for (int q = 0; q < c.ng; q++) {
            double* lg = &a.lamG[((size_t)j * 2 + q) * P + g];
            const double lgn = *lg + lr * c.dr[1 + q];
            lam += lgn * c.dg[q][0];
            *lg = lgn * c.dg[q][1];
        }
    }
    a.lamV[g] = lam;
}

// Records the pre-tick V and gate states (states 0 and 1 of every slot) of one tick and, when
// recIb is not NULL, the I_buffer the previous tick computed.
__global__ __launch_bounds__(256) void k_adj_record(Net d, double* recV, double* recG, double* recIb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    const size_t P = d.npad;
    recV[i] = d.V[i];
    if (recIb) recIb[i] = d.ibuf[i];
    for (uint32_t j = 0; j < d.nslot; j++)
        for (int q = 0; q < 2; q++) recG[((size_t)j * 2 + q) * P + i] = d.mstate[((size_t)j * 7 + q) * P + i];
}
// Loss terms of one tick: acc[c] += (V(soma of cell c) - target[c])^2; trace[c] = that V.
__global__ __launch_bounds__(256) void k_adj_loss(const double* V, const uint32_t* soma, uint32_t ncells,
                                                  const float* target, double* acc, float* trace) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const double v = V[soma[c]];
    const double e = v - (double)target[c];
    acc[c] += e * e;
    if (trace) trace[c] = (float)v;
}
// d loss / d V(after the tick) at the somas: coef * (V - target).
__global__ __launch_bounds__(256) void k_adj_seed(const double* V, const uint32_t* soma, uint32_t ncells,
                                                  const float* target, double coef, double* lamV) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    lamV[soma[c]] += coef * (V[soma[c]] - (double)target[c]);
}

}  // namespace bc
