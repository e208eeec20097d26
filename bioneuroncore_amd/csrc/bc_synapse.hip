// bc_synapse.hip — step 3 (synaptic release and receptors) and step 4 (membrane update), the
// tick counter and the stimulus-list map.
//
//   k_pre_summary    step 3: per presynaptic segment, the V-only values its synapses share
//   k_syn_pre        step 3, presynaptic half: delay-ring arrivals, bouton Ca, release draw
//   k_syn_post(_rec) step 3, post-synaptic receptors in synapse order per post segment + step 4
//   k_syn_plast      step 3, per-synapse spine Ca, cleft clearance, BCM/CaMKII/weight
//   k_syn_rrp        step 3, vesicle-pool recovery/depletion and delay-ring enqueue
//
// Ordering facts that make the split exact (reference src/simulation_host.cpp:322-593):
// a synapse's pre-side reads only pre V_m, extra Ca and Cav2.1 state, none of which step 3
// writes; the post-side of a segment depends only on its own incoming synapses in index order,
// and of each synapse's post-side only the receptors feed that order (the rest reads the
// synapse's own state and receptor current);
// the one cross-chain read is rrp_do_recovery's pre intra Ca, which k_syn_rrp takes from the
// recorded post-chain history (ca_hist) of the last earlier synapse onto that segment.
#include "bc_tick.h"

namespace bc {

// ------------------------------------------------------------------ step 3, pre-side
// Presynaptic summaries: one lane per segment that is the presynaptic segment of a synapse.
__global__ __launch_bounds__(256) void k_pre_summary(Net d) {
    exp_table_stage();
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.nwatched) return;
    const uint32_t pre = d.watched[w];
    const size_t P = d.npad;
    PreSum ps;
    ps.V = d.V[pre];
    ps.exca = d.ex[2 * P + pre];
    float vb = ps.V;
    ps.vn = vn_di(vb);
    ps.edi = bc::exp(-ps.vn);
    int cs = d.cav21_slot[pre];
    ps.has_cav = cs >= 0;
    ps.dens = 0.0f;
    ps.mi = 0.0f;
    ps.h0 = 0.0f;
    ps.h1 = 0.0;
    if (cs >= 0) {
        ps.dens = d.mdens[cs * P + pre];
        ps.h1 = d.mstate[(size_t)cs * 7 * P + 1 * P + pre];
        ps.mi = cav21_minf(vb);
        ps.h0 = 1.0f / (1.0f + bc::exp((vb - -50.0f) / 8.0f));
    }
    ps.ca_snap = d.in[2 * P + pre];
    d.ca_snap[w] = ps.ca_snap;
    d.presum[w] = ps;
}

// Gap junction of slot y (simulation_host.cpp:372-400): the slow Ca/CaM gate from both ends'
// intracellular Ca (run_tick_common.h:112-122), the connexin conductance (run_tick_common.h:
// 79-110) times the junction voltage, through a first-order filter of time constant `delay`
// when it has one. The current leaves the presynaptic end and enters the postsynaptic end in
// the post-side pass. Both ends' Ca is read at step-3 entry: exact unless an earlier synapse
// changed it (counted as an order hazard at load).
__device__ static void gap_junction(const Net& d, uint32_t y, float dt) {
    const size_t P = d.npad, NS = d.ns;
    const uint32_t pre = d.y_pre[y], post = d.y_post[y];
    const float vj = d.V[pre] - d.V[post];
    const float ca_pre = d.in[2 * P + pre], ca_post = d.in[2 * P + post];
    float m = d.y_dep[y];
    const float avg = 0.5f * (ca_pre + ca_post);
    const float cam = (avg * avg) / (avg * avg + 0.0003f * 0.0003f);
    m += (cam - m) * dt / 100.0f;
    if (m < 0.0f) m = 0.0f;
    if (m > 1.0f) m = 1.0f;
    d.y_dep[y] = m;
    uint32_t pm = d.mcount[pre], qm = d.mcount[post];
    if (pm > 32) pm = 0;
    if (qm > 32) qm = 0;
    float g = 0.0f;
    for (int cx = 0; cx < 2; cx++) {
        const uint8_t ty = cx == 0 ? M_CX36 : M_CX43;
        const float gamma = cx == 0 ? 14.0f : 100.0f;
        float dens = 0.0f, f = 1.0f;
        for (uint32_t j = 0; j < pm; j++)
            if (d.mtype[j * P + pre] == ty) {
                dens = d.mdens[j * P + pre];
                f *= d.mstate[((size_t)j * 7 + 0) * P + pre] * d.mstate[((size_t)j * 7 + 4) * P + pre];
            }
        for (uint32_t j = 0; j < qm; j++)
            if (d.mtype[j * P + post] == ty)
                f *= d.mstate[((size_t)j * 7 + 0) * P + post] * d.mstate[((size_t)j * 7 + 4) * P + post];
        if (cx == 1) {   // Cx43 closes with |Vj| (Boltzmann, half 40 mV, slope 5 mV)
            const float a = vj < 0.0f ? -vj : vj;
            const float fv = 1.0f / (1.0f + bc::exp((a - 40.0f) / 5.0f));
            f *= fv;
        }
        if (f < 0.01f) f = 0.01f;
        g += dens * f * gamma;
    }
    const float raw = g * 1e-3f * (1.0f - m) * vj;
    float out = raw;
    const float delay = d.y_delay[y];
    if (delay > 0.001f) {
        float flt = d.y_cleft[y];   // species 0 holds the filter state
        flt += (raw - flt) * dt / delay;
        d.y_cleft[y] = flt;
        out = flt;
    }
    d.y_igap[y] = out;
}

// Glial capture (simulation_host.cpp:472-490): of the cleft's glutamate, glycine and GABA after
// this tick's releases, the fraction 1 - exp(-dt / tau) (tau: the clearance time, 5 ms if not
// positive) goes to the glial segment nearest the post segment. Stored per synapse here; the
// post-side pass adds it to that segment's transmitters in synapse order and the plasticity
// pass takes it out of the cleft.
__device__ static void glial_capture(const Net& d, uint32_t y, float dt) {
    const size_t NS = d.ns;
    const uint16_t cm = d.y_cleft_mask[y];
    float tau = d.y_tau[y];
    tau = (tau > 0.0f) ? tau : 5.0f;
    const double keep = 1.0f - bc::exp(-dt / tau);
    uint8_t got = 0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int nt = q == 0 ? 0 : (q == 1 ? 7 : 1);
        if (!(cm & (1u << nt))) continue;   // outside the mask the cleft holds +0
        const float c = d.y_cleft[nt * NS + y];
        if (c <= 0.0f) continue;
        d.y_cap[q * NS + y] = (float)(c * keep);
        got |= 1u << q;
    }
    d.y_capm[y] = got;
}

// The vesicle-recovery rate of slot y with its presynaptic segment's CaMKII and calcineurin as
// they stand (kernel_dispatch.h:614-627): RRP_recovery + CaMKII * 0.02 + calcineurin * 0.01.
__device__ static inline float prim_live(const Net& d, uint32_t y, uint32_t pre) {
    const float a = d.y_rrec[y] + d.camkii[pre] * 0.02f;
    return a + d.calcn[pre] * 0.01f;
}

// Immediate release of probability p (simulation_host.cpp:412-456): vesicle recovery with the
// presynaptic Ca the reference would read at this synapse (rrp_do_recovery, kernel_dispatch.h:
// 612-631), depletion (rrp_do_release, :636-645) and the cleft deposit; y_relp = -1 marks it done.
__device__ static void release_now(const Net& d, uint32_t y, float p, float ca, float dt) {
    const size_t NS = d.ns;
    float rrp = d.y_rrp[y], rmax = d.y_rrp_max[y];
    if (!(rmax <= 0.0f || rrp >= rmax)) {
        float caf = ca / (0.0005f + ca);
        float prim = d.y_rrec ? prim_live(d, y, d.y_pre[y]) : d.y_prim0[y];
        prim *= caf;
        float rec = (rmax - rrp) * prim * dt;
        rrp += rec;
        if (rrp > rmax) rrp = rmax;
    }
    float amt;
    float amount = d.y_amount[y];
    if (rmax <= 0.0f) {
        amt = amount * p;
    } else {
        float av = amount;
        if (av > rrp) av = rrp;
        amt = av * p;
        rrp -= amt;
        if (rrp < 0) rrp = 0;
    }
    d.y_rrp[y] = rrp;
    uint8_t nt = d.y_rcount[y] > 0 ? nt_of_receptor(d.y_rtype[y]) : 0;
    d.y_cleft[nt * NS + y] += amt;
    d.y_relp[y] = -1.0f;   // finished here
}

// Presynaptic autoreceptors (simulation_host.cpp:430-436): each, in record order, takes the
// synapse's pre_area as state[7] and multiplies p by its release modulator's factor
// (receptors_biological.cpp:427-491; the CPU step's dispatch with no plugin registered,
// kernel_dispatch.h:575-604). Gi/o types: state[0] relaxes to the cleft activation
// nt / (nt + 0.5) with tau 50 ms, factor 1 - 0.7 * state[0] * density * area; Gs types the same
// with tau 30 ms and 1 + 0.5 * ...; nicotinic: 1 + 0.3 * activation * density * area, no state.
// area: state[7] as float, 1 if not positive. Other IDs: factor 1 (p unchanged).
__device__ static float release_modulators(const Net& d, uint32_t y, uint32_t a, float p, float dt) {
    const size_t NS = d.ns;
    const float pa = d.y_pre_area[y];
    for (uint32_t k = a >> 4, e = (a >> 4) + (a & 15u); k < e; k++) {
        d.ar_st7[k] = pa;
        int nt, kind;   // kind 0: Gi/o, 1: Gs, 2: ionotropic
        switch (d.ar_type[k]) {
            case 18: case 19: nt = 3; kind = 0; break;   // 5-HT1A/1B: serotonin
            case 22: nt = 2; kind = 0; break;            // D2: dopamine
            case 27: nt = 5; kind = 0; break;            // A1: adenosine
            case 46: nt = 5; kind = 1; break;            // A2A
            case 44: nt = 2; kind = 1; break;            // D1
            case 29: case 30: nt = 0; kind = 0; break;   // mGluR2/3: glutamate
            case 36: nt = 9; kind = 0; break;            // CB1: endocannabinoid
            case 38: nt = 8; kind = 0; break;            // H3: histamine
            case 13: nt = 4; kind = 2; break;            // nicotinic: acetylcholine
            case 35: nt = 6; kind = 0; break;            // alpha2: noradrenaline
            default: continue;
        }
        const float c = d.y_cleft[nt * NS + y];
        const float act = c / (c + 0.5f);
        const float area = (pa > 0.0f) ? pa : 1.0f;
        const double scale = d.ar_dens[k] * area;
        float f;
        if (kind == 2) {
            f = 1.0f + 0.3f * act * scale;
        } else {
            double s0 = d.ar_st0[k];
            s0 += (act - s0) * dt / (kind == 0 ? 50.0f : 30.0f);
            d.ar_st0[k] = s0;
            f = kind == 0 ? 1.0f - 0.7f * s0 * scale : 1.0f + 0.5f * s0 * scale;
        }
        p *= f;
    }
    return p;
}

__global__ __launch_bounds__(256) void k_syn_pre(Net d, float dt) {
    exp_table_stage();
    const uint64_t tick = *d.tick;
    const double e300 = bc::exp(-dt / 300.0f);   // Cav2.1 inactivation relaxation, tau 300 ms
    const size_t NS = d.ns;
    for (uint32_t y = blockIdx.x * blockDim.x + threadIdx.x; y < d.ns; y += gridDim.x * blockDim.x) {
        d.y_relp[y] = 0.0f;
        const uint8_t hz = d.y_haz ? d.y_haz[y] : 0;   // resolved in the ordered pass
        if (d.y_flags[y] & YF_SKIP) {
            if ((d.y_flags[y] & YF_GAP) && hz != HZ_GAP) gap_junction(d, y, dt);
            continue;
        }
        const float delay = d.y_delay[y];
        // delayed arrivals (delay_buffer.h:31-43)
        uint8_t mask = d.d_mask[y];
        if (delay > 0.0f && mask) {
            for (int s = 0; s < 4; s++) {
                if (!(mask & (1u << s))) continue;
                uint64_t arr = d.d_arr[s * NS + y];
                if (arr <= tick) {
                    uint8_t nt = d.d_nt[s * NS + y];
                    d.y_cleft[nt * NS + y] += d.d_amt[s * NS + y];
                    d.d_arr[s * NS + y] = 0;
                    mask &= ~(1u << s);
                }
            }
            d.d_mask[y] = mask;
        }
        // bouton Ca microdomain (modifier_helpers.h:764-804)
        const PreSum ps = d.presum[d.y_prew[y]];
        float vb = ps.V;
        float bca = d.y_bca[y];
        float dens = 0.0f, mg = 0.0f, hg = 0.0f;
        if (ps.has_cav) {
            dens = ps.dens;
            float hi = ps.h0;
            hi *= 1.0f / (1.0f + (bca * kMmToUm) / 3.0f);
            mg = ps.mi;
            hg = hi + (ps.h1 - hi) * e300;
        }
        float open = mg * mg * hg;
        double Pc = p_single(15.0f, 2.0f, kCbarCa);
        double flux;
        {   // ghk_divalent(V_bouton, bouton_Ca, extra Ca) with the per-segment exp
            const double V = vb, ci = bca, co = ps.exca, vn = ps.vn;
            const double kk = 4.0 * kFaraday / (kGas * kTemp);
            if (fabs(vn) < 1e-6) flux = 2.0 * kFaraday * (ci - co);
            else if (vn > 20.0) flux = kk * kFaraday * V * 0.001 * ci;
            else if (vn < -20.0) flux = kk * kFaraday * V * 0.001 * co;
            else flux = kk * kFaraday * V * 0.001 * (ci - co * ps.edi) / (1.0 - ps.edi);
        }
        double ia = Pc * flux * (dens * d.y_pre_area[y]) * open;
        float ica = (float)(ia * kAtoPa);
        bca -= ica * dt * 1000.0f / (2.0f * kFaraday * 0.1f);
        bca += (0.0001f - bca) * dt / 5.0f;
        if (bca < 0.0f) bca = 0.0f;
        d.y_bca[y] = bca;
        // Ca-triggered release (simulation_host.cpp:421-459)
        uint8_t refr = d.y_refr[y];
        if (refr == 0) {
            float cu = bca * kMmToUm;
            float act = pow4(cu) / (pow4(cu) + 10000.0);
            float p = d.y_prel[y] * act;
            if (d.y_ar && d.y_ar[y]) p = release_modulators(d, y, d.y_ar[y], p, dt);
            uint32_t seed = d.y_id[y] ^ ((uint32_t)tick) ^ (uint32_t)(int64_t)(ps.V * 1000.0f);
            if (xorshift_draw(seed) < p) {
                d.y_relp[y] = p;
                refr = (uint8_t)(uint32_t)(20.0f / dt + 0.5f);
                if (delay_ticks(delay, dt) == 0 && hz != HZ_RELEASE)   // pre Ca = value at step-3 start
                    release_now(d, y, p, (float)d.in[2 * (size_t)d.npad + d.y_pre[y]], dt);
            }
        }
        if (refr > 0) refr--;
        d.y_refr[y] = refr;
        if ((d.y_flags[y] & YF_GLIA) && hz != HZ_RELEASE) glial_capture(d, y, dt);
    }
}

// One chain step of the receptors-only post pass: every load of slot y at once. The slot arrays
// cover every slot and the receptor arrays d.nrec <= 2 receptors, so the loads a lane does not
// use are still in bounds.
struct PostStep {
    double st[2][3];
    double dens[2];
    float cl0, cl1, parea;
    uint32_t rc;
    uint16_t cm;
    uint8_t yf, rt[2], av[2];
};
__device__ __attribute__((always_inline)) inline void load_step(const Net& d, uint32_t y, PostStep& p) {
    const size_t NS = d.ns;
    p.yf = d.y_flags[y];
    p.cm = d.y_cleft_mask[y];
    p.cl0 = d.y_cleft[y];
    p.cl1 = d.y_cleft[NS + y];
    p.rc = d.y_rcount[y];
    p.parea = d.y_post_area[y];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        p.rt[j] = 0;
        p.av[j] = 0;
        p.dens[j] = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) p.st[j][k] = 0.0;
        if (j < (int)d.nrec) {
            const size_t b = (size_t)j * 7 * NS + y;
            p.rt[j] = d.y_rtype[j * NS + y];
            p.av[j] = d.y_rav[j * NS + y];
            p.dens[j] = d.y_rdens[j * NS + y];
#pragma unroll
            for (int k = 0; k < 3; k++) p.st[j][k] = d.y_rstate[b + (size_t)k * NS];
        }
    }
}

// Step 4 of segment i (simulation_host.cpp:596-649): I_buffer clamp, dV = I / C_m * dt clamped
// to +-50 mV, spike count, voltage clamps in index order (the last active one wins), V clamp.
__device__ static double step4(const Net& d, uint32_t i, double ibuf, double V, float dt) {
    if (ibuf != ibuf) ibuf = 0.0f;
    if (ibuf > 10000.0f) ibuf = 10000.0f;
    if (ibuf < -10000.0f) ibuf = -10000.0f;
    float cm = d.cap[i];
    float prev = V;
    float dV = ibuf / cm * dt;
    if (dV > 50.0f) dV = 50.0f;
    if (dV < -50.0f) dV = -50.0f;
    V += dV;
    if (prev <= -40.0f && V > -40.0f) d.spikes[i]++;
    const uint32_t sh = d.stim.any ? d.stim.head[i] : ~0u;
    if (sh != ~0u) {
        const uint64_t tick = *d.tick + d.stim.tick_shift;
        const uint32_t cnt = d.stim.ent[sh];
        for (uint32_t e = 1; e <= cnt; e++) {
            const uint32_t en = d.stim.ent[sh + e];
            if ((en >> 28) != SK_VCLAMP) continue;
            const VoltageClamp q = d.stim.vc[en & 0x0FFFFFFFu];
            if (tick >= q.start_tick && tick < q.end_tick) V = q.target_mV;
        }
    }
    if (V != V) V = -65.0f;
    else {
        if (V < -120.0f) V = -120.0f;
        if (V > 150.0f) V = 150.0f;
    }
    return V;
}

// ------------------------------------------------------------------ step 3 post-side + step 4
// kReceptorsOnly: every receptor slot holds AMPA, NMDA or GABA-A (a leaner dispatch).
// kPre (with kReceptorsOnly, at most two receptors per synapse): each chain step issues all
// of its loads together.
template <bool kReceptorsOnly, bool kPre = false>
__device__ __attribute__((always_inline)) inline void syn_post(const Net& d, float dt) {
    // kStage: the block's 256-segment window is written back through LDS once every lane is
    // done, one coalesced row per field; lanes walk the window in chain-length order and finish
    // at different times, so direct writes would leave lines partly written for long stretches.
    constexpr bool kStage = kReceptorsOnly;
    __shared__ double o_in[kStage ? 4 : 1][256];
    __shared__ float o_nt[kStage ? 15 : 1][256];
    __shared__ double o_V[256];
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t P = d.npad, NS = d.ns;
    if (lane < d.n) {
    const uint32_t i = d.walk[lane];
    const uint32_t pos = i - blockIdx.x * blockDim.x;   // the walk permutes inside the window
    const uint8_t fl = d.sflags[i];
    const bool ord = d.seg_ord && d.seg_ord[i];   // walked (and stepped) by the ordered pass
    double ibuf = d.ibuf[i];
    double V = d.V[i];
    if (d.ns > 0) {
        // With AMPA/NMDA/GABA-A receptors only, the walk reads and writes ions 0-3 and reads
        // transmitters 0-1; ions 4-6 are only floored and transmitters 2-14 only summed, both
        // done after the walk so the walk keeps fewer registers.
        constexpr int kIonWalk = kReceptorsOnly ? 4 : 7;
        constexpr int kNtWalk = kReceptorsOnly ? 2 : 15;
        SegRegs s;
        s.V = V;
#pragma unroll
        for (int k = 0; k < 7; k++) {   // the receptors read ions 0-3 only
            s.in[k] = k < kIonWalk ? d.in[k * P + i] : 1.0;
            s.ex[k] = k < kIonWalk ? d.ex[k * P + i] : 1.0;
        }
        s.area = d.area[i];
        s.vol = d.vol[i];
        if (!kReceptorsOnly) {   // channels in receptor slots read these and write ex / ATP
            s.atp = d.atp[i];
            s.pkc = d.pkc[i];
            s.pka = d.pka[i];
            s.camp = d.camp[i];
            s.camkii = d.camkii[i];
            s.camkii01 = d.camkii01[i];
        }
        seg_consts_init(s);
        ghk_cache_init(s);
        mg_cache_init(s);
#pragma unroll
        for (int k = 0; k < 15; k++) s.nt[k] = 0.0f;
        if (fl & SF_BATH_NMDA) s.nt[0] = 1.0f;
        if (fl & SF_BATH_5HT) s.nt[3] = 0.0f;
        const bool my = fl & SF_MYELIN;
        const uint32_t len = d.post_len[lane];
        const uint32_t hb = (fl & SF_WATCHED) ? d.hist_base[i] : 0;
        const uint32_t wl = threadIdx.x & 63;
        uint32_t off = d.wave_base[lane >> 6];
        bool ran = false;
        // electrical synapses whose presynaptic end is this segment take their current out of
        // I_buffer in synapse-index order among the chain (simulation_host.cpp:372-400)
        uint32_t go = 0, ge = 0;
        if (d.gap_off && !ord) { go = d.gap_off[i]; ge = d.gap_off[i + 1]; }
        // glial captures this segment receives, merged in synapse-index order (each synapse
        // captures after its own deposit)
        uint32_t co = 0, ce = 0;
        if (!kReceptorsOnly && d.glia_off && !ord) { co = d.glia_off[i]; ce = d.glia_off[i + 1]; }
        auto take = [&](uint32_t e) {
            const uint32_t sl = d.glia_slot[e];
            const uint8_t m = d.y_capm[sl];
            if (m & 1) s.nt[0] += d.y_cap[0 * NS + sl];
            if (m & 2) s.nt[7] += d.y_cap[1 * NS + sl];
            if (m & 4) s.nt[1] += d.y_cap[2 * NS + sl];
        };
        // kPre: the loads of chain step q + 1 are issued before step q is computed (the slot
        // of step q + 1 is known from the lanes still walking then), so each step's memory
        // round trip overlaps the previous step's receptor arithmetic.
        PostStep nx;
        if (kPre && len > 0 && !ord) load_step(d, off + wl, nx);
        for (uint32_t q = 0; q < len; q++) {
            // the lanes still walking are lanes 0..cnt-1 of this wavefront (sorted by length)
            const uint32_t cnt = (uint32_t)__popcll(__ballot(1));
            const uint32_t y = off + wl;
            off += cnt;
            PostStep c;
            if constexpr (kPre) {
                c = nx;
                if (q + 1 < len && !ord) load_step(d, off + wl, nx);
            }
            if (ord) continue;   // the ordered pass walks this chain
            if (go < ge) {   // this segment's junctions as the presynaptic end, up to slot y
                const uint32_t sid = d.y_sid[y];
                for (; go < ge && d.gap_sid[go] <= sid; go++) ibuf -= d.y_igap[d.gap_slot[go]];
            }
            if constexpr (kPre) {
                const uint8_t yf = c.yf;
                const uint16_t cm = c.cm;
                const float cl0 = c.cl0, cl1 = c.cl1;
                const uint32_t rc = c.rc;
                RegSlot rs[2];
                uint8_t rt[2];
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    rt[j] = c.rt[j];
                    rs[j].av = c.av[j];
                    rs[j].dens = c.dens[j];
#pragma unroll
                    for (int k = 0; k < 3; k++) rs[j].st[k] = c.st[j][k];
                    rs[j].mem = d.y_rstate + (size_t)j * 7 * NS + y;
                    rs[j].stride = NS;
                    rs[j].a7 = c.parea;
                }
                if (yf & YF_SKIP) {
                    if (yf & YF_GAP) ibuf += d.y_igap[y];
                    if (fl & SF_WATCHED) d.ca_hist[hb + q] = s.in[2];
                    continue;
                }
                if (cm & 1u) s.nt[0] += cl0;
                if (cm & 2u) s.nt[1] += cl1;
                float isyn = 0.0f;
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    if ((uint32_t)j >= rc) break;
                    if (my && is_ionotropic(rt[j])) continue;
                    isyn += dispatch_receptor(rt[j], rs[j], s, dt);
                    // the receptor set state 0, state 1 (AMPA, GABA-A: desensitisation) or 2 (NMDA)
                    double* st = d.y_rstate + (size_t)j * 7 * NS + y;
                    st[0] = rs[j].st[0];
                    if (rt[j] == M_NMDA) st[2 * NS] = rs[j].st[2];
                    else st[NS] = rs[j].st[1];
                    ran = true;
                }
                ibuf += isyn;
                d.y_isyn[y] = isyn;
                if (fl & SF_WATCHED) d.ca_hist[hb + q] = s.in[2];
                continue;
            }
            const uint8_t yf = d.y_flags[y];
            if (yf & YF_SKIP) {
                if (yf & YF_GAP) ibuf += d.y_igap[y];
                if (fl & SF_WATCHED) d.ca_hist[hb + q] = s.in[2];
                continue;
            }
            const uint16_t cm = d.y_cleft_mask[y];
            uint32_t sid = 0;
            if (co < ce) {
                sid = d.y_sid[y];
                for (; co < ce && d.glia_sid[co] < sid; co++) take(co);
            }
#pragma unroll
            for (int k = 0; k < kNtWalk; k++)
                if (cm & (1u << k)) s.nt[k] += d.y_cleft[k * NS + y];
            if (co < ce && d.glia_sid[co] == sid) take(co++);
            float isyn = 0.0f;
            const uint32_t rc = d.y_rcount[y];
            const float parea = d.y_post_area[y];
            for (uint32_t j = 0; j < rc; j++) {
                uint8_t rt = d.y_rtype[j * NS + y];
                if (my && is_ionotropic(rt)) continue;
                ModSlot m;
                m.st = d.y_rstate + (size_t)j * 7 * NS + y;
                m.stride = NS;
                m.dens = d.y_rdens[j * NS + y];
                m.a7 = parea;
                m.av = d.y_rav[j * NS + y];
                if (!kReceptorsOnly && (rt == 75 || rt == 76)) s.xa = msn_scale(d.mtype, d.mstate, P, i, d.mcount[i], rt);
                isyn += kReceptorsOnly ? dispatch_receptor(rt, m, s, dt) : dispatch(rt, m, s, dt);
                ran = true;
            }
            ibuf += isyn;
            d.y_isyn[y] = isyn;
            if (fl & SF_WATCHED) d.ca_hist[hb + q] = s.in[2];
        }
        for (; go < ge; go++) ibuf -= d.y_igap[d.gap_slot[go]];
        for (; co < ce; co++) take(co);
        // the same chain again, for the transmitters the receptors do not read; skipped by a
        // wavefront none of whose segments receives a synapse with such a transmitter live
        if (kReceptorsOnly && __ballot(fl & SF_NTX) != 0) {
            uint32_t off2 = d.wave_base[lane >> 6];
            for (uint32_t q = 0; q < len; q++) {
                const uint32_t cnt = (uint32_t)__popcll(__ballot(1));
                const uint32_t y = off2 + wl;
                off2 += cnt;
                if (ord || (d.y_flags[y] & YF_SKIP)) continue;
                const uint32_t cm = d.y_cleft_mask[y];
                if (!(cm & ~3u)) continue;
#pragma unroll
                for (int k = 2; k < 15; k++)
                    if (cm & (1u << k)) s.nt[k] += d.y_cleft[k * NS + y];
            }
            // floor7 of every receptor call left ions 4-6 floored once and otherwise untouched
#pragma unroll
            for (int k = 4; k < 7; k++) {
                double v = d.in[k * P + i];
                if (ran && v < 1e-9) d.in[k * P + i] = 1e-9;
            }
        }
        if (kStage) {
#pragma unroll
            for (int k = 0; k < 4; k++) o_in[k][pos] = s.in[k];
#pragma unroll
            for (int k = 0; k < 15; k++) o_nt[kStage ? k : 0][pos] = s.nt[k];
        } else {
#pragma unroll
            for (int k = 0; k < kIonWalk; k++) d.in[k * P + i] = s.in[k];
#pragma unroll
            for (int k = 0; k < 7; k++) d.ex[k * P + i] = s.ex[k];
            d.atp[i] = s.atp;
#pragma unroll
            for (int k = 0; k < 15; k++) d.nt[k * P + i] = s.nt[k];
        }
    } else {
        if (fl & SF_BATH_NMDA) d.nt[0 * P + i] = 1.0f;
        if (fl & SF_BATH_5HT) d.nt[3 * P + i] = 0.0f;
    }
    if (!ord) V = step4(d, i, ibuf, V, dt);
    if (kStage) o_V[pos] = V;
    else d.V[i] = V;
    }   // lane < n
    if (kStage) {
        __syncthreads();
        const uint32_t t = threadIdx.x, g = blockIdx.x * blockDim.x + t;
        if (g < d.n) {
            if (d.ns > 0) {
#pragma unroll
                for (int k = 0; k < 4; k++) d.in[k * P + g] = o_in[k][t];
#pragma unroll
                for (int k = 0; k < 15; k++) d.nt[k * P + g] = o_nt[kStage ? k : 0][t];
            }
            d.V[g] = o_V[t];
        }
    }
}

// AMPA/NMDA/GABA-A receptors only. kPre: at most two receptors per synapse, each chain step's
// loads issued together and one step ahead; three wavefronts per SIMD hold the look-ahead
// registers without spilling (four spill, and measure slower).
template <bool kPre>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_syn_post_rec(
    Net d, float dt) {
    exp_table_stage();
    syn_post<true, kPre>(d, dt);
}
// Any receptor type (full modifier dispatch).
__global__ __launch_bounds__(256) void k_syn_post(Net d, float dt) {
    exp_table_stage();
    syn_post<false>(d, dt);
}

// ------------------------------------------------------------------ step 3, ordered pass
// One lane walks, in synapse-index order, every event that touches an ordered segment (the
// post pass skipped those), as the reference's synapse loop does (simulation_host.cpp:356-592):
// gap junctions computed with both ends' Ca as of their position, currents into and out of
// I_buffer, immediate releases completed with the presynaptic Ca after the last earlier synapse
// onto that segment (its Ca-history entry: the parallel post pass wrote the entries of other
// chains, this walk those of the ordered ones, each before any later synapse reads it), glial
// captures, the post side with the full receptor dispatch, and the Ca history. Then step 4 of
// the ordered segments. Segment state stays in memory between events.
__global__ __launch_bounds__(64) void k_syn_ordered(Net d, float dt) {
    exp_table_stage();
    if (threadIdx.x != 0) return;
    const size_t P = d.npad, NS = d.ns;
    auto take = [&](uint32_t y, float* nt) {   // a glial capture into transmitters nt[0], nt[7], nt[1]
        const uint8_t m = d.y_capm[y];
        if (m & 1) nt[0] += d.y_cap[0 * NS + y];
        if (m & 2) nt[7] += d.y_cap[1 * NS + y];
        if (m & 4) nt[1] += d.y_cap[2 * NS + y];
    };
    for (uint32_t e = 0; e < d.ord_n; e++) {
        const uint32_t y = d.ord_ev[e];
        const uint8_t f = d.ord_fl[e];
        const uint8_t yf = d.y_flags[y];
        const uint32_t post = d.ord_post[e];
        const int32_t h = d.ord_hist[e];
        if (yf & YF_SKIP) {
            if (yf & YF_GAP) {
                if (f & EV_GAPHAZ) gap_junction(d, y, dt);
                const float ig = d.y_igap[y];
                if (f & EV_PREGAP) d.ibuf[d.y_pre[y]] -= ig;
                if (f & EV_POST) d.ibuf[post] += ig;
            }
            if (h >= 0) d.ca_hist[h] = d.in[2 * P + post];
            continue;
        }
        if (f & EV_HAZ) {
            const float p = d.y_relp[y];
            if (p > 0.0f) release_now(d, y, p, (float)d.ca_hist[d.y_src[y]], dt);
            if (yf & YF_GLIA) glial_capture(d, y, dt);
        }
        const bool glia_here = (f & EV_GLIA) && d.ord_glia[e] == post;
        if (f & EV_POST) {
            SegRegs s;
            s.V = d.V[post];
#pragma unroll
            for (int k = 0; k < 7; k++) {
                s.in[k] = d.in[k * P + post];
                s.ex[k] = d.ex[k * P + post];
            }
            s.area = d.area[post];
            s.vol = d.vol[post];
            s.atp = d.atp[post];
            s.pkc = d.pkc[post];
            s.pka = d.pka[post];
            s.camp = d.camp[post];
            s.camkii = d.camkii[post];
            s.camkii01 = d.camkii01[post];
            seg_consts_init(s);
            ghk_cache_init(s);
            mg_cache_init(s);
#pragma unroll
            for (int k = 0; k < 15; k++) s.nt[k] = d.nt[k * P + post];
            const uint16_t cm = d.y_cleft_mask[y];
#pragma unroll
            for (int k = 0; k < 15; k++)
                if (cm & (1u << k)) s.nt[k] += d.y_cleft[k * NS + y];
            if (glia_here) take(y, s.nt);
            const bool my = d.sflags[post] & SF_MYELIN;
            float isyn = 0.0f;
            const uint32_t rc = d.y_rcount[y];
            for (uint32_t j = 0; j < rc; j++) {
                uint8_t rt = d.y_rtype[j * NS + y];
                if (my && is_ionotropic(rt)) continue;
                ModSlot m;
                m.st = d.y_rstate + (size_t)j * 7 * NS + y;
                m.stride = NS;
                m.dens = d.y_rdens[j * NS + y];
                m.a7 = d.y_post_area[y];
                m.av = d.y_rav[j * NS + y];
                isyn += dispatch(rt, m, s, dt);
            }
            d.ibuf[post] += isyn;
            d.y_isyn[y] = isyn;
#pragma unroll
            for (int k = 0; k < 7; k++) {
                d.in[k * P + post] = s.in[k];
                d.ex[k * P + post] = s.ex[k];
            }
            d.atp[post] = s.atp;
#pragma unroll
            for (int k = 0; k < 15; k++) d.nt[k * P + post] = s.nt[k];
            if (h >= 0) d.ca_hist[h] = s.in[2];
        }
        if ((f & EV_GLIA) && !((f & EV_POST) && glia_here)) {
            const uint32_t g = d.ord_glia[e];
            float nt[15];
#pragma unroll
            for (int k = 0; k < 15; k++) nt[k] = d.nt[k * P + g];
            take(y, nt);
            d.nt[0 * P + g] = nt[0];
            d.nt[1 * P + g] = nt[1];
            d.nt[7 * P + g] = nt[7];
        }
    }
    for (uint32_t k = 0; k < d.ord_nseg; k++) {
        const uint32_t i = d.ord_seg[k];
        d.V[i] = step4(d, i, d.ibuf[i], d.V[i], dt);
    }
}

// ------------------------------------------------------------------ step 3, synapse plasticity
// eCB deposits (simulation_host.cpp:578-589): each chemical synapse onto a segment with eCB
// synthesis, in synapse-index order, reads the first such modifier's state[0] as float e and,
// if e > 0, adds e * 0.1 * dt to its cleft's endocannabinoid and multiplies state[0] by 0.9f.
// The level synapse y sees is state[0] after the multiplications of its `rank` predecessors;
// the last synapse writes the level after its own back in k_syn_rrp, when every deposit of the
// tick has read the stored value.
__device__ static double ecb_after(double s, uint32_t takes) {
    for (; takes > 0; takes--) {
        if (!((float)s > 0.0f)) break;
        s *= 0.9f;
    }
    return s;
}
__device__ static size_t ecb_at(const Net& d, uint32_t y) {
    const uint32_t post = d.y_post[y];
    return (size_t)d.seg_ecb[post] * 7 * d.npad + post;
}
__device__ static float ecb_level(const Net& d, uint32_t y, uint32_t rank) {
    return (float)ecb_after(d.mstate[ecb_at(d, y)], rank);
}

// The part of the post-side step that depends only on the synapse's own state and the receptor
// current the ordered pass stored: spine Ca, cleft clearance, BCM threshold, spine
// CaMKII/calcineurin and the weight (simulation_host.cpp:505-576). One lane per synapse.
__global__ __launch_bounds__(256) void k_syn_plast(Net d, float dt) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= d.ns) return;
    const uint8_t yf = d.y_flags[y];
    if (yf & YF_SKIP) return;
    const float isyn = d.y_isyn[y];
    const uint16_t cm = d.y_cleft_mask[y];
    const size_t NS = d.ns;
    // spine Ca (modifier_helpers.h:808-816)
    float sca = d.y_sca[y];
    {
        float ia = isyn < 0.0f ? -isyn : isyn;
        float icaf = ia * 0.1f;
        float influx = icaf * dt * 1000.0f / (2.0f * kFaraday * 0.05f * 100.0f);
        sca += influx;
        sca += (0.0001f - sca) * dt / 50.0f;
        if (sca < 0.0f) sca = 0.0f;
    }
    // cleft clearance (simulation_host.cpp:505-510)
    {
        const float tau = d.y_tau[y];
        const double E = bc::exp_t(-dt / tau, ExpTabGlobal{});
        // with glia, glutamate, glycine and GABA leave the cleft by capture only
        const uint8_t capm = (yf & YF_GLIA) ? d.y_capm[y] : 0;
#pragma unroll
        for (int k = 0; k < 15; k++)
            if (cm & (1u << k)) {
                float v = d.y_cleft[k * NS + y];
                if ((yf & YF_GLIA) && (k == 0 || k == 7 || k == 1)) {
                    const int q = k == 0 ? 0 : (k == 7 ? 1 : 2);
                    if (!(capm & (1u << q))) continue;
                    v -= d.y_cap[q * NS + y];
                } else {
                    v *= E;
                    if (v < 0.0f) v = 0.0f;
                }
                d.y_cleft[k * NS + y] = v;
            }
    }
    // BCM threshold + spine CaMKII/calcineurin + weight (simulation_host.cpp:527-576)
    float facil = d.y_fac[y];
    float cu = sca * 1000.0f;
    float csq = cu * cu;
    float tau_theta = 300000.0f;
    facil += (csq - facil) * dt / tau_theta;
    if (facil < 0.0f) facil = 0.0f;
    float dep = facil + 0.01f;
    float ck = d.y_camkii[y], cn = d.y_calcn[y];
    {
        float c1 = sca * kMmToUm;
        float c4 = c1 * c1 * c1 * c1;
        float ck_inf = c4 / (c4 + 1.0f + 1e-9f);
        ck += (ck_inf - ck) * dt / 2000.0f;
        if (ck < 0.0f) ck = 0.0f;
        if (ck > 1.0f) ck = 1.0f;
        float c2 = c1 * c1;
        float cn_inf = c2 / (c2 + 0.05f * 0.05f + 1e-9f);
        cn += (cn_inf - cn) * dt / 300.0f;
        if (cn < 0.0f) cn = 0.0f;
        if (cn > 1.0f) cn = 1.0f;
    }
    float dw = 0.0f;
    if (ck > 0.0f) dw += 0.0003f * ck * dt;
    if (cn > 0.0f) dw -= 0.0001f * cn * (d.y_postd ? d.pp1[d.y_postd[y]] : d.y_pp1[y]) * dt;
    if (yf & YF_INHIBITORY) dw = -dw;
    float w = d.y_weight[y] + dw;
    if (w != w) w = 0.5f;
    if (facil != facil) facil = 0.0f;
    if (dep != dep) dep = 0.01f;
    if (w < 0.01f) w = 0.01f;
    if (w > 10.0f) w = 10.0f;
    d.y_sca[y] = sca;
    d.y_fac[y] = facil;
    d.y_dep[y] = dep;
    d.y_camkii[y] = ck;
    d.y_calcn[y] = cn;
    d.y_weight[y] = w;
    if (d.y_ecb && d.y_ecb[y] != EC_NONE) {   // eCB retrograde deposit (simulation_host.cpp:578-589)
        const float e = ecb_level(d, y, d.y_ecb[y] & ~EC_LAST);
        if (e > 0.0f) d.y_cleft[9 * NS + y] += e * 0.1f * dt;
    }
}

// ------------------------------------------------------------------ step 3, vesicle pool
__global__ __launch_bounds__(256) void k_syn_rrp(Net d, float dt) {
    const uint64_t tick = *d.tick;
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= d.ns) return;
    if (d.y_flags[y] & YF_SKIP) return;
    if (d.y_ecb && (d.y_ecb[y] & EC_LAST) && d.y_ecb[y] != EC_NONE) {   // the chain's final eCB level
        const size_t at = ecb_at(d, y);
        d.mstate[at] = ecb_after(d.mstate[at], (d.y_ecb[y] & ~EC_LAST) + 1);
    }
    const float p = d.y_relp[y];
    if (p < 0.0f) return;   // immediate release, finished in k_syn_pre
    const size_t NS = d.ns;
    float rrp = d.y_rrp[y], rmax = d.y_rrp_max[y];
    const uint32_t pre = d.y_pre[y];
    if (!(rmax <= 0.0f || rrp >= rmax)) {   // rrp_do_recovery (kernel_dispatch.h:612-631)
        int32_t src = d.y_src[y];
        float ca = src >= 0 ? (float)d.ca_hist[src] : (float)d.ca_snap[d.y_prew[y]];
        float caf = ca / (0.0005f + ca);
        float prim = d.y_rrec ? prim_live(d, y, pre) : d.y_prim0[y];   // RRP_recovery + CaMKII(pre) * 0.02 + calcineurin(pre) * 0.01
        prim *= caf;
        float rec = (rmax - rrp) * prim * dt;
        rrp += rec;
        if (rrp > rmax) rrp = rmax;
    }
    if (p > 0.0f) {   // deferred release: rrp_do_release + enqueue (delay_buffer.h:47-58)
        float amt;
        float amount = d.y_amount[y];
        if (rmax <= 0.0f) {
            amt = amount * p;
        } else {
            float av = amount;
            if (av > rrp) av = rrp;
            amt = av * p;
            rrp -= amt;
            if (rrp < 0) rrp = 0;
        }
        uint8_t nt = d.y_rcount[y] > 0 ? nt_of_receptor(d.y_rtype[y]) : 0;
        uint64_t arrival = tick + delay_ticks(d.y_delay[y], dt);
        uint8_t mask = d.d_mask[y];
        for (int s = 0; s < 4; s++)
            if (!(mask & (1u << s))) {
                d.d_arr[s * NS + y] = arrival;
                d.d_amt[s * NS + y] = amt;
                d.d_nt[s * NS + y] = nt;
                d.d_mask[y] = mask | (1u << s);
                break;
            }
    }
    d.y_rrp[y] = rrp;
}

// ------------------------------------------------------------------ tick counter, stimulus lists
// The tick the kernels read lives in device memory so a sequence of ticks can be captured once
// and replayed; this one-lane kernel ends every tick.
__global__ void k_tick_advance(uint64_t* tick) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *tick += 1;
}

// Points the listed segments at their stimulus lists (heads == nullptr: clears them).
__global__ __launch_bounds__(256) void k_stim_map(const uint32_t* __restrict__ segs, const uint32_t* __restrict__ heads,
                                                  uint32_t count, uint32_t* __restrict__ head) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) head[segs[k]] = heads ? heads[k] : ~0u;
}

template __global__ void k_syn_post_rec<true>(Net, float);
template __global__ void k_syn_post_rec<false>(Net, float);

}  // namespace bc
