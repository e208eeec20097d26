// bc_kernels.hip — the per-tick step as gfx950 kernels.
//
//   k_modifiers      step 1: one lane per segment, modifier slots in order, stimuli
//   k_diffusion      step 2.5: one workgroup per tile of whole islands, all state in LDS
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
#include <hip/hip_runtime.h>
#include <type_traits>
#include "bc_device.h"
#include "bc_state.h"
#include "../../include/biocore_abi.h"

namespace bc {

__device__ static inline bool is_ionotropic(uint8_t t) { return ionotropic_type(t); }
__device__ static inline bool myelin_blocked(uint8_t t) { return myelin_blocks(t); }
__device__ static inline uint8_t nt_of_receptor(uint8_t rt) {   // run_tick_common.h:241-248
    if (rt == 11 || rt == 17) return 1;
    if (rt == 12) return 7;
    if (rt == 14) return 3;
    if (rt == 13) return 4;
    if (rt == 44 || rt == 22) return 2;
    return 0;
}
__device__ static inline uint64_t delay_ticks(float delay, float dt) {   // delay_buffer.h:61-64
    if (delay <= 0.0f || dt <= 0.0f) return 0;
    return (uint64_t)((delay / dt) + 0.5f);
}

// ------------------------------------------------------------------ step 1
// kNt: some segment carries a modifier that reads or takes up transmitters (SF_NTREAD); kExt:
// some segment carries a modifier outside the core set. The variants without keep those
// registers free.
template <bool kNt, bool kExt>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_modifiers(Net d, float dt) {
    exp_table_stage();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
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
        for (int k = 0; k < 12; k++) s.nt[k] = d.nt[k * P + i];
    }
    const bool my = sfl & SF_MYELIN;
    const uint32_t cnt = d.mcount[i];
    double itot = 0.0;
    for (uint32_t j = 0; j < cnt; j++) {
        const uint8_t t = d.mtype[j * P + i];
        if (my && t != 4 && myelin_blocked(t)) continue;
        ModSlot m;
        m.st = d.mstate + (size_t)j * 7 * P + i;
        m.stride = P;
        m.dens = d.mdens[j * P + i];
        m.a7 = s.area;
        m.av = d.mav[j * P + i];
        if (kExt && t == 144) {   // Cav1.4 reads the first CaBP4's state[0] as it stands now
            s.cabp4 = 0.0f;
            for (uint32_t q = 0; q < cnt; q++)
                if (d.mtype[q * P + i] == 145) {
                    s.cabp4 = d.mstate[(size_t)q * 7 * P + i];
                    break;
                }
        }
        if (kExt && t == 86) {   // eCB synthesis reads the first PLC/IP3/DAG modifier's state[1]
            s.dag = 0.0f;
            for (uint32_t q = 0; q < cnt; q++)
                if (d.mtype[q * P + i] == 136) {
                    s.dag = d.mstate[((size_t)q * 7 + 1) * P + i];
                    break;
                }
        }
        itot += dispatch<kNt, kExt>(t, m, s, dt);
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
    }
    d.ibuf[i] = itot;
}

// ------------------------------------------------------------------ step 2.5
// One workgroup of TS lanes owns one tile: whole islands, at most TS segments and TS
// connections, every per-segment and per-connection quantity of the step held in LDS.
// Lane t owns segment t and connection t.
//
// Valences are +1 (Na, K, H), -1 (Cl, A) and +2 (Ca, Mg): x = z*F/RT*dV takes three values
// per connection and B(-x) for z = -1 is B(x) for z = +1 bit for bit, so four Bernoulli
// evaluations serve all seven ions.
template <int TS>
struct DiffLds {
    // Each phase overlays its own planes on u; ion and u together stage the transmitter phase.
    double ion[7][TS];     // intracellular ions (intra phase), then extracellular (extra phase)
    union {
        struct {
            double sig[TS];        // sigma_e per segment
            double gphi[TS];       // phi conductance per connection
            double phi[2][TS + 1]; // entry TS of each plane: 0, the padding target
        } p;                       // phi_e
        struct {
            double V[TS];
            float flux[7][TS];     // per connection
            float vf[2][TS + 1];   // Jacobi iterates; entry TS of each plane: 0
            float cg[TS];
            float cterm[TS];
        } i;                       // intracellular
        double flux[7][TS];        // extracellular, per connection
    } u;
    float cac[TS];
    float cdist[TS];
    uint32_t cab[TS];      // a_local | b_local << 16
    uint16_t ent[2 * TS];  // incidences of each segment in connection order (ent_* below)
};

static_assert(sizeof(DiffLds<64>::ion) + sizeof(((DiffLds<64>*)nullptr)->u) >= 18 * 4 * 64,
              "transmitter staging needs 18 floats per segment in ion + u");

// Tiles per workgroup: 64-lane tiles are single wavefronts that synchronise at wavefront level,
// so up to four can share a workgroup (blockDim.x / 64 of them: a launch-schedule choice).
constexpr int kDiffMaxTiles64 = 4;

// Barrier of one tile. A 64-lane tile is one wavefront, whose LDS accesses complete in issue
// order: a wavefront-scope fence pair around a wave barrier orders them for the compiler.
template <int TS>
__device__ __attribute__((always_inline)) inline void tile_sync() {
    if (TS == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Incidence entry: tile-local connection << 2 | intracellular << 1 | role (0 = end a, 1 = end b).
BC_DEV uint32_t ent_conn(uint32_t en) { return en >> 2; }
BC_DEV bool ent_intra(uint32_t en) { return (en >> 1) & 1; }
template <int TS>
BC_DEV uint32_t ent_other(const DiffLds<TS>& L, uint32_t en) {   // the segment at the other end
    uint32_t ab = L.cab[en >> 2];
    return (en & 1) ? (ab & 0xFFFF) : (ab >> 16);
}

struct BPair { float m, p; };   // B(-x), B(x)

__device__ static inline void bern4(float dV_or_phi_x1, float x2, BPair& b1, BPair& b2) {
    // the electrodiffusion kernel keeps its LDS for state: exp reads the global table
    b1.m = bernoulli(-dV_or_phi_x1, ExpTabGlobal{});
    b1.p = bernoulli(dV_or_phi_x1, ExpTabGlobal{});
    b2.m = bernoulli(-x2, ExpTabGlobal{});
    b2.p = bernoulli(x2, ExpTabGlobal{});
}
// (B(-x), B(x)) of ion k given the +1 and +2 pairs.
__device__ static inline BPair bsel(int k, const BPair& b1, const BPair& b2) {
    if (k == 2 || k == 4) return b2;
    if (k == 3 || k == 5) return BPair{b1.p, b1.m};
    return b1;
}

// All lanes of a tile: pred holds on every lane (a barrier for multi-wavefront tiles).
template <int TS>
__device__ __attribute__((always_inline)) inline bool tile_all(bool pred) {
    if (TS == 64) return __ballot(!pred) == 0;
    return __syncthreads_and(pred) != 0;
}

// Exact early end of the fixed-count (50-sweep) Jacobi loops. A sweep is a pure function of the
// tile's iterate: every other input (diagonal, conductances, right-hand side) is fixed across the
// sweeps, and islands never cross a tile. So when sweep j returns x_j equal to x_{j-1} bit for bit
// on every lane of the tile, every later sweep returns x_j; when x_j equals x_{j-2}, the iterates
// alternate x_j, x_{j-1} from then on and sweep 50 returns x_j for even 50 - j, else x_{j-1}.
// `it` is j - 1. d1: this lane's x_j differs from x_{j-1}; d2: from x_{j-2}. The result is
// uniform over the tile: 0 go on, 1 fixed point, 2 two-cycle.
// For tiles of several wavefronts, each wavefront publishes its two ballots in sflag (two
// sweep-parity buffers of TS/64 words) and the barrier that ends the sweep orders them before
// every wavefront reads all of them; a buffer is rewritten two sweeps later, after the next
// barrier, when every wavefront has read it. The TS = 64 form takes no barrier and no sflag.
template <int TS>
__device__ __attribute__((always_inline)) inline int sweep_settled(bool d1, bool d2, int it,
                                                                   uint32_t* sflag = nullptr) {
    const uint64_t b1 = __ballot(d1), b2 = __ballot(d2);
    uint32_t any = (b1 != 0) | ((b2 != 0) << 1);
    if (TS > 64) {
        uint32_t* f = sflag + (it & 1) * (TS / 64);
        if ((threadIdx.x & 63) == 0) f[threadIdx.x >> 6] = any;
        __syncthreads();
        any = 0;
#pragma unroll
        for (int w = 0; w < TS / 64; w++) any |= f[w];
    }
    if (!(any & 1)) return 1;
    if (it >= 1 && !(any & 2)) return 2;
    return 0;
}
BC_DEV bool bits_ne(double a, double b) { return __double_as_longlong(a) != __double_as_longlong(b); }
BC_DEV bool bits_ne(float a, float b) { return __float_as_uint(a) != __float_as_uint(b); }
// Register form for a one-wavefront tile: nx = x_j, p1 = x_{j-1}, p2 = x_{j-2}.
template <class T>
__device__ __attribute__((always_inline)) inline int sweep_period(T nx, T p1, T p2, int it) {
    return sweep_settled<64>(bits_ne(nx, p1), bits_ne(nx, p2), it);
}
// The value sweep 50 would return, given the settled state found after sweep j = it + 1.
template <class T>
__device__ __attribute__((always_inline)) inline T sweep_final(int settled, T nx, T p1, int it) {
    return (settled == 2 && ((50 - (it + 1)) & 1)) ? p1 : nx;
}

// Residual-terminated sweeps. An island is done when every lane's update is <= tol (NaN is
// not), so each sweep needs per island only whether some lane exceeded: one ballot per
// wavefront, masked with the lanes of this lane's island (islands are contiguous lane ranges of
// the tile; il is the island's first lane), and one LDS store per island and wavefront, by the
// island's first lane in it.
// island_lanes: this lane's island's lanes within its wavefront, as a ballot mask.
__device__ __attribute__((always_inline)) inline uint64_t island_lanes(uint32_t t, uint32_t il) {
    const int lane = (int)(t & 63u);
    const uint64_t starts = __ballot(t == il) | 1ull;   // lane 0 opens its wavefront's first island
    const int s0 = max((int)il - (int)(t - (uint32_t)lane), 0);
    const uint64_t above = starts & ~((2ull << s0) - 1);
    const int e = above ? __ffsll((unsigned long long)above) - 1 : 64;
    const uint64_t upto = e == 64 ? ~0ull : ((1ull << e) - 1);
    return upto & ~((1ull << s0) - 1);
}
// Marks the island's slot when some lane of it (in this wavefront) exceeded tol. over: this
// lane's update is not <= tol (0 for lanes that contribute nothing). Every lane of the
// wavefront must call it.
__device__ __attribute__((always_inline)) inline void island_over(unsigned long long* slots, uint32_t t, uint32_t il,
                                                                  uint64_t imask, bool over) {
    const uint64_t b = __ballot(over);
    if ((b & imask) && (uint32_t)__ffsll((unsigned long long)imask) - 1 == (t & 63u)) slots[il] = 1;
}

// kRes: residual-terminated sweeps (an extension; the reference sweeps 50 times). Each island
// (cell) stops after the first sweep whose largest update is <= d.jac_tol, or after d.jac_max
// sweeps; per sweep, the islands with an update above tol are marked in LDS beyond the tile's
// DiffLds (island_over).
// kRec: cable mode for the adjoint, which records every Jacobi iterate in d.xrec.
template <int TS, bool kRes, bool kRec = false>
__global__ __launch_bounds__(TS == 64 ? 64 * kDiffMaxTiles64 : TS) void k_diffusion(Net d, TileSet ts, float dt) {
    // 64-lane tiles: each wavefront of the workgroup owns its own tile and LDS slice.
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t sub = TS == 64 ? threadIdx.x >> 6 : 0;
    DiffLds<TS>& L = *reinterpret_cast<DiffLds<TS>*>(lds_raw + sub * sizeof(DiffLds<TS>));
    unsigned long long* ires = nullptr;   // [2][TS]: residual (double bits) per island, by first lane
    if (kRes)
        ires = reinterpret_cast<unsigned long long*>(lds_raw + (blockDim.x / TS) * sizeof(DiffLds<TS>)) + sub * 2 * TS;
    const uint32_t tile = blockIdx.x * (TS == 64 ? blockDim.x >> 6 : 1) + sub;
    if (tile >= ts.count) return;   // whole wavefront: no workgroup barrier follows for TS = 64
    const uint32_t s0 = ts.seg0[tile];
    const uint32_t nseg = ts.nseg[tile];
    const uint32_t c0 = ts.conn0[tile];
    const uint32_t nconn = ts.nconn[tile];
    const uint32_t t = threadIdx.x & (TS - 1);
    const bool act = t < nseg;
    const bool cact = t < nconn;
    const uint32_t g = s0 + t;
    const size_t P = d.npad;
    const uint32_t E0 = d.csr_off[s0];
    const uint32_t nent = d.csr_off[s0 + nseg] - E0;
    uint32_t* sflag = nullptr;   // settle flags of the sweeps (tiles of several wavefronts)
    if constexpr (TS > 64) {
        __shared__ uint32_t sflag_[2 * (TS / 64)];
        sflag = sflag_;
    }

    double ibuf = 0.0, diag = 0.0, sg = 0.0, vcur = 0.0;
    float area = 0.0f, cap = 0.0f, volf = 0.0f;
    double rvol = 0.0;   // RN(1 / volf): divisions by volf below are div_f24 / div_wide
    uint32_t e0 = 0, e1 = 0;
    if (act) {
        vcur = d.V[g];
#pragma unroll
        for (int k = 0; k < 7; k++) {
            L.ion[k][t] = d.in[k * P + g];
            sg += kZ[k] * kZ[k] * kDiff[k] * d.ex[k * P + g];
        }
        ibuf = d.ibuf[g];
        area = d.area[g];
        cap = d.cap[g];
        volf = (float)d.vol[g];
        if (volf < 1e-6f) volf = 1e-6f;
        rvol = 1.0 / (double)volf;
        e0 = d.csr_off[g] - E0;
        e1 = d.csr_off[g + 1] - E0;
    }
    uint32_t ca_ = 0, cb_ = 0;
    float cac = 0.0f, cdist = 0.0f;
    bool cintra = false;
    if (cact) {
        uint32_t ab = d.conn_ab[c0 + t];
        L.cab[t] = ab;
        cintra = d.conn_intra[c0 + t] != 0;
        // connection geometry (diffusion_kernel.h:72-80,162-172)
        ca_ = ab & 0xFFFF;
        cb_ = ab >> 16;
        float ra = d.radius[s0 + ca_], rb = d.radius[s0 + cb_];
        float rmin = ra < rb ? ra : rb;
        cac = kPi * rmin * rmin;
        if (cac < 1e-6f) cac = 1e-6f;
        cdist = (d.length[s0 + ca_] + d.length[s0 + cb_]) * 0.5f;
        if (cdist < 0.1f) cdist = 0.1f;
        L.cac[t] = cac;
        L.cdist[t] = cdist;
    }
    // ---- phi_e (diffusion_kernel.h:41-104) ----
    double* sig = L.u.p.sig;
    if (act) {
        sg = kSigmaE * sg;
        sig[t] = sg;
        diag = sg * area / 1.0f;
        L.u.p.phi[0][t] = 0.0;
    }
    if (t == 0) {
        L.u.p.phi[0][TS] = 0.0;
        L.u.p.phi[1][TS] = 0.0;
    }
    tile_sync<TS>();
    for (uint32_t e = t; e < nent; e += TS) {
        uint32_t en = d.csr_ent[E0 + e];
        uint32_t c = en >> 1;
        L.ent[e] = (uint16_t)((c << 2) | ((d.conn_intra[c0 + c] != 0) << 1) | (en & 1));
    }
    double* gphi = L.u.p.gphi;
    if (cact) gphi[t] = 0.5 * (sig[ca_] + sig[cb_]) * cac / cdist;
    tile_sync<TS>();
    // The first D incidences of a lane stay in registers across the sweeps, D = 2, 3 or 4 from
    // the tile's largest incidence count (chains 2, trees 3); the rest are read from LDS in order
    // after them. Unused register slots read the zero entry with weight 0: a sum that starts at
    // +0 is never -0, so adding +0 leaves it unchanged, and every read of a sweep issues at once.
    const uint32_t maxinc = ts.maxinc[tile];
    const int dsel = maxinc > 3 ? 4 : (maxinc > 2 ? 3 : 2);
    int cur = 0;
    auto phi_sweeps = [&](auto dc) {
        constexpr int D = decltype(dc)::value;
        double pg[D];
        uint32_t po[D];
#pragma unroll
        for (int q = 0; q < D; q++) { pg[q] = 0.0; po[q] = TS; }
        int pdeg = 0;
        uint32_t pe_rest = e1;
        if (act)
            for (uint32_t e = e0; e < e1; e++) {
                uint32_t en = L.ent[e];
                double gg = gphi[ent_conn(en)];
                diag += gg;
                if (gg <= 0.0) continue;
                if (pdeg < D) {
                    pg[pdeg] = gg;
                    po[pdeg] = ent_other(L, en);
                    pdeg++;
                    if (pdeg == D) pe_rest = e + 1;
                }
            }
        const double ydiag = rcp_prep(diag);   // diag is fixed across the sweeps
        if constexpr (kRes) {
            const uint32_t il = act ? d.isl_lane0[g] : t;
            const uint64_t imask = island_lanes(t, il);
            ires[t] = 0;
            ires[TS + t] = 0;
            tile_sync<TS>();
            bool done = !act;
            for (int it = 0; it < d.jac_max; it++) {
                const int rb = it & 1;
                bool over = false;
                if (!done) {
                    double pn = 0.0;
#pragma unroll
                    for (int q = 0; q < D; q++) pn += pg[q] * L.u.p.phi[cur][po[q]];
                    for (uint32_t e = pe_rest; e < e1; e++) {
                        uint32_t en = L.ent[e];
                        double gg = gphi[ent_conn(en)];
                        if (gg <= 0.0) continue;
                        pn += gg * L.u.p.phi[cur][ent_other(L, en)];
                    }
                    double ph = (diag > 1e-9) ? div_rcp(pn - ibuf, diag, ydiag) : 0.0;
                    if (ph > 10.0) ph = 10.0;
                    if (ph < -10.0) ph = -10.0;
                    L.u.p.phi[cur ^ 1][t] = ph;
                    over = !(fabs(ph - L.u.p.phi[cur][t]) <= d.jac_tol);
                } else if (act) {
                    L.u.p.phi[cur ^ 1][t] = L.u.p.phi[cur][t];
                }
                island_over(ires + rb * TS, t, il, imask, over);
                tile_sync<TS>();
                if (!done) done = ires[rb * TS + il] == 0;
                ires[(rb ^ 1) * TS + t] = 0;
                cur ^= 1;
                if (tile_all<TS>(done)) break;
            }
            tile_sync<TS>();
            return;
        }
        if constexpr (TS == 64) {
            if (maxinc <= 4) {   // registers + ds_bpermute, as the Jacobi sweeps below
                int pa[D];
#pragma unroll
                for (int q = 0; q < D; q++) pa[q] = (int)((q < pdeg ? po[q] : t) * 4);
                double pi = 0.0, pp = 0.0;   // x_{j-1}, x_{j-2}
                for (int it = 0; it < 50; it++) {
                    const int lo = (int)(uint32_t)__double_as_longlong(pi);
                    const int hi = (int)(uint32_t)((uint64_t)__double_as_longlong(pi) >> 32);
                    double pv[D];
#pragma unroll
                    for (int q = 0; q < D; q++) {
                        uint32_t vlo = (uint32_t)__builtin_amdgcn_ds_bpermute(pa[q], lo);
                        uint32_t vhi = (uint32_t)__builtin_amdgcn_ds_bpermute(pa[q], hi);
                        pv[q] = __longlong_as_double((long long)(((uint64_t)vhi << 32) | vlo));
                    }
                    double pn = 0.0;
#pragma unroll
                    for (int q = 0; q < D; q++) {
                        double s = pn + pg[q] * pv[q];
                        pn = q < pdeg ? s : pn;
                    }
                    double ph = (diag > 1e-9) ? div_rcp(pn - ibuf, diag, ydiag) : 0.0;
                    if (ph > 10.0) ph = 10.0;
                    if (ph < -10.0) ph = -10.0;
                    const double nx = act ? ph : pi;
                    // checked after every second sweep: a fixed point or two-cycle, once
                    // reached, persists, and sweep_final takes the sweep it is found at
                    const int pfx = (it & 1) ? sweep_period(nx, pi, pp, it) : 0;
                    if (pfx) {
                        pi = sweep_final(pfx, nx, pi, it);
                        break;
                    }
                    pp = pi;
                    pi = nx;
                }
                if (act) L.u.p.phi[0][t] = pi;
                cur = 0;
                tile_sync<TS>();
                return;
            }
        }
        double p1 = 0.0, p2 = 0.0;   // this lane's x_{j-1}, x_{j-2}
        for (int it = 0; it < 50; it++) {
            double ph = p1;
            if (act) {
                double pv[D];
#pragma unroll
                for (int q = 0; q < D; q++) pv[q] = L.u.p.phi[cur][po[q]];
                double pn = 0.0;
#pragma unroll
                for (int q = 0; q < D; q++) pn += pg[q] * pv[q];
                for (uint32_t e = pe_rest; e < e1; e++) {
                    uint32_t en = L.ent[e];
                    double gg = gphi[ent_conn(en)];
                    if (gg <= 0.0) continue;
                    pn += gg * L.u.p.phi[cur][ent_other(L, en)];
                }
                ph = (diag > 1e-9) ? div_rcp(pn - ibuf, diag, ydiag) : 0.0;
                if (ph > 10.0) ph = 10.0;
                if (ph < -10.0) ph = -10.0;
                L.u.p.phi[cur ^ 1][t] = ph;
            }
            // the reduction is also the barrier before the next sweep reads the new plane
            const int settled = sweep_settled<TS>(bits_ne(ph, p1), bits_ne(ph, p2), it, sflag);
            if (TS == 64) tile_sync<TS>();
            cur ^= 1;
            if (settled) {   // phi[cur] holds x_j, phi[cur ^ 1] x_{j-1}
                if (sweep_final(settled, 1, 0, it) == 0) cur ^= 1;
                break;
            }
            p2 = p1;
            p1 = ph;
        }
    };
    if (d.cable) {   // phi_e = 0 (both planes hold zeros)
    } else if (dsel == 2) phi_sweeps(std::integral_constant<int, 2>{});
    else if (dsel == 3) phi_sweeps(std::integral_constant<int, 3>{});
    else phi_sweeps(std::integral_constant<int, 4>{});

    // phi of both ends of each connection stays in registers; the phi planes are dead after.
    double pha = 0.0, phb = 0.0;
    if (cact) {
        pha = L.u.p.phi[cur][ca_];
        phb = L.u.p.phi[cur][cb_];
    }
    tile_sync<TS>();
    if (act) L.u.i.V[t] = vcur;
    if (t == 0) {
        L.u.i.vf[0][TS] = 0.0f;
        L.u.i.vf[1][TS] = 0.0f;
    }
    tile_sync<TS>();

    // ---- intracellular Nernst-Planck, 5 substeps (diffusion_kernel.h:151-238) ----
    const int kSub = 5;
    const float dts = dt / kSub;
    float cdt = cap / dts;
    if (cdt < 1e-9f) cdt = 1e-9f;
    float* cfin = &L.u.i.flux[0][0];   // [7][TS]
    float kin[7];
#pragma unroll
    for (int k = 0; k < 7; k++) kin[k] = kDiff[k] * cac * dts / cdist;
    const double rdts = 1.0 / (double)dts;                  // divisions by dts and cdist in the
    const double rcdist = cact ? 1.0 / (double)cdist : 0.0;  // substeps are div_wide
    for (int sub = 0; sub < kSub; sub++) {
        if (cact && cintra) {
            const uint32_t a = ca_, b = cb_;
            float sgi = 0.0f;
#pragma unroll
            for (int k = 0; k < 7; k++) {
                float cbar = 0.5f * (L.ion[k][a] + L.ion[k][b]);
                sgi += kZ[k] * kZ[k] * kDiff[k] * cbar;
            }
            float gax = div_wide(kSigmaE * sgi * cac, cdist, rcdist);
            const float csc = d.conn_scale ? d.conn_scale[c0 + t] : 1.0f;
            if (d.conn_scale) gax *= csc;
            L.u.i.cg[t] = gax;
            float dV = (L.u.i.V[a] + pha) - (L.u.i.V[b] + phb);
            BPair b1, b2;
            bern4(1.0f * kFrtMv * dV, 2.0f * kFrtMv * dV, b1, b2);
            float inp = 0.0f;
#pragma unroll
            for (int k = 0; k < 7; k++) {
                float ca = L.ion[k][a], cb = L.ion[k][b];
                BPair bb = bsel(k, b1, b2);
                float fl = kin[k] * (ca * bb.m - cb * bb.p);
                if (d.conn_scale) fl *= csc;
                cfin[k * TS + t] = fl;
                inp += div_wide(kZ[k] * kChargeToPa * fl, dts, rdts);
            }
            L.u.i.cterm[t] = inp - gax * dV;
        }
        tile_sync<TS>();
        float gdiag = 0.0f, inpd = 0.0f;
        double dacc[7];
#pragma unroll
        for (int k = 0; k < 7; k++) dacc[k] = 0.0;
        float vold = 0.0f;
        if (act) {
            for (uint32_t e = e0; e < e1; e++) {
                uint32_t en = L.ent[e];
                uint32_t c = ent_conn(en);
                if (!ent_intra(en)) continue;
                gdiag += L.u.i.cg[c];
                if (en & 1) {
                    inpd -= L.u.i.cterm[c];
#pragma unroll
                    for (int k = 0; k < 7; k++) dacc[k] += div_f24((double)cfin[k * TS + c], (double)volf, rvol);
                } else {
                    inpd += L.u.i.cterm[c];
#pragma unroll
                    for (int k = 0; k < 7; k++) dacc[k] -= div_f24((double)cfin[k * TS + c], (double)volf, rvol);
                }
            }
            vold = (float)vcur;
            L.u.i.vf[0][t] = vold;
            if (kRec) d.xrec[(size_t)sub * 51 * P + g] = vold;   // the adjoint's recording
        }
        tile_sync<TS>();
        float vnew = vold;
        const float den = cdt + gdiag;
        const float base = cdt * vold + inpd;
        const double rden = 1.0 / (double)den;
        auto jacobi = [&](auto dc) {   // register incidences as in phi_sweeps
            constexpr int D = decltype(dc)::value;
            float jg[D];
            uint32_t jo[D];
#pragma unroll
            for (int q = 0; q < D; q++) { jg[q] = 0.0f; jo[q] = TS; }
            int jdeg = 0;
            uint32_t je_rest = e1;
            if (act)
                for (uint32_t e = e0; e < e1 && jdeg < D; e++) {
                    uint32_t en = L.ent[e];
                    uint32_t c = ent_conn(en);
                    if (!ent_intra(en)) continue;
                    jg[jdeg] = L.u.i.cg[c];
                    jo[jdeg] = ent_other(L, en);
                    jdeg++;
                    if (jdeg == D) je_rest = e + 1;
                }
            if constexpr (kRes) {   // residual-terminated, as the phi sweeps
                const uint32_t il = act ? d.isl_lane0[g] : t;
                const uint64_t imask = island_lanes(t, il);
                ires[t] = 0;
                ires[TS + t] = 0;
                tile_sync<TS>();
                bool done = !act;
                int vc = 0;
                for (int it = 0; it < d.jac_max; it++) {
                    const int rb = it & 1;
                    bool over = false;
                    if (!done) {
                        float off = 0.0f;
#pragma unroll
                        for (int q = 0; q < D; q++) off += jg[q] * L.u.i.vf[vc][jo[q]];
                        for (uint32_t e = je_rest; e < e1; e++) {
                            uint32_t en = L.ent[e];
                            uint32_t c = ent_conn(en);
                            if (!ent_intra(en)) continue;
                            off += L.u.i.cg[c] * L.u.i.vf[vc][ent_other(L, en)];
                        }
                        const float v = div_wide(base + off, den, rden);
                        L.u.i.vf[vc ^ 1][t] = v;
                        over = !((double)fabsf(v - vnew) <= d.jac_tol);
                        vnew = v;
                    } else if (act) {
                        L.u.i.vf[vc ^ 1][t] = vnew;
                    }
                    island_over(ires + rb * TS, t, il, imask, over);
                    tile_sync<TS>();
                    if (!done) done = ires[rb * TS + il] == 0;
                    ires[(rb ^ 1) * TS + t] = 0;
                    vc ^= 1;
                    if (tile_all<TS>(done)) break;
                }
                tile_sync<TS>();
                return;
            }
            if constexpr (TS == 64) {
                if (maxinc <= 4) {
                    // The tile is one wavefront and every incidence has a register slot: the
                    // iterates stay in registers, neighbours' values come by ds_bpermute. Padded
                    // slots read the lane itself and are not added.
                    int ja[D];
#pragma unroll
                    for (int q = 0; q < D; q++) ja[q] = (int)((q < jdeg ? jo[q] : t) * 4);
                    float vi = vold, vp = vold;   // x_{j-1}, x_{j-2}
                    for (int it = 0; it < 50; it++) {
                        float jv[D];
#pragma unroll
                        for (int q = 0; q < D; q++)
                            jv[q] = __int_as_float(__builtin_amdgcn_ds_bpermute(ja[q], __float_as_int(vi)));
                        // padded slots hold g = 0 and read the lane itself: off starts at +0, so
                        // it is never -0 and adding 0 * v (v finite) leaves it unchanged
                        float off = 0.0f;
#pragma unroll
                        for (int q = 0; q < D; q++) off = off + jg[q] * jv[q];
                        float nv = div_wide(base + off, den, rden);
                        const float nx = act ? nv : vi;
                        if (kRec) {   // the adjoint's recording keeps every sweep
                            if (act) d.xrec[((size_t)sub * 51 + it + 1) * P + g] = nv;
                        } else {
                            const int settled = (it & 1) ? sweep_period(nx, vi, vp, it) : 0;
                            if (settled) {
                                vi = sweep_final(settled, nx, vi, it);
                                break;
                            }
                        }
                        vp = vi;
                        vi = nx;
                    }
                    vnew = vi;
                    return;
                }
            }
            int vc = 0;
            float vp = vnew;   // this lane's x_{j-2} (vnew holds x_{j-1})
            for (int it = 0; it < 50; it++) {
                float nx = vnew;
                if (act) {
                    float jv[D];
#pragma unroll
                    for (int q = 0; q < D; q++) jv[q] = L.u.i.vf[vc][jo[q]];
                    float off = 0.0f;
#pragma unroll
                    for (int q = 0; q < D; q++) off += jg[q] * jv[q];
                    for (uint32_t e = je_rest; e < e1; e++) {
                        uint32_t en = L.ent[e];
                        uint32_t c = ent_conn(en);
                        if (!ent_intra(en)) continue;
                        off += L.u.i.cg[c] * L.u.i.vf[vc][ent_other(L, en)];
                    }
                    nx = div_wide(base + off, den, rden);
                    L.u.i.vf[vc ^ 1][t] = nx;
                    if (kRec) d.xrec[((size_t)sub * 51 + it + 1) * P + g] = nx;
                }
                if (kRec) {   // the adjoint's recording keeps every sweep
                    tile_sync<TS>();
                } else {
                    // the reduction is also the barrier before the next sweep reads the new plane
                    const int settled = sweep_settled<TS>(bits_ne(nx, vnew), bits_ne(nx, vp), it, sflag);
                    if (TS == 64) tile_sync<TS>();
                    if (settled) {
                        vnew = sweep_final(settled, nx, vnew, it);
                        break;
                    }
                }
                vp = vnew;
                vnew = nx;
                vc ^= 1;
            }
        };
        if (dsel == 2) jacobi(std::integral_constant<int, 2>{});
        else if (dsel == 3) jacobi(std::integral_constant<int, 3>{});
        else jacobi(std::integral_constant<int, 4>{});
        if (act) {
            vcur = vnew;
            L.u.i.V[t] = vcur;
            if (!d.cable)
#pragma unroll
                for (int k = 0; k < 7; k++) L.ion[k][t] += (float)dacc[k];
        }
        tile_sync<TS>();
    }
    // intracellular ions are final: store them, then reuse the plane for extracellular ions
    if (d.cable) {   // concentrations, ATP, temperature and transmitters are held
        if (act) d.V[g] = vcur;
        return;
    }
    if (act) {
        d.V[g] = vcur;
#pragma unroll
        for (int k = 0; k < 7; k++) {
            d.in[k * P + g] = L.ion[k][t];
            L.ion[k][t] = d.ex[k * P + g];
        }
    }
    tile_sync<TS>();

    // ---- extracellular Nernst-Planck, 5 substeps (diffusion_kernel.h:245-281) ----
    // phi is fixed across the substeps, so each connection's Bernoulli factors are too.
    BPair e1p, e2p;
    if (cact) {
        double dphi = pha - phb;
        bern4(1.0f * kFrtMv * dphi, 2.0f * kFrtMv * dphi, e1p, e2p);
    }
    for (int sub = 0; sub < kSub; sub++) {
        if (cact) {
#pragma unroll
            for (int k = 0; k < 7; k++) {
                double ca = L.ion[k][ca_], cb = L.ion[k][cb_];
                BPair bb = bsel(k, e1p, e2p);
                L.u.flux[k][t] = kin[k] * (ca * bb.m - cb * bb.p);
            }
        }
        tile_sync<TS>();
        if (act) {
            double dx[7];
#pragma unroll
            for (int k = 0; k < 7; k++) dx[k] = 0.0;
            const float ve = volf * kEvr;
            const double rve = 1.0 / (double)ve;
            for (uint32_t e = e0; e < e1; e++) {
                uint32_t en = L.ent[e];
                uint32_t c = ent_conn(en);
                if (en & 1) {
#pragma unroll
                    for (int k = 0; k < 7; k++) dx[k] += div_f24(L.u.flux[k][c], (double)ve, rve);
                } else {
#pragma unroll
                    for (int k = 0; k < 7; k++) dx[k] -= div_f24(L.u.flux[k][c], (double)ve, rve);
                }
            }
            // Only this lane reads its own ex during the gather; neighbours read cbuf.
#pragma unroll
            for (int k = 0; k < 7; k++) L.ion[k][t] += dx[k];
        }
        tile_sync<TS>();
    }
    if (act) {
#pragma unroll
        for (int k = 0; k < 7; k++) {
            double x = L.ion[k][t];
            x += (kBath[k] - x) * dt / 1.0f;
            if (x < 1e-9) x = 1e-9;
            if (x > 1000.0) x = 1000.0;
            d.ex[k * P + g] = x;
        }
    }
    tile_sync<TS>();

    // ---- transmitters, ATP, temperature, transcription (diffusion_kernel.h:284-345) ----
    // Staged in LDS (ion + cbuf planes are free now): 15 + 3 floats per segment, then 15.
    float* st = reinterpret_cast<float*>(&L.ion[0][0]);   // ion + u: 28 floats per segment
    auto fick = [&](int nsp, int base_sp, bool do_scal, float* dsp, float& datp, float& dtemp, float& dtxn,
                    const float* dcoef, bool intra_only) {
        for (uint32_t e = e0; e < e1; e++) {
            uint32_t en = L.ent[e];
            uint32_t c = ent_conn(en);
            const bool role_b = en & 1;
            uint32_t a = L.cab[c] & 0xFFFF, b = L.cab[c] >> 16;
            float ac = L.cac[c], dist = L.cdist[c];
            const bool intra = ent_intra(en);
            if (!intra_only || intra) {
                for (int k = 0; k < nsp; k++) {
                    float dC = st[k * TS + a] - st[k * TS + b];
                    if (dC < 1e-12f && dC > -1e-12f) continue;
                    float fl = dcoef[base_sp + k] * ac * dC * dt / dist;
                    if (role_b) dsp[k] += div_wide(fl, volf, rvol);
                    else dsp[k] -= div_wide(fl, volf, rvol);
                }
            }
            if (do_scal && intra) {
                float d1 = st[15 * TS + a] - st[15 * TS + b];
                if (d1 > 1e-12f || d1 < -1e-12f) {
                    float fl = 0.14f * ac * d1 * dt / dist;
                    if (role_b) datp += div_wide(fl, volf, rvol); else datp -= div_wide(fl, volf, rvol);
                }
                float d2 = st[16 * TS + a] - st[16 * TS + b];
                if (d2 > 0.001f || d2 < -0.001f) {
                    float fl = 0.0014f * ac * d2 * dt / dist;
                    if (role_b) dtemp += div_wide(fl, volf, rvol); else dtemp -= div_wide(fl, volf, rvol);
                }
                float d3 = st[17 * TS + a] - st[17 * TS + b];
                if (d3 > 1e-9f || d3 < -1e-9f) {
                    float fl = 0.01f * ac * d3 * dt / dist;
                    if (role_b) dtxn += div_wide(fl, volf, rvol); else dtxn -= div_wide(fl, volf, rvol);
                }
            }
        }
    };
    float dnt[15], dnt_in[15];
#pragma unroll
    for (int k = 0; k < 15; k++) { dnt[k] = 0.0f; dnt_in[k] = 0.0f; }
    float datp = 0.0f, dtemp = 0.0f, dtxn = 0.0f, z0 = 0.0f, z1 = 0.0f, z2 = 0.0f;
    float own_nt[15], own_scal[3];
    const bool nt_ex = d.nt_ex_live != 0, nt_in = d.nt_in_live != 0;
    if (act) {
        if (nt_ex)
            for (int k = 0; k < 15; k++) st[k * TS + t] = own_nt[k] = d.nt[k * P + g];
        st[15 * TS + t] = own_scal[0] = d.atp[g];
        st[16 * TS + t] = own_scal[1] = d.temp[g];
        st[17 * TS + t] = own_scal[2] = d.txn[g];
    }
    tile_sync<TS>();
    if (nt_ex) fick(15, 0, true, dnt, datp, dtemp, dtxn, d.nt_diff, false);
    else fick(0, 0, true, dnt, datp, dtemp, dtxn, d.nt_diff, false);
    if (nt_in) {
        tile_sync<TS>();
        if (act)
            for (int k = 0; k < 15; k++) st[k * TS + t] = d.nt_in[k * P + g];
        tile_sync<TS>();
        fick(15, 0, false, dnt_in, z0, z1, z2, d.nt_diff, true);
    }
    if (!act) return;
    if (nt_in)
        for (int k = 0; k < 15; k++) d.nt_in[k * P + g] = st[k * TS + t] + dnt_in[k];
    if (nt_ex)
        for (int k = 0; k < 15; k++) d.nt[k * P + g] = own_nt[k] + dnt[k];
    d.atp[g] = own_scal[0] + datp;
    d.temp[g] = own_scal[1] + dtemp;
    d.txn[g] = own_scal[2] + dtxn;
}

// Without any connection step 2.5 is skipped entirely (diffusion_kernel.h:118); nothing here.

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

// Immediate release of probability p (simulation_host.cpp:412-456): vesicle recovery with the
// presynaptic Ca the reference would read at this synapse (rrp_do_recovery, kernel_dispatch.h:
// 612-631), depletion (rrp_do_release, :636-645) and the cleft deposit; y_relp = -1 marks it done.
__device__ static void release_now(const Net& d, uint32_t y, float p, float ca, float dt) {
    const size_t NS = d.ns;
    float rrp = d.y_rrp[y], rmax = d.y_rrp_max[y];
    if (!(rmax <= 0.0f || rrp >= rmax)) {
        float caf = ca / (0.0005f + ca);
        float prim = d.y_prim0[y];
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
        for (uint32_t q = 0; q < len; q++) {
            // the lanes still walking are lanes 0..cnt-1 of this wavefront (sorted by length)
            const uint32_t cnt = (uint32_t)__popcll(__ballot(1));
            const uint32_t y = off + wl;
            off += cnt;
            if (ord) continue;   // the ordered pass walks this chain
            if (go < ge) {   // this segment's junctions as the presynaptic end, up to slot y
                const uint32_t sid = d.y_sid[y];
                for (; go < ge && d.gap_sid[go] <= sid; go++) ibuf -= d.y_igap[d.gap_slot[go]];
            }
            if constexpr (kPre) {
                // every load of the step at once (one memory round trip per step): the slot
                // arrays cover every slot and the receptor arrays d.nrec <= 2 receptors, so the
                // loads a lane does not use are still in bounds
                const uint8_t yf = d.y_flags[y];
                const uint16_t cm = d.y_cleft_mask[y];
                const float cl0 = d.y_cleft[y], cl1 = d.y_cleft[NS + y];
                const uint32_t rc = d.y_rcount[y];
                const float parea = d.y_post_area[y];
                RegSlot rs[2];
                uint8_t rt[2];
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    rt[j] = 0;
                    if (j < (int)d.nrec) {
                        const size_t b = (size_t)j * 7 * NS + y;
                        rt[j] = d.y_rtype[j * NS + y];
                        rs[j].av = d.y_rav[j * NS + y];
                        rs[j].dens = d.y_rdens[j * NS + y];
#pragma unroll
                        for (int k = 0; k < 3; k++) rs[j].st[k] = d.y_rstate[b + (size_t)k * NS];
                        rs[j].mem = d.y_rstate + b;
                        rs[j].stride = NS;
                        rs[j].a7 = parea;
                    }
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

// AMPA/NMDA/GABA-A receptors only: four wavefronts per SIMD fit in the register budget.
// kPre: at most two receptors per synapse, each chain step's loads issued together.
template <bool kPre>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_syn_post_rec(
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
    if (cn > 0.0f) dw -= 0.0001f * cn * d.y_pp1[y] * dt;
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
        float ca = src >= 0 ? (float)d.ca_hist[src] : (float)d.presum[d.y_prew[y]].ca_snap;
        float caf = ca / (0.0005f + ca);
        float prim = d.y_prim0[y];   // RRP_recovery + CaMKII(pre) * 0.02 + calcineurin(pre) * 0.01
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

}  // namespace bc
