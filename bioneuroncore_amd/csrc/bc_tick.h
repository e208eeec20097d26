// bc_tick.h — device code shared by the tick kernels, and the kernels' declarations.
//
//   diffusion_body    step 2.5 of one tile of whole islands, all state in LDS
//
// The kernels are defined in their own translation units, compiled separately:
//   bc_step1.hip      k_modifiers (step 1)
//   bc_diffusion.hip  k_diffusion (step 2.5 of 64/128/256-segment tiles)
//   bc_synapse.hip    step 3 and step 4: k_pre_summary, k_syn_pre, k_syn_post(_rec),
//                     k_syn_ordered, k_syn_plast, k_syn_rrp; k_tick_advance, k_stim_map
//   bc_bigisland.hip  k_diffusion_big (islands beyond a tile)
#pragma once
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
template <int TS, bool kRes, bool kRec>
__device__ __attribute__((always_inline)) inline void diffusion_body(const Net& d, const TileSet& ts, float dt) {
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


// ------------------------------------------------------------------ kernel declarations
template <bool kNt, bool kExt>
__global__ void k_modifiers(Net d, float dt, uint32_t first, uint32_t last);
template <int TS, bool kRes, bool kRec = false>
__global__ void k_diffusion(Net d, TileSet ts, float dt);
__global__ void k_pre_summary(Net d);
__global__ void k_syn_pre(Net d, float dt);
template <bool kPre>
__global__ void k_syn_post_rec(Net d, float dt);
__global__ void k_syn_post(Net d, float dt);
__global__ void k_syn_ordered(Net d, float dt);
__global__ void k_syn_plast(Net d, float dt);
__global__ void k_syn_rrp(Net d, float dt);
__global__ void k_tick_advance(uint64_t* tick);
__global__ void k_stim_map(const uint32_t* __restrict__ segs, const uint32_t* __restrict__ heads, uint32_t count,
                           uint32_t* __restrict__ head);

}  // namespace bc
