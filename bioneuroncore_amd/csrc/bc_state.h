// bc_state.h — device-resident network state (structure of arrays) and the per-launch views
// the tick kernels receive by value.
//
// Segments are stored in "device order": connected islands (components of the
// IntraConnection graph) are contiguous and packed into tiles of at most kTileSeg segments /
// kTileConn connections; one workgroup owns one tile in the electrodiffusion kernel.
// Synapses are stored in "slot order": wavefront-interleaved for the post-side pass. Lane l of
// that pass walks the incoming synapses of segment walk[l] in synapse index order; lanes are
// sorted by chain length inside each wavefront, and step q of a wavefront's lanes occupies one
// contiguous run of slots (see k_syn_post).
#pragma once
#include <stdint.h>
#include <stddef.h>
#include "../../include/biocore_abi.h"

namespace bc {

// Tile size classes of the electrodiffusion kernel (segments = connections = lanes).
constexpr int kTileClasses = 3;
constexpr int kTileSize[kTileClasses] = {64, 128, 256};
constexpr int kTileSeg = 256;   // largest island one tile holds
constexpr int kNtCount = 15;

// Segment flag bits.
enum : uint8_t {
    SF_MYELIN = 1,        // carries MOD_MYELIN: ionotropic + listed channels blocked
    SF_BATH_NMDA = 2,     // carries an NMDA segment modifier: glutamate bath 1.0
    SF_BATH_5HT = 4,      // carries a 5-HT7 segment modifier: serotonin bath 0.0
    SF_WATCHED = 8,       // is the presynaptic segment of some synapse
    SF_NTREAD = 16,       // a modifier reads (receptors) or takes up (transporters) transmitters
    SF_NTX = 32,          // an incoming synapse carries a cleft transmitter other than glutamate/GABA
    SF_ER = 64,           // carries SERCA, IP3R or RyR: step 1 reads and writes its ER Ca2+ store
    SF_TT = 128,          // carries the TRPV1 receptor, somatic ATP synthesis, thermoregulation,
                          // CREB, NF-kB or a physical sensor: step 1 reads and writes its
                          // temperature and transcription signal
};

// Synapse flag bits (static).
enum : uint8_t {
    YF_INHIBITORY = 1,    // a post receptor is GABA-A / GABA-B / glycine
    YF_SKIP = 2,          // pre/post out of range or electrical: no chemical step runs
    YF_GAP = 4,           // electrical (gap junction); also YF_SKIP
    YF_GLIA = 8,          // its network has glia: glutamate, glycine and GABA are captured
};

// Tiles of one size class, launched together.
struct TileSet {
    const uint32_t* seg0;
    const uint32_t* nseg;
    const uint32_t* conn0;
    const uint32_t* nconn;
    const uint32_t* maxinc;   // largest incidence count of a segment in the tile
    uint32_t count;           // tiles in the set
};

// Per-tick summary of one presynaptic segment: everything the bouton/release computation of
// its outgoing synapses reads, including the V-only exponentials (64 bytes, one cache line).
struct PreSum {
    double V;      // V_m after step 2.5 (seed, bouton voltage)
    double exca;   // extracellular Ca
    double h1;     // Cav2.1 inactivation state[1] (last Cav2.1 modifier)
    double vn;     // 2F*V/RT of the divalent GHK for V rounded to float
    double edi;    // exp(-vn)
    float dens;    // Cav2.1 density (0 without Cav2.1)
    float mi;      // Cav2.1 m_inf(V)
    float h0;      // 1 / (1 + exp((V + 50) / 8)), the voltage factor of h_inf
    uint32_t has_cav;
    double ca_snap;  // intra Ca at the start of step 3 (RRP recovery without an earlier synapse)
};
static_assert(sizeof(PreSum) == 64, "PreSum layout");

// Stimuli as per-segment ordered lists. head[s] is ~0u, or the index in ent of a count word
// followed by that many entries (kind << 28 | payload index), sorted by kind in the order the
// tick applies them and, within a kind, by stimulus index - the reference's matching order.
enum : uint32_t { SK_CUR = 0, SK_RAMP = 1, SK_ION = 2, SK_LIG = 3, SK_VCLAMP = 4 };
struct StimDev {
    const uint32_t* head;   // [n]
    const uint32_t* ent;
    const CurrentStimulus* cur;
    const RampStimulus* ramp;
    const VoltageClamp* vc;
    const IonInjection* ion;
    const LigandInjection* lig;
    uint64_t inj_tick;      // one-shot injections apply on this tick only
    uint32_t any;           // some segment has a list
    uint32_t tick_shift;    // time windows are matched against tick + tick_shift (0: the CPU tick)
};

struct Net {
    // sizes
    uint32_t n;          // segments
    uint32_t npad;       // stride of segment SoA arrays
    uint32_t nslot;      // modifier slots per segment (max modifier_count)
    uint32_t ns;         // synapses
    uint32_t nrec;       // receptor slots per synapse (max post_receptor_count)
    uint32_t ntiles;
    uint32_t has_conn;   // any IntraConnection at all (step 2.5 runs)
    // Transmitter fields step 2.5 must diffuse. With synapses, step 3 rebuilds every segment's
    // extracellular transmitters from zero before anything reads them, so their diffusion is
    // dead; intra_nt stays +0.0 (and its diffusion an exact no-op) until something sources it.
    uint32_t nt_ex_live;
    uint32_t nt_in_live;
    // Residual-terminated Jacobi (an extension; off: the reference's fixed 50 sweeps)
    double jac_tol;
    int32_t jac_max;
    const uint8_t* isl_lane0;   // per segment: tile-local index of its island's first segment
    // Cable mode (an extension, the tick the adjoint differentiates): ion concentrations and
    // ATP held at their values, phi_e = 0; the axial current is the intracellular SG flux and
    // chord term at those concentrations. conn_scale (per connection, device order; null = 1)
    // multiplies the axial conductance and the intracellular ion fluxes of each connection.
    uint32_t cable;
    const float* conn_scale;
    float* xrec;         // cable mode: the Jacobi iterates of the tick [5][51][npad] (NULL: none)

    // segment dynamic state
    double* V;
    double* in;          // [7][npad]
    double* ex;          // [7][npad]
    float* atp;
    float* temp;
    float* txn;
    float* er;           // ER_Ca2, the ER Ca2+ store SERCA fills and IP3R/RyR release
    float* nt;           // neurotransmitter [15][npad]
    float* nt_in;        // intra_nt [15][npad]
    double* ibuf;
    uint32_t* spikes;

    const uint64_t* tick;   // device tick counter (advanced by k_tick_advance after each tick)
    StimDev stim;

    // segment static data
    const float* radius;
    const float* length;
    const uint32_t* segid;   // Segment.id of each record (thermal noise seeds its draw with it)
    const uint8_t* stype;
    const float* area;
    const double* vol;
    const float* cap;
    float* pkc;                 // state[0] of the first PKC modifier (0 if none); step 1 keeps it current
    const float* pka;
    const float* camp;
    float* pp1;                 // state[3] of the last PP1 modifier (0.5 if none); step 1 keeps it current
    float* camkii;              // state[0] of the last CaMKII modifier (0 if none); step 1 keeps it current
    float* camkii01;            // its state[0] + state[1] (0 if none)
    float* calcn;               // state[0] of the last calcineurin modifier (0 if none)
    const uint8_t* sflags;
    const int8_t* cav21_slot;   // last Cav2.1 modifier slot, -1 if none

    // modifiers [slot][npad] and states [slot][7][npad]
    const uint8_t* mcount;
    const uint8_t* mtype;
    const uint8_t* mav;
    const double* mdens;
    double* mstate;

    // electrodiffusion tiles (see TileSet); connections stored tile by tile
    const uint32_t* conn_ab;     // (a_local | b_local << 16), tile-local, global connection order
    const uint8_t* conn_intra;
    const uint32_t* csr_off;     // [n+1] per device segment
    const uint32_t* csr_ent;     // (tile-local connection << 1) | role (0 = a, 1 = b)

    // synapses, slot order
    const uint32_t* y_pre;       // device segment index
    const uint32_t* y_prew;      // index into the presynaptic summaries
    const uint32_t* watched;     // device segment of each presynaptic summary
    uint32_t nwatched;
    PreSum* presum;
    double* ca_snap;     // each summary's ca_snap again, compact (the vesicle pass gathers only it)
    const uint32_t* y_id;
    const uint8_t* y_flags;
    const float* y_delay;
    const float* y_pre_area;
    const float* y_post_area;
    const float* y_prel;
    const float* y_amount;
    const float* y_tau;
    const float* y_rrp_max;
    const float* y_prim0;        // RRP_recovery + CaMKII(pre) * 0.02 + calcineurin(pre) * 0.01
    const float* y_rrec;         // RRP_recovery alone when some segment carries CaMKII or
                                 // calcineurin (then the sum is formed each tick), else null
    const uint16_t* y_cleft_mask;
    const int32_t* y_src;        // Ca-history entry of the last earlier synapse whose post is my
                                 // pre (see ca_hist), -1 if none
    const uint32_t* hist_base;   // per segment: its first Ca-history entry (presynaptic segments)
    double* ca_hist;             // post intra Ca after each synapse onto a presynaptic segment, in
                                 // chain order per segment
    const uint32_t* walk;        // lane -> segment for the post-side pass: each 256-segment
                                 // window ordered by incoming-synapse count, longest first
    const uint32_t* post_len;    // per lane: incoming synapses of walk[lane]
    const uint32_t* wave_base;   // per 64-lane wavefront: its first slot
    uint8_t* y_rcount;
    const uint8_t* y_rtype;      // [nrec][ns]
    const uint8_t* y_rav;
    const double* y_rdens;
    double* y_rstate;            // [nrec][7][ns]
    float* y_cleft;              // [15][ns]
    float* y_weight;
    float* y_fac;
    float* y_dep;
    uint8_t* y_refr;
    float* y_rrp;
    float* y_bca;
    float* y_sca;
    float* y_camkii;
    float* y_calcn;
    float* y_isyn;               // receptor current of the synapse this tick (post pass -> plasticity)
    const float* y_pp1;          // PP1 activity of the post segment
    const uint32_t* y_postd;     // post segment of each slot when some segment carries PP1 (its
                                 // activity then read from pp1 each tick), else null
    float* y_relp;               // release probability of a deferred release this tick, 0 if none
    uint8_t* d_mask;             // occupied delay slots (bit s)
    uint64_t* d_arr;             // [4][ns]
    float* d_amt;                // [4][ns]
    uint8_t* d_nt;               // [4][ns]
    // presynaptic autoreceptors (null when no synapse has one): the release modulators of slot
    // y are entries y_ar[y] >> 4 .. + (y_ar[y] & 15) of the ar_* arrays, in record order
    const uint32_t* y_ar;
    const uint8_t* ar_type;
    const double* ar_dens;
    double* ar_st0;              // state[0]
    double* ar_st7;              // state[7]: the synapse's pre_area once a release check ran
    // eCB retrograde deposits (null when no post segment carries eCB synthesis): per slot, the
    // number of earlier chemical synapses onto the same post segment (EC_LAST: it is the last);
    // EC_NONE: the post segment has no eCB synthesis. seg_ecb: per segment, the slot index of
    // its first eCB synthesis modifier.
    const uint32_t* y_ecb;
    const uint8_t* seg_ecb;
    // gap junctions (null when the network has none; y_post also set for eCB deposits)
    const uint32_t* y_post;      // device post segment of each slot
    float* y_igap;               // junction current this tick (pA, into the post, out of the pre)
    const uint32_t* y_sid;       // synapse index of each slot
    const uint32_t* gap_off;     // [n+1] per segment: its electrical synapses as the presynaptic end
    const uint32_t* gap_slot;    //   their slots
    const uint32_t* gap_sid;     //   their synapse indices (ascending per segment)
    // glial capture (null when no network has glia): each chemical synapse of a network with
    // GLAST segments hands part of its cleft glutamate (0), glycine (7) and GABA (1) to the
    // glial segment nearest its post segment in depth
    float* y_cap;                // [3][ns] amount captured this tick (species 0, 7, 1)
    uint8_t* y_capm;             // bit q: species q was captured this tick
    const uint32_t* glia_off;    // [n+1] per segment: the synapses whose captures it receives
    const uint32_t* glia_slot;   //   their slots
    const uint32_t* glia_sid;    //   their synapse indices (ascending per segment)

    // Ordered pass (k_syn_ordered; null / 0 when no synapse needs it). A synapse whose effect
    // depends on another post chain's progress within step 3 - an immediate release whose
    // vesicle recovery reads a presynaptic Ca that an earlier synapse's receptors change, or a
    // gap junction behind such a synapse - makes its post chain (and the segments its effect
    // reaches) "ordered": the parallel post pass skips them, and one lane then walks every
    // event touching them in synapse-index order.
    const uint8_t* seg_ord;      // per segment: 1 = walked by the ordered pass
    const uint8_t* y_haz;        // per slot: HZ_RELEASE or HZ_GAP (resolved in the ordered pass)
    const uint32_t* ord_ev;      // events in synapse-index order: the slot
    const uint8_t* ord_fl;       //   EV_* flags
    const int32_t* ord_hist;     //   Ca-history entry written after the post side (-1: none)
    const uint32_t* ord_glia;    //   glial segment receiving the capture (EV_GLIA)
    const uint32_t* ord_post;    //   the post segment (device index)
    const uint32_t* ord_seg;     // the ordered segments, for step 4 after the events
    uint32_t ord_n, ord_nseg;

    float nt_diff[kNtCount];     // transmitter diffusion coefficients
};

enum : uint8_t { HZ_RELEASE = 1, HZ_GAP = 2 };
enum : uint32_t { EC_LAST = 1u << 31, EC_NONE = ~0u };
enum : uint8_t {
    EV_POST = 1,     // the post segment is ordered: cleft, captures, receptors, I_buffer, Ca history
    EV_PREGAP = 2,   // gap junction whose presynaptic end is ordered: its current leaves that end
    EV_GLIA = 4,     // the capture goes to an ordered glial segment
    EV_HAZ = 8,      // immediate release completed here (vesicle recovery with the Ca history)
    EV_GAPHAZ = 16,  // gap junction computed here, with both ends' Ca as of its position
};

}  // namespace bc
