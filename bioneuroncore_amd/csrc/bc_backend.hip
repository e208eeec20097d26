// bc_backend.hip — host side of libbiocore_hip_backend.so: BION v14 loading, the device
// layout (segment tiles, synapse slot order, CSR lists), host<->device coherence, and the
// C ABI of BioNeuronCore's GPU backend (src/gpu_backend_cuda.cu:316-390) plus extensions.
#include <hip/hip_runtime.h>
#include <dlfcn.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <new>
#include <numeric>
#include <string>
#include <vector>
#include "../../include/biocore_abi.h"
#include "../../include/biocore_hip_backend.h"
#include "bc_tick.h"
#include "bc_bigisland.h"
#include "bc_adjoint.hip"

// Registered plugin functions (src/plugin_manager.cpp:37-75). The reference's CPU tick calls them
// in place of the built-in modifier current or release modulator of their type
// (include/kernel_dispatch.h:337-343); its device code never does. A network with an instance
// of such a type is counted as unsupported work at load (refused unless bchip_allow_unsupported),
// so a registered plugin cannot make the device silently depart from the CPU tick.
static std::mutex g_plugin_mu;
static ModifierFunc g_mod_table[128];
static ReleaseModulatorFunc g_rel_table[128];
struct PluginTypes {
    bool mod[256] = {}, rel[256] = {};
};
static PluginTypes plugin_types() {
    PluginTypes p;
    std::lock_guard<std::mutex> lk(g_plugin_mu);
    for (int t = 0; t < 128; t++) {
        p.mod[t] = g_mod_table[t] != nullptr;
        p.rel[t] = g_rel_table[t] != nullptr;
    }
    return p;
}

namespace bc {

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "biocore_hip: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                 \
            throw std::runtime_error("hip");                                             \
        }                                                                                \
    } while (0)

}  // namespace bc
#include "bc_bion.h"
namespace bc {

// ------------------------------------------------------------------ host-side static precompute
static inline bool seg_sphere(const Segment& s) { return s.segment_type == 1 || s.segment_type == 3 || s.segment_type == 4; }
static inline float seg_area(const Segment& s) {
    return seg_sphere(s) ? 4.0f * kPi * s.radius * s.radius : 2.0f * kPi * s.radius * s.length;
}
static inline double seg_vol(const Segment& s) {
    if (seg_sphere(s)) return (4.0 / 3.0) * kPi * (double)s.radius * (double)s.radius * (double)s.radius;
    return kPi * (double)s.radius * (double)s.radius * (double)s.length;
}
static inline float seg_cap(const Segment& s) {   // run_tick_common.h:216-231
    float c = seg_area(s) * 0.01f;
    if (c < 1e-4f) c = 1e-4f;
    for (uint8_t j = 0; j < s.modifier_count && j < 16; j++)
        if (s.modifiers[j].type == 4) {
            float gr = s.modifiers[j].state[0];
            if (gr <= 0.0f || gr >= 1.0f) gr = 0.6f;
            float factor = 100.0f * (1.0f / gr - 1.0f) / (1.0f / 0.6f - 1.0f);
            c /= factor;
            break;
        }
    return c;
}
static inline uint8_t host_nt_of_receptor(uint8_t rt) {
    if (rt == 11 || rt == 17) return 1;
    if (rt == 12) return 7;
    if (rt == 14) return 3;
    if (rt == 13) return 4;
    if (rt == 44 || rt == 22) return 2;
    return 0;
}

struct Sim;
static std::mutex g_reg_mu;

struct Sim {
    int device = 0;
    hipStream_t stream = nullptr;
    BionData host{};
    bool loaded = false;
    Net d{};
    std::vector<void*> dmem;
    std::vector<uint32_t> dev2host, host2dev, slot2syn, conn2dev;
    uint32_t ntiles = 0;
    uint32_t class_tile0[kTileClasses] = {}, class_ntiles[kTileClasses] = {};
    std::vector<uint32_t> h_tile64_seg0;   // first device segment of each 64-segment tile, then the end
    // Second stream of the tick: step 2.5 of the 64-segment tiles, chunk by chunk, while step 1
    // of the next chunk runs on `stream` (sched.step1_chunks > 1). Default priority, as `stream`.
    hipStream_t stream2 = nullptr;
    static constexpr uint32_t kMaxChunks = 16;
    hipEvent_t fork_ev[kMaxChunks + 1] = {};
    BigSet big{};   // islands larger than a tile (k_diffusion_big); big.count = 0 without
    TileSet tiles[kTileClasses] = {};
    size_t lds_bytes = 0;
    uint64_t tick = 0;
    uint64_t ticks_run = 0;
    GpuStimulusData* stim_host = nullptr;   // last gpu_backend_upload_stimuli payload
    bool stim_present = false, stim_dirty = false;
    uint64_t* tick_dev = nullptr;
    // stimulus lists (see StimDev): payloads, list words and the segments whose head is set
    std::vector<CurrentStimulus> st_cur;
    std::vector<RampStimulus> st_ramp;
    std::vector<VoltageClamp> st_vc;
    std::vector<IonInjection> st_ion;
    std::vector<LigandInjection> st_lig;
    std::vector<uint32_t> st_ent, st_segs[2], st_heads;
    int st_cur_set = 0;
    struct DBuf {
        void* p = nullptr;
        size_t cap = 0;
    };
    DBuf sb_cur, sb_ramp, sb_vc, sb_ion, sb_lig, sb_ent, sb_segs[2], sb_heads;
    DBuf sb_segord, sb_yhaz, sb_oev, sb_ofl, sb_ohist, sb_oglia, sb_opost, sb_oseg;
    uint32_t* stim_head = nullptr;
    std::vector<std::pair<uint64_t, uint32_t>> id2dev;   // (network << 32 | Segment.id, device index), sorted
    // Networks of a batch (a plain load is one network): host index offsets into the union,
    // and each network's stimuli (unbounded; Segment.id matched within the network).
    std::vector<uint32_t> net_seg_off{0, 0}, net_conn_off{0, 0}, net_syn_off{0, 0};
    struct NetStim {
        std::vector<CurrentStimulus> cur;
        std::vector<RampStimulus> ramp;
        std::vector<VoltageClamp> vc;
        std::vector<IonInjection> ion;
        std::vector<LigandInjection> lig;
    };
    std::vector<NetStim> net_stim{NetStim{}};
    bool inj_committed = false;   // injections are in the device lists (cleared after their tick)
    // Launch schedule (bchip_autotune picks it; results do not depend on it).
    BchipSchedule sched{256, 1, 8192, 256, 256, 1};
    // Captured tick sequence (bchip_run_ticks): replayed while the launch inputs are unchanged.
    hipGraphExec_t graph = nullptr;
    uint32_t graph_ticks = 0;
    Net graph_net{};
    float graph_dt = 0.0f;
    BchipSchedule graph_sched{};
    bool use_graphs = true;
    bool cable = false;               // bchip_set_cable_mode
    float* conn_scale_dev = nullptr;
    uint64_t adj_record_cap = ~0ull;   // bchip_set_adjoint_memory
    uint32_t stim_tick_shift = 0;      // bchip_set_stimulus_convention  // bchip_set_coupling_scale (device order), in dmem
    bool jac_res = false;   // residual-terminated Jacobi (bchip_set_jacobi)
    double jac_tol = 0.0;
    int jac_max = 50;
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_ticks = 0;
    }
    int coherence = BCHIP_COHERENCE_COMPAT;
    bool host_stale = false;
    uint64_t unsupported_mod = 0, unsupported_syn = 0, hazard_syn = 0;
    bool allow_unsupported = false;   // bchip_allow_unsupported
    uint32_t* host2dev_dev = nullptr;   // device copies of host2dev and conn2dev (the adjoint's
    uint32_t* conn2dev_dev = nullptr;   // soma mapping and gradient order)
    // Ordered pass planning (see plan_ordered): per slot the Ca-history source, the effective
    // RRP_max, flags and glial target; per segment the Ca-history base and flags; the chemical
    // synapses that become hazards when their release is immediate, and the gap junction hazards.
    std::vector<int32_t> h_src;
    std::vector<float> h_rmax;
    std::vector<uint8_t> h_yflags, h_sflags;
    char* stage = nullptr;        // pinned staging of the COMPAT transfers (hipHostMalloc)
    size_t stage_bytes = 0;
    std::vector<uint32_t> h_ar;   // per slot: first autoreceptor << 4 | count
    size_t nar = 0;               // autoreceptors on the device
    std::vector<uint32_t> h_glia, h_hist_base, h_cand, h_gaphaz;
    float ord_dt = NAN;   // dt the ordered pass was planned for (NaN: not planned)
    uint32_t ord_events = 0, ord_segments = 0;
    std::vector<uint8_t> o_segord, o_yhaz, o_fl;
    std::vector<uint32_t> o_ev, o_glia, o_post, o_seg;
    std::vector<int32_t> o_hist;
    bool receptors_only = true;   // every receptor slot is AMPA, NMDA or GABA-A
    bool nt_read = false;         // some segment modifier reads or takes up transmitters
    bool protein_map = false;     // nucleus proteins set the release parameters (see release_params)
    bool ext_mods = false;        // some segment modifier is outside the core channel set
    std::vector<uint8_t> seg_dispatched;   // [n*32] slot ran at least once (state[7] written)
    hipEvent_t ev[8] = {};
    float last_ms[8] = {};
    int timing = 0;

    template <class T>
    T* dalloc(size_t n) {
        void* p = nullptr;
        size_t bytes = std::max<size_t>(n * sizeof(T), 64);
        HIPCHK(hipMalloc(&p, bytes));
        HIPCHK(hipMemsetAsync(p, 0, bytes, stream));
        dmem.push_back(p);
        return static_cast<T*>(p);
    }
    template <class T>
    T* dupload(const std::vector<T>& v) {
        T* p = dalloc<T>(v.size());
        if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return p;
    }
    void free_device() {
        for (void* p : dmem) (void)hipFree(p);
        dmem.clear();
        memset(&d, 0, sizeof(d));
    }
    // Device copy of v in b (grown as needed); the copy reads v asynchronously, so v must stay
    // unchanged until the stream has passed it.
    template <class T>
    T* dput(DBuf& b, const std::vector<T>& v) {
        size_t bytes = std::max<size_t>(v.size() * sizeof(T), 64);
        if (b.cap < bytes) {
            if (b.p) HIPCHK(hipStreamSynchronize(stream));
            if (b.p) (void)hipFree(b.p);
            b.p = nullptr;
            b.cap = 0;
            HIPCHK(hipMalloc(&b.p, bytes + bytes / 2));
            b.cap = bytes + bytes / 2;
        }
        if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        return static_cast<T*>(b.p);
    }
    void free_stim_buffers() {
        for (DBuf* b : {&sb_cur, &sb_ramp, &sb_vc, &sb_ion, &sb_lig, &sb_ent, &sb_segs[0], &sb_segs[1], &sb_heads,
                        &sb_segord, &sb_yhaz, &sb_oev, &sb_ofl, &sb_ohist, &sb_oglia, &sb_opost, &sb_oseg}) {
            if (b->p) (void)hipFree(b->p);
            b->p = nullptr;
            b->cap = 0;
        }
        st_segs[0].clear();
        st_segs[1].clear();
    }
    void free_stage() {
        if (stage) (void)hipHostFree(stage);
        stage = nullptr;
        stage_bytes = 0;
    }
    ~Sim() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        free_device();
        free_stage();
        free_stim_buffers();
        delete stim_host;
        if (stream2) (void)hipStreamSynchronize(stream2);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : fork_ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        if (stream2) (void)hipStreamDestroy(stream2);
        free_bion(&host);
    }
};

// Union-find over all connections: islands of the connection graph.
static uint32_t uf_find(std::vector<uint32_t>& p, uint32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

// Release parameters of every synapse slot: RRP_max, release probability and the vesicle
// priming rate RRP_recovery + CaMKII(pre) 0.02 + calcineurin(pre) 0.01 (kernel_dispatch.h:
// 612-631). With protein mapping on (the reference applies it while its consolidation runs,
// simulation_host.cpp:407-414, kernel_dispatch.h:651-663), a chemical synapse whose presynaptic
// neuron has a nucleus takes RRP_max, the release probability and RRP_recovery from that
// nucleus's Munc13-1, synaptotagmin-1 and complexin levels; the nucleus state does not change
// on the device, so the mapped values are constant and computed here once.
static void protein_mapped(const Synapse& y, const BionData& B, float& rmax, float& prel, float& rrec) {
    rmax = y.RRP_max; prel = y.release_probability; rrec = y.RRP_recovery;
    if (y.type == 1 || y.type == 2 || !B.nuclei) return;
    const uint32_t nid = B.segments[y.pre_segment].neuron_id;
    if (nid >= B.num_nuclei) return;
    const float* pl = B.nuclei[nid].protein_level;
    float munc13 = pl[42];
    rmax = 10.0f + munc13 * 40.0f;
    float syt1 = pl[40];
    prel = 0.3f + syt1 * 0.7f;
    float cplx1 = pl[44];
    rrec = 0.02f + munc13 * 0.08f - cplx1 * 0.01f;
    if (rrec < 0.005f) rrec = 0.005f;
    if (rrec > 0.2f) rrec = 0.2f;
}
static void release_params(const Sim& S, std::vector<float>& rmax, std::vector<float>& prel,
                           std::vector<float>& prim0, std::vector<float>& rrecs) {
    const BionData& B = S.host;
    const uint32_t ns = B.num_synapses;
    rmax.resize(ns); prel.resize(ns); prim0.resize(ns); rrecs.resize(ns);
    for (uint32_t k = 0; k < ns; k++) {
        const Synapse& y = B.synapses[S.slot2syn[k]];
        float rrec;
        if (S.protein_map) protein_mapped(y, B, rmax[k], prel[k], rrec);
        else { rmax[k] = y.RRP_max; prel[k] = y.release_probability; rrec = y.RRP_recovery; }
        const Segment& pre = B.segments[y.pre_segment];
        float ck = 0.0f, cn = 0.0f;
        for (uint32_t j = 0; j < pre.modifier_count; j++) {
            if (pre.modifiers[j].type == 82) ck = pre.modifiers[j].state[0];
            if (pre.modifiers[j].type == 83) cn = pre.modifiers[j].state[0];
        }
        float a = rrec + ck * 0.02f;
        prim0[k] = a + cn * 0.01f;
        rrecs[k] = rrec;
    }
}

static int build_device(Sim& S) {
    BionData& B = S.host;
    const uint32_t n = B.num_segments, nc = B.num_connections, ns = B.num_synapses;
    Segment* sg = B.segments;
    IntraConnection* cn = B.connections;
    Synapse* sy = B.synapses;

    // ---- islands, device order, tiles ----
    std::vector<uint32_t> par(n);
    std::iota(par.begin(), par.end(), 0u);
    for (uint32_t i = 0; i < nc; i++) {
        uint32_t a = uf_find(par, cn[i].segment_a), b = uf_find(par, cn[i].segment_b);
        if (a != b) par[std::max(a, b)] = std::min(a, b);
    }
    std::vector<uint32_t> root(n), comp_size(n, 0), comp_conns(n, 0);
    for (uint32_t i = 0; i < n; i++) { root[i] = uf_find(par, i); comp_size[root[i]]++; }
    for (uint32_t i = 0; i < nc; i++) comp_conns[root[cn[i].segment_a]]++;
    // size class of each island: the smallest tile that holds its segments and connections
    std::vector<int> comp_class(n, -1);
    std::vector<uint32_t> comps;
    for (uint32_t i = 0; i < n; i++) {
        if (root[i] != i) continue;
        int cls = kTileClasses;   // past every tile class: a big island (one workgroup each)
        for (int k = 0; k < kTileClasses && cls == kTileClasses; k++)
            if (comp_size[i] <= (uint32_t)kTileSize[k] && comp_conns[i] <= (uint32_t)kTileSize[k]) cls = k;
        comp_class[i] = cls;
        comps.push_back(i);
    }
    // device order: islands by (size class, smallest segment index), segments in index order
    std::stable_sort(comps.begin(), comps.end(), [&](uint32_t a, uint32_t b) { return comp_class[a] < comp_class[b]; });
    std::vector<uint32_t> comp_start(n, 0);
    S.dev2host.assign(n, 0);
    S.host2dev.assign(n, 0);
    {
        uint32_t pos = 0;
        for (uint32_t r : comps) { comp_start[r] = pos; pos += comp_size[r]; }
        std::vector<uint32_t> fill(n, 0);
        for (uint32_t i = 0; i < n; i++) {
            uint32_t r = root[i];
            uint32_t dpos = comp_start[r] + fill[r]++;
            S.dev2host[dpos] = i;
            S.host2dev[i] = dpos;
        }
    }
    std::vector<uint32_t> tile_seg0, tile_nseg, tile_conn0, tile_nconn, seg_tile(n);
    for (int k = 0; k < kTileClasses; k++) S.class_tile0[k] = S.class_ntiles[k] = 0;
    {
        uint32_t cur_s = 0, cur_c = 0, start = 0;
        int cur_cls = -1;
        auto flush = [&]() {
            if (!cur_s) return;
            tile_seg0.push_back(start); tile_nseg.push_back(cur_s); tile_nconn.push_back(cur_c);
            S.class_ntiles[cur_cls]++;
            start += cur_s; cur_s = 0; cur_c = 0;
        };
        for (uint32_t r : comps) {
            int cls = comp_class[r];
            if (cls == kTileClasses) break;   // the big islands follow
            uint32_t cap = kTileSize[cls];
            if (cls != cur_cls) {
                flush();
                cur_cls = cls;
                S.class_tile0[cls] = (uint32_t)tile_seg0.size();
            }
            if (cur_s + comp_size[r] > cap || cur_c + comp_conns[r] > cap) flush();
            cur_s += comp_size[r]; cur_c += comp_conns[r];
        }
        flush();
    }
    S.ntiles = (uint32_t)tile_seg0.size();
    // islands are placed by size class, the 64-segment class first, from device segment 0
    S.h_tile64_seg0.assign(tile_seg0.begin(), tile_seg0.begin() + S.class_ntiles[0]);
    S.h_tile64_seg0.push_back(S.class_ntiles[0] ? tile_seg0[S.class_ntiles[0] - 1] + tile_nseg[S.class_ntiles[0] - 1] : 0);
    // each big island is a pseudo-tile past the real ones (placement and incidence lists below
    // treat it as a tile; only k_diffusion_big runs on it)
    uint32_t nbig = 0;
    for (uint32_t r : comps)
        if (comp_class[r] == kTileClasses) {
            tile_seg0.push_back(comp_start[r]);
            tile_nseg.push_back(comp_size[r]);
            tile_nconn.push_back(comp_conns[r]);
            nbig++;
        }
    const uint32_t nall = S.ntiles + nbig;
    for (uint32_t t = 0; t < nall; t++)
        for (uint32_t k = 0; k < tile_nseg[t]; k++) seg_tile[tile_seg0[t] + k] = t;
    std::vector<uint8_t> isl_lane0(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t di = S.host2dev[i];
        isl_lane0[di] = (uint8_t)(comp_start[root[i]] - tile_seg0[seg_tile[di]]);
    }
    tile_conn0.resize(nall);
    {
        uint32_t acc = 0;
        for (uint32_t t = 0; t < nall; t++) { tile_conn0[t] = acc; acc += tile_nconn[t]; }
    }
    // connections placed per tile in global index order
    std::vector<uint32_t> conn_ab(nc), conn_local(nc);
    S.conn2dev.assign(nc, 0);
    std::vector<uint8_t> conn_intra(nc);
    {
        std::vector<uint32_t> fill(nall, 0);
        for (uint32_t i = 0; i < nc; i++) {
            uint32_t da = S.host2dev[cn[i].segment_a], db = S.host2dev[cn[i].segment_b];
            uint32_t t = seg_tile[da];
            uint32_t loc = fill[t]++;
            conn_local[i] = loc;
            uint32_t pos = tile_conn0[t] + loc;
            S.conn2dev[i] = pos;
            conn_ab[pos] = (da - tile_seg0[t]) | ((db - tile_seg0[t]) << 16);
            conn_intra[pos] = cn[i].is_intracellular ? 1 : 0;
        }
    }
    // per-segment incidence lists in global connection order (role a before role b)
    std::vector<uint32_t> csr_off(n + 1, 0), csr_ent;
    {
        std::vector<uint32_t> deg(n, 0);
        for (uint32_t i = 0; i < nc; i++) { deg[S.host2dev[cn[i].segment_a]]++; deg[S.host2dev[cn[i].segment_b]]++; }
        for (uint32_t i = 0; i < n; i++) csr_off[i + 1] = csr_off[i] + deg[i];
        csr_ent.resize(csr_off[n]);
        std::vector<uint32_t> fill(n, 0);
        for (uint32_t i = 0; i < nc; i++) {
            uint32_t da = S.host2dev[cn[i].segment_a], db = S.host2dev[cn[i].segment_b];
            csr_ent[csr_off[da] + fill[da]++] = (conn_local[i] << 1) | 0u;
            csr_ent[csr_off[db] + fill[db]++] = (conn_local[i] << 1) | 1u;
        }
    }
    // per tile: the largest incidence count of its segments (selects the sweep variant)
    std::vector<uint32_t> tile_maxinc(nall, 0);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t& m = tile_maxinc[seg_tile[i]];
        m = std::max<uint32_t>(m, csr_off[i + 1] - csr_off[i]);
    }

    // ---- segment SoA ----
    const uint32_t npad = (n + 63) & ~63u;
    uint32_t nslot = 1;
    for (uint32_t i = 0; i < n; i++) nslot = std::max<uint32_t>(nslot, sg[i].modifier_count);
    Net& d = S.d;
    d.n = n; d.npad = npad; d.nslot = nslot; d.ns = ns; d.ntiles = S.ntiles; d.has_conn = nc > 0;
    d.nt_ex_live = ns == 0;
    d.nt_in_live = 1;   // until the first upload of the state decides
    std::vector<float> radius(npad, 0), length(npad, 0), area(npad, 0), cap(npad, 1), pkc(npad, 0), pka(npad, 0),
        camp(npad, 0.1f), pp1(npad, 0.5f), camkii(npad, 0), calcn(npad, 0),
        camkii01(npad, 0);
    std::vector<double> vol(npad, 0);
    std::vector<uint8_t> stype(npad, 0), sflags(npad, 0), mcount(npad, 0);
    std::vector<uint32_t> segid(npad, 0);
    std::vector<int8_t> cav21(npad, -1);
    std::vector<uint8_t> mtype((size_t)nslot * npad, 0), mav((size_t)nslot * npad, 0);
    std::vector<double> mdens((size_t)nslot * npad, 0.0);
    S.unsupported_mod = 0;
    const PluginTypes plug = plugin_types();
    S.nt_read = false;
    S.ext_mods = false;
    bool any_ck = false, any_pp1 = false;   // CaMKII / calcineurin, PP1 on some segment
    for (uint32_t di = 0; di < n; di++) {
        const Segment& s = sg[S.dev2host[di]];
        radius[di] = s.radius; length[di] = s.length; stype[di] = s.segment_type; segid[di] = s.id;
        area[di] = seg_area(s); vol[di] = seg_vol(s); cap[di] = seg_cap(s);
        mcount[di] = s.modifier_count;
        bool seen_pkc = false, seen_pka = false;
        uint8_t fl = 0;
        for (uint32_t j = 0; j < s.modifier_count; j++) {
            const Modifier& m = s.modifiers[j];
            mtype[(size_t)j * npad + di] = m.type;
            mav[(size_t)j * npad + di] = m.active_vars;
            mdens[(size_t)j * npad + di] = m.density;
            if (m.type == 4) fl |= SF_MYELIN;
            if (m.type == 9) fl |= SF_BATH_NMDA;
            if (m.type == 43) fl |= SF_BATH_5HT;
            if (m.type == 137 && !seen_pkc) { pkc[di] = m.state[0]; seen_pkc = true; }
            if (m.type == 135 && !seen_pka) { pka[di] = m.state[1]; camp[di] = m.state[0]; seen_pka = true; }
            if (m.type == 113) pp1[di] = m.state[3];
            if (m.type == 82) {
                camkii[di] = m.state[0];
                camkii01[di] = m.state[0] + m.state[1];
            }
            if (m.type == 83) calcn[di] = m.state[0];
            if (m.type == 82 || m.type == 83) any_ck = true;
            if (m.type == 113) any_pp1 = true;
            if (m.type == M_CAV21) cav21[di] = (int8_t)j;
            if (modifier_reads_nt(m.type)) fl |= SF_NTREAD;
            if (m.type == 126 || m.type == 127 || m.type == 128) fl |= SF_ER;
            if (m.type == 77 || m.type == 78 || m.type == 79 || m.type == 16 || m.type == 88 || m.type == 6)
                fl |= SF_TT;
            if (!modifier_core(m.type) && modifier_supported(m.type)) S.ext_mods = true;
            if (!modifier_supported(m.type) || plug.mod[m.type]) S.unsupported_mod++;
            // G-protein receptors read state[active_vars + 1]: past the record's state array
            if (gprotein_type(m.type) && m.active_vars >= 7) S.unsupported_mod++;
            // PMCA's Hill powers are exact squares only for n = 2 and a float-valued K_m
            if (m.type == M_PMCA && ((m.state[2] > 0.0f && m.state[2] != 2.0) ||
                                     (m.state[1] > 0.0f && (double)(float)m.state[1] != m.state[1])))
                S.unsupported_mod++;
        }
        sflags[di] = fl;
        if (fl & SF_NTREAD) S.nt_read = true;
    }
    // ---- synapses, slot order ----
    uint32_t nrec = 1;
    for (uint32_t i = 0; i < ns; i++) nrec = std::max<uint32_t>(nrec, sy[i].post_receptor_count);
    d.nrec = nrec;
    // Incoming synapses of each device segment, in synapse index order (CSR).
    std::vector<uint32_t> post_off(n + 1, 0), by_post(ns);
    for (uint32_t i = 0; i < ns; i++) post_off[S.host2dev[sy[i].post_segment] + 1]++;
    for (uint32_t i = 0; i < n; i++) post_off[i + 1] += post_off[i];
    {
        std::vector<uint32_t> fill(post_off.begin(), post_off.end() - 1);
        for (uint32_t i = 0; i < ns; i++) by_post[fill[S.host2dev[sy[i].post_segment]]++] = i;
    }
    // Lanes of the post-side pass: each 256-segment window ordered by chain length, longest
    // first, so in every 64-lane wavefront the lanes still walking at step q are a prefix.
    std::vector<uint32_t> walk(n), post_len(n);
    std::iota(walk.begin(), walk.end(), 0u);
    for (uint32_t w0 = 0; w0 < n; w0 += 256) {
        uint32_t w1 = std::min<uint32_t>(n, w0 + 256);
        std::stable_sort(walk.begin() + w0, walk.begin() + w1, [&](uint32_t a, uint32_t b) {
            return post_off[a + 1] - post_off[a] > post_off[b + 1] - post_off[b];
        });
    }
    for (uint32_t l = 0; l < n; l++) post_len[l] = post_off[walk[l] + 1] - post_off[walk[l]];
    // Slot order: wavefront by wavefront; inside one, step q of every lane still walking, lane by
    // lane. Lane j of wavefront g reads step q at wave_base[g] + (steps before q) + j, so each
    // step of the walk touches one contiguous run of slots.
    const uint32_t nwave = (n + 63) / 64;
    std::vector<uint32_t> wave_base(nwave + 1, 0);
    S.slot2syn.assign(ns, 0);
    {
        uint32_t off = 0;
        for (uint32_t g = 0; g < nwave; g++) {
            wave_base[g] = off;
            const uint32_t l0 = g * 64, l1 = std::min<uint32_t>(n, l0 + 64);
            for (uint32_t q = 0;; q++) {
                uint32_t cnt = 0;
                for (uint32_t l = l0; l < l1 && post_len[l] > q; l++, cnt++)
                    S.slot2syn[off + cnt] = by_post[post_off[walk[l]] + q];
                if (cnt == 0) break;
                off += cnt;
            }
        }
        wave_base[nwave] = off;
        if (off != ns) return BION_ERR_ALLOC;   // unreachable: every synapse has one slot
    }
    std::vector<uint32_t> syn2slot(ns);
    for (uint32_t k = 0; k < ns; k++) syn2slot[S.slot2syn[k]] = k;
    // Ca history: one entry per synapse onto a presynaptic segment, each such segment's entries
    // in its chain order; src = the entry of the last earlier synapse onto my pre segment.
    std::vector<uint8_t> is_pre(n, 0);
    for (uint32_t i = 0; i < ns; i++) is_pre[S.host2dev[sy[i].pre_segment]] = 1;
    std::vector<uint32_t> hist_base(n, 0), chain_rank(ns, 0);
    uint32_t nhist = 0;
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t q = post_off[i]; q < post_off[i + 1]; q++) chain_rank[by_post[q]] = q - post_off[i];
        if (is_pre[i]) { hist_base[i] = nhist; nhist += post_off[i + 1] - post_off[i]; }
    }
    std::vector<int32_t> src(ns, -1);
    {
        std::vector<int64_t> last_onto(n, -1);
        for (uint32_t i = 0; i < ns; i++) {
            int64_t j = last_onto[sy[i].pre_segment];
            src[syn2slot[i]] = j >= 0 ? (int32_t)(hist_base[S.host2dev[sy[j].post_segment]] + chain_rank[j]) : -1;
            last_onto[sy[i].post_segment] = i;
        }
    }
    // Glia: per network, its first 256 segments carrying GLAST (simulation_host.cpp:350-354).
    // Every chemical synapse of such a network captures into the glial segment nearest its post
    // segment in depth, a strictly nearer one replacing the first (simulation_host.cpp:477-482).
    std::vector<uint32_t> y_glia(ns, UINT32_MAX);   // per slot: device index of that glial segment
    uint32_t nglia_syn = 0;
    for (size_t k = 0; k + 1 < S.net_seg_off.size(); k++) {
        const uint32_t s0 = S.net_seg_off[k], s1 = S.net_seg_off[k + 1];
        std::vector<uint32_t> glia;
        for (uint32_t h = s0; h < s1 && glia.size() < 256; h++)
            for (uint32_t j = 0; j < sg[h].modifier_count; j++)
                if (sg[h].modifiers[j].type == M_GLAST) { glia.push_back(h); break; }
        if (glia.empty()) continue;
        std::vector<uint32_t> best_of(s1 - s0, UINT32_MAX);   // per post segment
        for (uint32_t i = S.net_syn_off[k]; i < S.net_syn_off[k + 1]; i++) {
            const Synapse& y = sy[i];
            if (y.type == 1 || y.type == 2 || y.pre_segment >= n || y.post_segment >= n) continue;
            uint32_t& b = best_of[y.post_segment - s0];
            if (b == UINT32_MAX) {
                const float pz = sg[y.post_segment].position.z;
                float best = 1e30f;
                b = glia[0];
                for (uint32_t g : glia) {
                    const float dz = fabsf(sg[g].position.z - pz);
                    if (dz < best) { best = dz; b = g; }
                }
            }
            y_glia[syn2slot[i]] = S.host2dev[b];
            nglia_syn++;
        }
    }
    std::vector<uint32_t> y_pre(ns), y_id(ns), y_prew(ns), watched;
    std::vector<int64_t> watch_idx(n, -1);
    std::vector<uint8_t> y_flags(ns, 0), y_rcount(ns, 0), y_refr(ns, 0);
    std::vector<float> y_delay(ns), y_pa(ns), y_qa(ns), y_prel(ns), y_amt(ns), y_tau(ns), y_rmax(ns);
    std::vector<float> y_w(ns), y_fac(ns), y_dep(ns), y_rrp(ns), y_bca(ns), y_sca(ns), y_ck(ns), y_cn(ns);
    std::vector<float> y_pp1(ns), y_prim0(ns);
    std::vector<uint16_t> y_mask(ns, 0);
    std::vector<uint8_t> y_rtype((size_t)nrec * ns, 0), y_rav((size_t)nrec * ns, 0);
    std::vector<double> y_rdens((size_t)nrec * ns, 0.0), y_rstate((size_t)nrec * 7 * ns, 0.0);
    std::vector<float> y_cleft((size_t)15 * ns, 0.0f);
    std::vector<uint32_t> y_ar(ns, 0);   // autoreceptors: first << 4 | count
    std::vector<uint8_t> ar_type;
    std::vector<double> ar_dens, ar_st0, ar_st7;
    S.unsupported_syn = 0;
    S.hazard_syn = 0;
    S.h_cand.clear();
    S.h_gaphaz.clear();
    S.ord_dt = NAN;
    S.receptors_only = true;
    uint32_t ngap = 0;
    for (uint32_t k = 0; k < ns; k++) {
        const Synapse& y = sy[S.slot2syn[k]];
        y_pre[k] = S.host2dev[y.pre_segment];
        sflags[y_pre[k]] |= SF_WATCHED;
        if (watch_idx[y_pre[k]] < 0) { watch_idx[y_pre[k]] = (int64_t)watched.size(); watched.push_back(y_pre[k]); }
        y_prew[k] = (uint32_t)watch_idx[y_pre[k]];
        y_id[k] = y.id;
        uint8_t fl = 0;
        if (y.type == 1 || y.type == 2) { fl |= YF_SKIP | YF_GAP; ngap++; }
        if (y_glia[k] != UINT32_MAX) fl |= YF_GLIA;
        if (y.pre_autoreceptor_count > 0) {
            y_ar[k] = (uint32_t)(ar_type.size() << 4) | y.pre_autoreceptor_count;
            for (uint32_t j = 0; j < y.pre_autoreceptor_count; j++) {
                const Modifier& m = y.pre_autoreceptors[j];
                if (plug.rel[m.type]) S.unsupported_mod++;   // a plugin modulator replaces the built-in one
                ar_type.push_back(m.type);
                ar_dens.push_back(m.density);
                ar_st0.push_back(m.state[0]);
                ar_st7.push_back(m.state[7]);
            }
        }
        for (uint32_t j = 0; j < y.post_receptor_count; j++) {
            uint8_t rt = y.post_receptors[j].type;
            if (rt == 11 || rt == 17 || rt == 12) fl |= YF_INHIBITORY;
            y_rtype[(size_t)j * ns + k] = rt;
            y_rav[(size_t)j * ns + k] = y.post_receptors[j].active_vars;
            y_rdens[(size_t)j * ns + k] = y.post_receptors[j].density;
            for (int q = 0; q < 7; q++) y_rstate[((size_t)j * 7 + q) * ns + k] = y.post_receptors[j].state[q];
            if (!modifier_supported(rt) || plug.mod[rt]) S.unsupported_mod++;
            if (rt == 86) S.unsupported_mod++;   // eCB synthesis as a receptor would read the post's DAG
            // the ER store and PLC as receptors would read the post's ER, IP3 and Gq pools
            if (rt == 126 || rt == 127 || rt == 128 || rt == 136) S.unsupported_mod++;
            // and somatic ATP, thermoregulation and CREB the post's temperature and transcription
            if (rt == 77 || rt == 78 || rt == 79 || rt == 16 || rt == 88) S.unsupported_mod++;
            // and the cascades that read or write sibling modifiers of the post segment
            if (rt == 87 || rt == 90 || rt == 91 || rt == 92) S.unsupported_mod++;
            // and the cascades whose activity the post segment's readers take from its own slots
            if (rt == 82 || rt == 83 || rt == 113 || rt == 137) S.unsupported_mod++;
            // and the sensors, which read the post's ATP, temperature, length and id or its light sensor
            if (rt == 6 || rt == 94 || rt == 95 || rt == 108 || rt == 111 || (rt >= 115 && rt <= 117) || rt == 120)
                S.unsupported_mod++;
            if (gprotein_type(rt) && y.post_receptors[j].active_vars >= 7) S.unsupported_mod++;
            if (rt != M_AMPA && rt != M_NMDA && rt != M_GABAA) S.receptors_only = false;
        }
        y_flags[k] = fl;
        y_rcount[k] = y.post_receptor_count;
        y_delay[k] = y.delay; y_pa[k] = y.pre_area; y_qa[k] = y.post_area; y_prel[k] = y.release_probability;
        y_amt[k] = y.release_amount; y_tau[k] = y.tau_clearance; y_rmax[k] = y.RRP_max;
        y_w[k] = y.weight; y_fac[k] = y.facilitation; y_dep[k] = y.depression; y_refr[k] = y.refractory;
        y_rrp[k] = y.RRP; y_bca[k] = y.bouton_Ca; y_sca[k] = y.spine_Ca; y_ck[k] = y.spine_camkii;
        y_cn[k] = y.spine_calcineurin;
        y_pp1[k] = pp1[S.host2dev[y.post_segment]];
        {
            const uint32_t p = S.host2dev[y.pre_segment];
            float a = y.RRP_recovery + camkii[p] * 0.02f;
            y_prim0[k] = a + calcn[p] * 0.01f;
        }
        uint16_t mask = 0;
        for (int q = 0; q < 15; q++) {
            y_cleft[(size_t)q * ns + k] = y.cleft_neurotransmitter[q];
            uint32_t bits;
            memcpy(&bits, &y.cleft_neurotransmitter[q], 4);
            if (bits) mask |= (1u << q);
        }
        mask |= 1u << (y.post_receptor_count > 0 ? host_nt_of_receptor(y.post_receptors[0].type) : 0);
        const Segment& post = sg[y.post_segment];
        for (uint32_t j = 0; j < post.modifier_count; j++)
            if (post.modifiers[j].type == 86) { mask |= 1u << 9; break; }   // eCB deposits
        y_mask[k] = mask;
        if (mask & ~3u) sflags[S.host2dev[y.post_segment]] |= SF_NTX;
        // An immediate release whose vesicle recovery reads a presynaptic Ca that an earlier
        // synapse may change: the ordered pass completes it (plan_ordered, per dt).
        if (src[k] >= 0 && !(fl & YF_SKIP)) S.h_cand.push_back(k);
    }
    if (nglia_syn) S.receptors_only = false;   // captures merge into the full post-side walk

    // transmitter diffusion coefficients (types.h:436-443), host powf as the reference
    static const float kMass[15] = {147.0f, 103.0f, 153.0f, 176.0f, 146.0f, 267.0f, 169.0f, 75.0f,
                                    111.0f, 348.0f, 507.0f, 232.0f, 24.0f, 162.0f, 0.0f};
    for (int k = 0; k < 15; k++) {
        float mass = kMass[k];
        if (mass < 1.0f) mass = 1.0f;
        d.nt_diff[k] = 1.33f * powf(23.0f / mass, 1.0f / 3.0f);
    }

    // ---- device allocation + upload ----
    S.tick_dev = S.dalloc<uint64_t>(1);
    d.tick = S.tick_dev;
    S.stim_head = S.dalloc<uint32_t>(npad);
    HIPCHK(hipMemsetAsync(S.stim_head, 0xFF, (size_t)npad * 4, S.stream));
    memset(&d.stim, 0, sizeof(d.stim));
    d.stim.tick_shift = S.stim_tick_shift;
    d.stim.head = S.stim_head;
    d.stim.inj_tick = ~0ull;
    S.id2dev.resize(n);
    for (uint32_t di = 0; di < n; di++) {
        const uint32_t h = S.dev2host[di];
        const uint64_t net = (uint64_t)(std::upper_bound(S.net_seg_off.begin(), S.net_seg_off.end(), h) -
                                        S.net_seg_off.begin() - 1);
        S.id2dev[di] = {(net << 32) | sg[h].id, di};
    }
    std::sort(S.id2dev.begin(), S.id2dev.end());
    d.V = S.dalloc<double>(npad);
    d.in = S.dalloc<double>((size_t)7 * npad);
    d.ex = S.dalloc<double>((size_t)7 * npad);
    d.atp = S.dalloc<float>(npad);
    d.temp = S.dalloc<float>(npad);
    d.txn = S.dalloc<float>(npad);
    d.er = S.dalloc<float>(npad);
    d.nt = S.dalloc<float>((size_t)15 * npad);
    d.nt_in = S.dalloc<float>((size_t)15 * npad);
    d.ibuf = S.dalloc<double>(npad);
    d.spikes = S.dalloc<uint32_t>(npad);
    d.radius = S.dupload(radius);
    d.length = S.dupload(length);
    d.segid = S.dupload(segid);
    d.stype = S.dupload(stype);
    d.area = S.dupload(area);
    d.vol = S.dupload(vol);
    d.cap = S.dupload(cap);
    d.pkc = S.dupload(pkc);
    d.pka = S.dupload(pka);
    d.camp = S.dupload(camp);
    d.pp1 = S.dupload(pp1);
    d.camkii = S.dupload(camkii);
    d.camkii01 = S.dupload(camkii01);
    d.calcn = S.dupload(calcn);
    d.sflags = S.dupload(sflags);
    d.cav21_slot = S.dupload(cav21);
    d.mcount = S.dupload(mcount);
    d.mtype = S.dupload(mtype);
    d.mav = S.dupload(mav);
    d.mdens = S.dupload(mdens);
    d.mstate = S.dalloc<double>((size_t)nslot * 7 * npad);
    {
        const uint32_t* ts0 = S.dupload(tile_seg0);
        const uint32_t* tns = S.dupload(tile_nseg);
        const uint32_t* tc0 = S.dupload(tile_conn0);
        const uint32_t* tnc = S.dupload(tile_nconn);
        const uint32_t* tmi = S.dupload(tile_maxinc);
        for (int k = 0; k < kTileClasses; k++) {
            uint32_t o = S.class_tile0[k];
            S.tiles[k] = TileSet{ts0 + o, tns + o, tc0 + o, tnc + o, tmi + o, S.class_ntiles[k]};
        }
    }
    d.isl_lane0 = S.dupload(isl_lane0);
    // ---- big islands: bounds, connection ends, scratch planes ----
    S.big = BigSet{};
    if (nbig) {
        BigSet& b = S.big;
        b.count = nbig;
        b.seg_base = tile_seg0[S.ntiles];
        b.conn_base = tile_conn0[S.ntiles];
        b.nseg_all = n - b.seg_base;
        b.nconn_all = nc - b.conn_base;
        std::vector<uint32_t> s0(tile_seg0.begin() + S.ntiles, tile_seg0.end()),
            sn(tile_nseg.begin() + S.ntiles, tile_nseg.end()), c0(tile_conn0.begin() + S.ntiles, tile_conn0.end()),
            cnn(tile_nconn.begin() + S.ntiles, tile_nconn.end()), ca(std::max<uint32_t>(b.nconn_all, 1)),
            cb(std::max<uint32_t>(b.nconn_all, 1));
        for (uint32_t i = 0; i < nc; i++) {
            const uint32_t pos = S.conn2dev[i];
            if (pos < b.conn_base) continue;
            ca[pos - b.conn_base] = S.host2dev[cn[i].segment_a];
            cb[pos - b.conn_base] = S.host2dev[cn[i].segment_b];
        }
        b.seg0 = S.dupload(s0);
        b.nseg = S.dupload(sn);
        b.conn0 = S.dupload(c0);
        b.nconn = S.dupload(cnn);
        b.ca = S.dupload(ca);
        b.cb = S.dupload(cb);
        const size_t SN = b.nseg_all, CN = std::max<uint32_t>(b.nconn_all, 1);
        b.phi = S.dalloc<double>(2 * SN);
        b.diag = S.dalloc<double>(SN);
        b.vf = S.dalloc<float>(2 * SN);
        b.base = S.dalloc<float>(SN);
        b.den = S.dalloc<float>(SN);
        b.dacc = S.dalloc<double>(7 * SN);
        b.dsp = S.dalloc<float>(33 * SN);
        b.gphi = S.dalloc<double>(CN);
        b.cac = S.dalloc<float>(CN);
        b.cdist = S.dalloc<float>(CN);
        b.flux = S.dalloc<float>(7 * CN);
        b.cg = S.dalloc<float>(CN);
        b.cterm = S.dalloc<float>(CN);
        b.eflux = S.dalloc<double>(7 * CN);
    }
    if (d.ns) S.cable = false;   // a network with synapses is not a cable: the mode ends with it
    d.cable = S.cable ? 1 : 0;
    S.conn_scale_dev = nullptr;
    d.conn_scale = nullptr;
    d.jac_tol = S.jac_res ? S.jac_tol : 0.0;
    d.jac_max = S.jac_res ? S.jac_max : 50;
    d.conn_ab = S.dupload(conn_ab);
    d.conn_intra = S.dupload(conn_intra);
    d.csr_off = S.dupload(csr_off);
    d.csr_ent = S.dupload(csr_ent);
    d.y_pre = S.dupload(y_pre);
    d.y_prew = S.dupload(y_prew);
    d.watched = S.dupload(watched);
    d.nwatched = (uint32_t)watched.size();
    d.presum = S.dalloc<PreSum>(watched.size());
    d.ca_snap = S.dalloc<double>(std::max<size_t>(watched.size(), 1));
    d.y_id = S.dupload(y_id);
    d.y_flags = S.dupload(y_flags);
    d.y_delay = S.dupload(y_delay);
    d.y_pre_area = S.dupload(y_pa);
    d.y_post_area = S.dupload(y_qa);
    d.y_amount = S.dupload(y_amt);
    d.y_tau = S.dupload(y_tau);
    std::vector<float> y_rrecs;
    release_params(S, y_rmax, y_prel, y_prim0, y_rrecs);
    d.y_prel = S.dupload(y_prel);
    d.y_rrp_max = S.dupload(y_rmax);
    d.y_prim0 = S.dupload(y_prim0);
    d.y_rrec = any_ck ? S.dupload(y_rrecs) : nullptr;
    d.y_cleft_mask = S.dupload(y_mask);
    d.y_src = S.dupload(src);
    d.post_len = S.dupload(post_len);
    d.wave_base = S.dupload(wave_base);
    d.walk = S.dupload(walk);
    d.y_rcount = S.dupload(y_rcount);
    d.y_rtype = S.dupload(y_rtype);
    d.y_rav = S.dupload(y_rav);
    d.y_rdens = S.dupload(y_rdens);
    d.y_rstate = S.dupload(y_rstate);
    d.y_cleft = S.dupload(y_cleft);
    d.y_weight = S.dupload(y_w);
    d.y_fac = S.dupload(y_fac);
    d.y_dep = S.dupload(y_dep);
    d.y_refr = S.dupload(y_refr);
    d.y_rrp = S.dupload(y_rrp);
    d.y_bca = S.dupload(y_bca);
    d.y_sca = S.dupload(y_sca);
    d.y_camkii = S.dupload(y_ck);
    d.y_calcn = S.dupload(y_cn);
    d.y_relp = S.dalloc<float>(ns);
    // ---- gap junctions: post end per slot, and each segment's junctions as presynaptic end ----
    d.y_post = nullptr; d.y_igap = nullptr; d.y_sid = nullptr;
    d.gap_off = nullptr; d.gap_slot = nullptr; d.gap_sid = nullptr;
    if (ngap) {
        std::vector<uint32_t> y_post(ns), y_sid(ns), goff(n + 1, 0), gslot, gsid;
        for (uint32_t k = 0; k < ns; k++) {
            y_sid[k] = S.slot2syn[k];
            y_post[k] = S.host2dev[sy[S.slot2syn[k]].post_segment];
        }
        // the first synapse index at which each segment's Ca can change in step 3: an incoming
        // chemical synapse with a receptor (every receptor call floors the ions at 1e-9,
        // modifier_helpers.h:1117-1121, and NMDA-type ones move Ca)
        std::vector<uint32_t> first_ca(n, UINT32_MAX);
        for (uint32_t i = 0; i < ns; i++) {
            const Synapse& y = sy[i];
            if (y.type == 1 || y.type == 2 || y.post_receptor_count == 0) continue;
            uint32_t& f = first_ca[S.host2dev[y.post_segment]];
            f = std::min(f, i);
        }
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> out(n);   // (synapse, slot)
        for (uint32_t k = 0; k < ns; k++) {
            if (!(y_flags[k] & YF_GAP)) continue;
            const uint32_t i = S.slot2syn[k];
            out[y_pre[k]].push_back({i, k});
            if (first_ca[y_pre[k]] < i || first_ca[y_post[k]] < i) S.h_gaphaz.push_back(k);
        }
        for (uint32_t s = 0; s < n; s++) {
            std::sort(out[s].begin(), out[s].end());
            goff[s + 1] = goff[s] + (uint32_t)out[s].size();
            for (auto& e : out[s]) { gsid.push_back(e.first); gslot.push_back(e.second); }
        }
        d.y_post = S.dupload(y_post);
        d.y_sid = S.dupload(y_sid);
        d.gap_off = S.dupload(goff);
        d.gap_slot = S.dupload(gslot);
        d.gap_sid = S.dupload(gsid);
        d.y_igap = S.dalloc<float>(ns);
    }
    // ---- presynaptic autoreceptors, and the eCB deposit chains in synapse-index order ----
    d.y_ar = nullptr; d.ar_type = nullptr; d.ar_dens = nullptr; d.ar_st0 = nullptr; d.ar_st7 = nullptr;
    if (!ar_type.empty()) {
        if (ar_type.size() >= (1u << 28)) return BION_ERR_ALLOC;
        d.y_ar = S.dupload(y_ar);
        d.ar_type = S.dupload(ar_type);
        d.ar_dens = S.dupload(ar_dens);
        d.ar_st0 = S.dupload(ar_st0);
        d.ar_st7 = S.dupload(ar_st7);
    }
    S.nar = ar_type.size();
    S.h_ar = std::move(y_ar);
    d.y_ecb = nullptr; d.seg_ecb = nullptr;
    {
        std::vector<uint8_t> seg_ecb(n, 0xFF);   // the first eCB synthesis slot
        bool any = false;
        for (uint32_t h = 0; h < n; h++)
            for (uint32_t j = 0; j < sg[h].modifier_count; j++)
                if (sg[h].modifiers[j].type == 86) {
                    seg_ecb[S.host2dev[h]] = (uint8_t)j;
                    any = true;
                    break;
                }
        if (any && ns) {
            std::vector<uint32_t> y_ecb(ns, EC_NONE), cnt(n, 0), last(n, UINT32_MAX);
            for (uint32_t i = 0; i < ns; i++) {
                const Synapse& y = sy[i];
                if (y.type == 1 || y.type == 2 || y.pre_segment >= n || y.post_segment >= n) continue;
                const uint32_t p = S.host2dev[y.post_segment];
                if (seg_ecb[p] == 0xFF) continue;
                const uint32_t k = syn2slot[i];
                y_ecb[k] = cnt[p]++;
                last[p] = k;
            }
            for (uint32_t p = 0; p < n; p++)
                if (last[p] != UINT32_MAX) y_ecb[last[p]] |= EC_LAST;
            d.y_ecb = S.dupload(y_ecb);
            d.seg_ecb = S.dupload(seg_ecb);
            if (!d.y_post) {
                std::vector<uint32_t> y_post(ns);
                for (uint32_t k = 0; k < ns; k++) y_post[k] = S.host2dev[sy[S.slot2syn[k]].post_segment];
                d.y_post = S.dupload(y_post);
            }
        }
    }
    // ---- glial captures: per receiving segment, its synapses in synapse-index order ----
    d.y_cap = nullptr; d.y_capm = nullptr;
    d.glia_off = nullptr; d.glia_slot = nullptr; d.glia_sid = nullptr;
    if (nglia_syn) {
        std::vector<uint32_t> coff(n + 1, 0), cslot(nglia_syn), csid(nglia_syn);
        for (uint32_t k = 0; k < ns; k++)
            if (y_glia[k] != UINT32_MAX) coff[y_glia[k] + 1]++;
        for (uint32_t i = 0; i < n; i++) coff[i + 1] += coff[i];
        std::vector<uint32_t> fill(coff.begin(), coff.end() - 1);
        for (uint32_t i = 0; i < ns; i++) {
            const uint32_t k = syn2slot[i];
            if (y_glia[k] == UINT32_MAX) continue;
            const uint32_t e = fill[y_glia[k]]++;
            cslot[e] = k;
            csid[e] = i;
        }
        if (!d.y_sid) {
            std::vector<uint32_t> y_sid(S.slot2syn.begin(), S.slot2syn.end());
            d.y_sid = S.dupload(y_sid);
        }
        d.glia_off = S.dupload(coff);
        d.glia_slot = S.dupload(cslot);
        d.glia_sid = S.dupload(csid);
        d.y_cap = S.dalloc<float>((size_t)3 * ns);
        d.y_capm = S.dalloc<uint8_t>(ns);
    }
    d.y_isyn = S.dalloc<float>(ns);
    d.y_pp1 = S.dupload(y_pp1);
    d.y_postd = nullptr;
    if (any_pp1) {
        std::vector<uint32_t> yp(ns);
        for (uint32_t k = 0; k < ns; k++) yp[k] = S.host2dev[sy[S.slot2syn[k]].post_segment];
        d.y_postd = S.dupload(yp);
    }
    d.hist_base = S.dupload(hist_base);
    d.ca_hist = S.dalloc<double>(std::max<uint32_t>(nhist, 1));
    d.d_mask = S.dalloc<uint8_t>(ns);
    d.d_arr = S.dalloc<uint64_t>((size_t)4 * ns);
    d.d_amt = S.dalloc<float>((size_t)4 * ns);
    d.d_nt = S.dalloc<uint8_t>((size_t)4 * ns);
    S.lds_bytes = sizeof(DiffLds<256>);
    S.seg_dispatched.assign((size_t)n * 32, 0);
    S.host2dev_dev = S.dupload(S.host2dev);
    S.conn2dev_dev = S.dupload(S.conn2dev);
    // kept for planning the ordered pass per dt
    S.h_src = std::move(src);
    S.h_rmax = std::move(y_rmax);
    S.h_yflags = std::move(y_flags);
    S.h_sflags = std::move(sflags);
    S.h_hist_base = std::move(hist_base);
    if (nglia_syn) S.h_glia = std::move(y_glia);
    else S.h_glia.clear();
    // hazards at any dt: the gap junctions, and the candidates whose release is always immediate
    S.hazard_syn = S.h_gaphaz.size();
    for (uint32_t k : S.h_cand)
        if (sy[S.slot2syn[k]].delay <= 0.0f && S.h_rmax[k] > 0.0f) S.hazard_syn++;
    d.seg_ord = nullptr; d.y_haz = nullptr;
    d.ord_ev = nullptr; d.ord_fl = nullptr; d.ord_hist = nullptr; d.ord_glia = nullptr; d.ord_post = nullptr;
    d.ord_seg = nullptr; d.ord_n = 0; d.ord_nseg = 0;
    return 0;
}

// ------------------------------------------------------------------ ordered pass planning
static uint64_t host_delay_ticks(float delay, float dt) {   // delay_buffer.h:61-64
    if (delay <= 0.0f || dt <= 0.0f) return 0;
    return (uint64_t)((delay / dt) + 0.5f);
}

// The hazards at this dt, the ordered segments they make and the events touching those, in
// synapse-index order (k_syn_ordered). Runs when dt differs from the last plan; with no hazard
// the ordered pass is off (null pointers).
static void plan_ordered(Sim& S, float dt) {
    if (S.ord_dt == dt) return;
    Net& d = S.d;
    const BionData& B = S.host;
    const uint32_t n = d.n, ns = d.ns;
    S.ord_dt = dt;
    d.seg_ord = nullptr; d.y_haz = nullptr;
    d.ord_ev = nullptr; d.ord_fl = nullptr; d.ord_hist = nullptr; d.ord_glia = nullptr; d.ord_post = nullptr;
    d.ord_seg = nullptr; d.ord_n = 0; d.ord_nseg = 0;
    S.ord_events = 0;
    S.ord_segments = 0;
    if (!ns) return;
    std::vector<uint8_t>& haz = S.o_yhaz;
    std::vector<uint8_t>& ser = S.o_segord;
    haz.assign(ns, 0);
    ser.assign(n, 0);
    uint64_t nhaz = 0;
    for (uint32_t k : S.h_cand) {
        const Synapse& y = B.synapses[S.slot2syn[k]];
        if (!(S.h_rmax[k] > 0.0f) || host_delay_ticks(y.delay, dt) != 0) continue;
        haz[k] = HZ_RELEASE;
        ser[S.host2dev[y.post_segment]] = 1;
        if (!S.h_glia.empty() && S.h_glia[k] != UINT32_MAX) ser[S.h_glia[k]] = 1;
        nhaz++;
    }
    for (uint32_t k : S.h_gaphaz) {
        const Synapse& y = B.synapses[S.slot2syn[k]];
        haz[k] = HZ_GAP;
        ser[S.host2dev[y.pre_segment]] = 1;
        ser[S.host2dev[y.post_segment]] = 1;
        nhaz++;
    }
    S.hazard_syn = nhaz;
    if (!nhaz) return;
    std::vector<uint32_t> syn2slot(ns), rank(n, 0);
    for (uint32_t k = 0; k < ns; k++) syn2slot[S.slot2syn[k]] = k;
    S.o_ev.clear(); S.o_fl.clear(); S.o_hist.clear(); S.o_glia.clear(); S.o_post.clear(); S.o_seg.clear();
    for (uint32_t i = 0; i < ns; i++) {
        const uint32_t k = syn2slot[i];
        const Synapse& y = B.synapses[i];
        const uint32_t post = S.host2dev[y.post_segment], pre = S.host2dev[y.pre_segment];
        const uint32_t q = rank[post]++;   // position in the post segment's chain
        const bool gap = S.h_yflags[k] & YF_GAP;
        const uint32_t g = S.h_glia.empty() ? UINT32_MAX : S.h_glia[k];
        uint8_t f = 0;
        if (ser[post]) f |= EV_POST;
        if (gap && ser[pre]) f |= EV_PREGAP;
        if (g != UINT32_MAX && ser[g]) f |= EV_GLIA;
        if (haz[k] == HZ_RELEASE) f |= EV_HAZ;
        if (haz[k] == HZ_GAP) f |= EV_GAPHAZ;
        if (!f) continue;
        S.o_ev.push_back(k);
        S.o_fl.push_back(f);
        S.o_hist.push_back((f & EV_POST) && (S.h_sflags[post] & SF_WATCHED) ? (int32_t)(S.h_hist_base[post] + q) : -1);
        S.o_glia.push_back(g);
        S.o_post.push_back(post);
    }
    for (uint32_t i = 0; i < n; i++)
        if (ser[i]) S.o_seg.push_back(i);
    d.seg_ord = S.dput(S.sb_segord, ser);
    d.y_haz = S.dput(S.sb_yhaz, haz);
    d.ord_ev = S.dput(S.sb_oev, S.o_ev);
    d.ord_fl = S.dput(S.sb_ofl, S.o_fl);
    d.ord_hist = S.dput(S.sb_ohist, S.o_hist);
    d.ord_glia = S.dput(S.sb_oglia, S.o_glia);
    d.ord_post = S.dput(S.sb_opost, S.o_post);
    d.ord_seg = S.dput(S.sb_oseg, S.o_seg);
    d.ord_n = (uint32_t)S.o_ev.size();
    d.ord_nseg = (uint32_t)S.o_seg.size();
    S.ord_events = d.ord_n;
    S.ord_segments = d.ord_nseg;
    HIPCHK(hipStreamSynchronize(S.stream));
}

// ------------------------------------------------------------------ coherence
// Host staging of the transferred state: every array below in one block, 256-byte aligned.
struct XferBufs {
    double *V, *in, *ex, *ms, *yrs, *a0, *a7;
    float *atp, *temp, *txn, *er, *nt, *nti, *yw, *yf, *yd, *yr, *yb, *ys, *yk, *yc, *ycl;
    uint8_t* yrf;
};
// Lays the arrays out from base (nullptr: only sizes them); returns the block size.
static size_t xfer_layout(const Sim& S, char* base, XferBufs& b) {
    const Net& d = S.d;
    const size_t npad = d.npad, ns = d.ns;
    size_t off = 0;
    auto take = [&](auto*& p, size_t count) {
        using T = std::remove_reference_t<decltype(*p)>;
        p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (count * sizeof(T) + 255) & ~(size_t)255;
    };
    take(b.V, npad); take(b.in, 7 * npad); take(b.ex, 7 * npad); take(b.ms, (size_t)d.nslot * 7 * npad);
    take(b.atp, npad); take(b.temp, npad); take(b.txn, npad); take(b.er, npad); take(b.nt, 15 * npad);
    take(b.nti, 15 * npad);
    take(b.yw, ns); take(b.yf, ns); take(b.yd, ns); take(b.yr, ns); take(b.yb, ns); take(b.ys, ns);
    take(b.yk, ns); take(b.yc, ns); take(b.ycl, 15 * ns); take(b.yrf, ns); take(b.yrs, (size_t)d.nrec * 7 * ns);
    take(b.a0, S.nar); take(b.a7, S.nar);
    return off;
}

// Runs f(lo, hi) over [0, n) in up to 16 host threads, one contiguous range each.
template <class F>
static void par_for(size_t n, F f) {
    unsigned T = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    if (n < 32768 || T == 1) {
        f((size_t)0, n);
        return;
    }
    const size_t chunk = (n + T - 1) / T;
    std::vector<std::thread> th;
    for (unsigned k = 1; k < T; k++) {
        const size_t lo = k * chunk, hi = std::min(n, lo + chunk);
        if (lo < hi) th.emplace_back(f, lo, hi);
    }
    f((size_t)0, std::min(n, chunk));
    for (auto& x : th) x.join();
}

// Transfers the dynamic state between the host AoS and the device SoA (both directions use
// the same field list, so what the device computes is exactly what callers can read/write).
// COMPAT coherence transfers every tick through one pinned staging block kept for the loaded
// network; other transfers stage in pageable memory. Packing and unpacking run per record
// range in host threads.
static void xfer(Sim& S, bool to_device) {
    Net& d = S.d;
    BionData& B = S.host;
    const uint32_t n = d.n, npad = d.npad, ns = d.ns, nslot = d.nslot, nrec = d.nrec;
    XferBufs b{};
    const size_t need = xfer_layout(S, nullptr, b);
    std::unique_ptr<char[]> tmp;
    char* base;
    if (S.coherence == BCHIP_COHERENCE_COMPAT) {
        if (S.stage_bytes < need) {
            S.free_stage();
            HIPCHK(hipHostMalloc((void**)&S.stage, need, hipHostMallocDefault));
            S.stage_bytes = need;
            par_for(need, [&](size_t lo, size_t hi) { memset(S.stage + lo, 0, hi - lo); });
        }
        base = S.stage;
    } else {
        tmp.reset(new char[need]);
        base = tmp.get();
        if (to_device) par_for(need, [&](size_t lo, size_t hi) { memset(base + lo, 0, hi - lo); });
    }
    xfer_layout(S, base, b);
    auto cp = [&](void* dev, void* hostbuf, size_t bytes) {
        if (!bytes) return;
        if (to_device) HIPCHK(hipMemcpyAsync(dev, hostbuf, bytes, hipMemcpyHostToDevice, S.stream));
        else HIPCHK(hipMemcpyAsync(hostbuf, dev, bytes, hipMemcpyDeviceToHost, S.stream));
    };
    if (to_device) {
        par_for(n, [&](size_t lo, size_t hi) {
            for (size_t di = lo; di < hi; di++) {
                const Segment& s = B.segments[S.dev2host[di]];
                b.V[di] = s.V_m;
                for (int k = 0; k < 7; k++) {
                    b.in[(size_t)k * npad + di] = s.intra_ions[k];
                    b.ex[(size_t)k * npad + di] = s.extra_ions[k];
                }
                b.atp[di] = s.ATP_level; b.temp[di] = s.temperature; b.txn[di] = s.transcription_signal;
                b.er[di] = s.ER_Ca2;
                for (int k = 0; k < 15; k++) {
                    b.nt[(size_t)k * npad + di] = s.neurotransmitter[k];
                    b.nti[(size_t)k * npad + di] = s.intra_nt[k];
                }
                for (uint32_t j = 0; j < s.modifier_count; j++)
                    for (int q = 0; q < 7; q++) b.ms[((size_t)j * 7 + q) * npad + di] = s.modifiers[j].state[q];
            }
        });
        par_for(ns, [&](size_t lo, size_t hi) {
            for (size_t k = lo; k < hi; k++) {
                const Synapse& y = B.synapses[S.slot2syn[k]];
                b.yw[k] = y.weight; b.yf[k] = y.facilitation; b.yd[k] = y.depression; b.yr[k] = y.RRP;
                b.yb[k] = y.bouton_Ca; b.ys[k] = y.spine_Ca; b.yk[k] = y.spine_camkii;
                b.yc[k] = y.spine_calcineurin; b.yrf[k] = y.refractory;
                for (int q = 0; q < 15; q++) b.ycl[(size_t)q * ns + k] = y.cleft_neurotransmitter[q];
                for (uint32_t j = 0; j < y.post_receptor_count; j++)
                    for (int q = 0; q < 7; q++) b.yrs[((size_t)j * 7 + q) * ns + k] = y.post_receptors[j].state[q];
                for (uint32_t j = 0, a = S.nar ? S.h_ar[k] : 0; j < (a & 15u); j++) {
                    b.a0[(a >> 4) + j] = y.pre_autoreceptors[j].state[0];
                    b.a7[(a >> 4) + j] = y.pre_autoreceptors[j].state[7];
                }
            }
        });
    }
    cp(d.V, b.V, (size_t)npad * 8);
    cp(d.in, b.in, (size_t)7 * npad * 8);
    cp(d.ex, b.ex, (size_t)7 * npad * 8);
    cp(d.mstate, b.ms, (size_t)nslot * 7 * npad * 8);
    cp(d.atp, b.atp, (size_t)npad * 4);
    cp(d.temp, b.temp, (size_t)npad * 4);
    cp(d.txn, b.txn, (size_t)npad * 4);
    cp(d.er, b.er, (size_t)npad * 4);
    cp(d.nt, b.nt, (size_t)15 * npad * 4);
    cp(d.nt_in, b.nti, (size_t)15 * npad * 4);
    if (ns) {
        cp(d.y_weight, b.yw, (size_t)ns * 4);
        cp(d.y_fac, b.yf, (size_t)ns * 4);
        cp(d.y_dep, b.yd, (size_t)ns * 4);
        cp(d.y_rrp, b.yr, (size_t)ns * 4);
        cp(d.y_bca, b.yb, (size_t)ns * 4);
        cp(d.y_sca, b.ys, (size_t)ns * 4);
        cp(d.y_camkii, b.yk, (size_t)ns * 4);
        cp(d.y_calcn, b.yc, (size_t)ns * 4);
        cp(d.y_refr, b.yrf, ns);
        cp(d.y_cleft, b.ycl, (size_t)15 * ns * 4);
        cp(d.y_rstate, b.yrs, (size_t)nrec * 7 * ns * 8);
        cp(d.ar_st0, b.a0, S.nar * 8);
        cp(d.ar_st7, b.a7, S.nar * 8);
    }
    if (to_device) {   // intra_nt is live once any value is not +0.0 (read while the copy runs)
        std::atomic<bool> live{false};
        par_for((size_t)15 * npad, [&](size_t lo, size_t hi) {
            for (size_t q = lo; q < hi && !live.load(std::memory_order_relaxed); q++) {
                uint32_t bits;
                memcpy(&bits, &b.nti[q], 4);
                if (bits) live = true;
            }
        });
        d.nt_in_live = live ? 1 : 0;
    }
    HIPCHK(hipStreamSynchronize(S.stream));
    if (to_device) return;
    const bool ran = S.ticks_run > 0;
    par_for(n, [&](size_t lo, size_t hi) {
        for (size_t di = lo; di < hi; di++) {
            Segment& s = B.segments[S.dev2host[di]];
            s.V_m = b.V[di];
            for (int k = 0; k < 7; k++) {
                s.intra_ions[k] = b.in[(size_t)k * npad + di];
                s.extra_ions[k] = b.ex[(size_t)k * npad + di];
            }
            s.ATP_level = b.atp[di]; s.temperature = b.temp[di]; s.transcription_signal = b.txn[di];
            s.ER_Ca2 = b.er[di];
            for (int k = 0; k < 15; k++) {
                s.neurotransmitter[k] = b.nt[(size_t)k * npad + di];
                s.intra_nt[k] = b.nti[(size_t)k * npad + di];
            }
            if (ran) s.I_buffer = 0.0;
            bool my = false;
            for (uint32_t j = 0; j < s.modifier_count; j++) my |= s.modifiers[j].type == 4;
            for (uint32_t j = 0; j < s.modifier_count; j++) {
                Modifier& m = s.modifiers[j];
                for (int q = 0; q < 7; q++) m.state[q] = b.ms[((size_t)j * 7 + q) * npad + di];
                uint8_t t = m.type;
                bool blocked = my && t != 4 && myelin_blocks(t);
                if (ran && !blocked) m.state[7] = seg_area(s);
            }
        }
    });
    par_for(ns, [&](size_t lo, size_t hi) {
        for (size_t k = lo; k < hi; k++) {
            Synapse& y = B.synapses[S.slot2syn[k]];
            if (ran && S.protein_map) protein_mapped(y, B, y.RRP_max, y.release_probability, y.RRP_recovery);
            y.weight = b.yw[k]; y.facilitation = b.yf[k]; y.depression = b.yd[k]; y.RRP = b.yr[k];
            y.bouton_Ca = b.yb[k]; y.spine_Ca = b.ys[k]; y.spine_camkii = b.yk[k];
            y.spine_calcineurin = b.yc[k]; y.refractory = b.yrf[k];
            for (int q = 0; q < 15; q++) y.cleft_neurotransmitter[q] = b.ycl[(size_t)q * ns + k];
            const Segment& post = B.segments[y.post_segment];
            bool my = false;
            for (uint32_t j = 0; j < post.modifier_count; j++) my |= post.modifiers[j].type == 4;
            for (uint32_t j = 0; j < y.post_receptor_count; j++) {
                Modifier& m = y.post_receptors[j];
                for (int q = 0; q < 7; q++) m.state[q] = b.yrs[((size_t)j * 7 + q) * ns + k];
                uint8_t t = m.type;
                bool blocked = my && ionotropic_type(t);
                if (ran && !(y.type == 1 || y.type == 2) && !blocked) m.state[7] = y.post_area;
            }
            for (uint32_t j = 0, a = S.nar ? S.h_ar[k] : 0; j < (a & 15u); j++) {
                y.pre_autoreceptors[j].state[0] = b.a0[(a >> 4) + j];
                y.pre_autoreceptors[j].state[7] = b.a7[(a >> 4) + j];
            }
        }
    });
}

static void sync_host(Sim& S) {
    if (S.host_stale) { xfer(S, false); S.host_stale = false; }
}

// ------------------------------------------------------------------ stimuli
// One stimulus entry bound to one device segment: key = kind << 28 | index into that kind's
// payload array. Sorting by (segment, key) gives each segment's list in the reference's order.
struct StimTrip {
    uint32_t dev, key;
    bool operator<(const StimTrip& o) const { return dev != o.dev ? dev < o.dev : key < o.key; }
};

// Appends network net's entries of one kind, matched by Segment.id through the sorted
// (network << 32 | id, device) map; index0 is where they start in the combined payload array.
template <class T>
static void stim_match(const std::vector<T>& v, uint32_t kind, uint64_t net, size_t index0,
                       const std::vector<std::pair<uint64_t, uint32_t>>& ids, std::vector<StimTrip>& out) {
    for (size_t k = 0; k < v.size(); k++) {
        const uint64_t key = (net << 32) | v[k].segment_id;
        auto lo = std::lower_bound(ids.begin(), ids.end(), std::make_pair(key, 0u));
        for (; lo != ids.end() && lo->first == key; ++lo)
            out.push_back({lo->second, (kind << 28) | (uint32_t)(index0 + k)});
    }
}

// Builds the per-segment lists from S.st_* and trips, uploads them and repoints the heads
// (segments of the previous lists are cleared first). Injections apply on tick inj_tick.
static void stim_commit(Sim& S, std::vector<StimTrip>& trips, uint64_t inj_tick) {
    Net& d = S.d;
    std::sort(trips.begin(), trips.end());
    const int nxt = S.st_cur_set ^ 1;
    std::vector<uint32_t>& segs = S.st_segs[nxt];
    segs.clear();
    S.st_heads.clear();
    S.st_ent.clear();
    for (size_t a = 0; a < trips.size();) {
        size_t b = a;
        while (b < trips.size() && trips[b].dev == trips[a].dev) b++;
        segs.push_back(trips[a].dev);
        S.st_heads.push_back((uint32_t)S.st_ent.size());
        S.st_ent.push_back((uint32_t)(b - a));
        for (size_t q = a; q < b; q++) S.st_ent.push_back(trips[q].key);
        a = b;
    }
    d.stim.cur = S.dput(S.sb_cur, S.st_cur);
    d.stim.ramp = S.dput(S.sb_ramp, S.st_ramp);
    d.stim.vc = S.dput(S.sb_vc, S.st_vc);
    d.stim.ion = S.dput(S.sb_ion, S.st_ion);
    d.stim.lig = S.dput(S.sb_lig, S.st_lig);
    d.stim.ent = S.dput(S.sb_ent, S.st_ent);
    const uint32_t* nsegs = S.dput(S.sb_segs[nxt], segs);
    const uint32_t* nheads = S.dput(S.sb_heads, S.st_heads);
    const std::vector<uint32_t>& prev = S.st_segs[S.st_cur_set];
    if (!prev.empty())
        k_stim_map<<<((uint32_t)prev.size() + 255) / 256, 256, 0, S.stream>>>(
            static_cast<const uint32_t*>(S.sb_segs[S.st_cur_set].p), nullptr, (uint32_t)prev.size(), S.stim_head);
    if (!segs.empty())
        k_stim_map<<<((uint32_t)segs.size() + 255) / 256, 256, 0, S.stream>>>(nsegs, nheads, (uint32_t)segs.size(),
                                                                               S.stim_head);
    S.st_cur_set = nxt;
    d.stim.head = S.stim_head;
    d.stim.any = segs.empty() ? 0 : 1;
    d.stim.inj_tick = inj_tick;
    for (const LigandInjection& q : S.st_lig)   // sources intra_nt from that tick on
        if (q.intracellular) d.nt_in_live = 1;
}

// All networks' stimuli -> one set of lists (Segment.id matching of the CPU tick, per network).
template <class T>
static void stim_cat(std::vector<T>& dst, const std::vector<T>& src, std::vector<size_t>& at) {
    at.push_back(dst.size());
    dst.insert(dst.end(), src.begin(), src.end());
}
static void stim_rebuild(Sim& S) {
    S.st_cur.clear(); S.st_ramp.clear(); S.st_vc.clear(); S.st_ion.clear(); S.st_lig.clear();
    std::vector<StimTrip> trips;
    std::vector<size_t> a, b, c, e, f;
    for (uint64_t k = 0; k < S.net_stim.size(); k++) {
        const Sim::NetStim& ns = S.net_stim[k];
        stim_cat(S.st_cur, ns.cur, a);
        stim_cat(S.st_ramp, ns.ramp, b);
        stim_cat(S.st_ion, ns.ion, c);
        stim_cat(S.st_lig, ns.lig, e);
        stim_cat(S.st_vc, ns.vc, f);
        stim_match(ns.cur, SK_CUR, k, a.back(), S.id2dev, trips);
        stim_match(ns.ramp, SK_RAMP, k, b.back(), S.id2dev, trips);
        stim_match(ns.ion, SK_ION, k, c.back(), S.id2dev, trips);
        stim_match(ns.lig, SK_LIG, k, e.back(), S.id2dev, trips);
        stim_match(ns.vc, SK_VCLAMP, k, f.back(), S.id2dev, trips);
    }
    const bool inj = !S.st_ion.empty() || !S.st_lig.empty();
    stim_commit(S, trips, inj ? S.tick : ~0ull);
    S.inj_committed = inj;
}
// After a tick has run: committed injections were one-shot (simulation_host.cpp:304-306); the
// device lists keep them but their tick has passed, and they leave the next rebuild.
static void stim_after_tick(Sim& S) {
    if (!S.inj_committed) return;
    for (Sim::NetStim& ns : S.net_stim) { ns.ion.clear(); ns.lig.clear(); }
    S.inj_committed = false;
}

// ------------------------------------------------------------------ tick
static void launch_modifiers(Sim& S, float dt, uint32_t first, uint32_t last) {
    if (first >= last) return;
    const Net& d = S.d;
    const uint32_t B = S.sched.modifiers_block;
    const uint32_t gm = (last - first + B - 1) / B;
    if (S.ext_mods) {
        if (S.nt_read) k_modifiers<true, true><<<gm, B, 0, S.stream>>>(d, dt, first, last);
        else k_modifiers<false, true><<<gm, B, 0, S.stream>>>(d, dt, first, last);
    } else {
        if (S.nt_read) k_modifiers<true, false><<<gm, B, 0, S.stream>>>(d, dt, first, last);
        else k_modifiers<false, false><<<gm, B, 0, S.stream>>>(d, dt, first, last);
    }
}

// Step 2.5 of the 64-segment tiles [t0, t1) on stream st.
template <bool kRes, bool kRec>
static void launch_diffusion64(Sim& S, float dt, uint32_t t0, uint32_t t1, hipStream_t st) {
    if (t0 >= t1) return;
    TileSet ts = S.tiles[0];
    ts.seg0 += t0; ts.nseg += t0; ts.conn0 += t0; ts.nconn += t0; ts.maxinc += t0;
    ts.count = t1 - t0;
    const uint32_t W = S.sched.diffusion_tiles;
    const size_t lds = W * (sizeof(DiffLds<64>) + (kRes ? 2 * 64 * 8 : 0));
    k_diffusion<64, kRes, kRec><<<(ts.count + W - 1) / W, 64 * W, lds, st>>>(S.d, ts, dt);
}

// Step 2.5 of the other size classes and the big islands.
template <bool kRes, bool kRec>
static void launch_diffusion_rest(Sim& S, float dt) {
    const Net& d = S.d;
    auto lds = [](size_t tile, uint32_t ts) { return tile + (kRes ? 2 * ts * 8 : 0); };
    if (S.class_ntiles[1])
        k_diffusion<128, kRes, kRec><<<S.class_ntiles[1], 128, lds(sizeof(DiffLds<128>), 128), S.stream>>>(
            d, S.tiles[1], dt);
    if (S.class_ntiles[2])
        k_diffusion<256, kRes, kRec><<<S.class_ntiles[2], 256, lds(sizeof(DiffLds<256>), 256), S.stream>>>(
            d, S.tiles[2], dt);
    if (!kRec && S.big.count) k_diffusion_big<kRes><<<S.big.count, kBigThreads, 0, S.stream>>>(d, S.big, dt);
}

// Steps 1 and 2.5. With sched.step1_chunks = C > 1 the 64-segment tiles are cut into C chunks of
// consecutive tiles: step 1 of chunk c runs on `stream`, and step 2.5 of chunk c on `stream2`
// once step 1 of chunk c is done, so the VALU-bound sweeps of chunk c overlap the memory-bound
// modifier streams of chunk c + 1. Step 1 of a segment reads and writes only that segment, and
// step 2.5 of a tile only the tile's segments, so results do not depend on C. `stream2` joins
// `stream` again before step 3.
template <bool kRes, bool kRec>
static void launch_step12(Sim& S, float dt, bool tm) {
    const Net& d = S.d;
    const uint32_t nt0 = S.class_ntiles[0];
    const uint32_t C = d.has_conn ? std::min(std::min(S.sched.step1_chunks, Sim::kMaxChunks), nt0) : 0;
    if (C <= 1) {
        launch_modifiers(S, dt, 0, d.n);
        if (tm) HIPCHK(hipEventRecord(S.ev[1], S.stream));
        if (d.has_conn) {
            launch_diffusion64<kRes, kRec>(S, dt, 0, nt0, S.stream);
            launch_diffusion_rest<kRes, kRec>(S, dt);
        }
        return;
    }
    for (uint32_t c = 0; c < C; c++) {
        const uint32_t t0 = (uint32_t)((uint64_t)nt0 * c / C), t1 = (uint32_t)((uint64_t)nt0 * (c + 1) / C);
        launch_modifiers(S, dt, S.h_tile64_seg0[t0], S.h_tile64_seg0[t1]);
        HIPCHK(hipEventRecord(S.fork_ev[c], S.stream));
        HIPCHK(hipStreamWaitEvent(S.stream2, S.fork_ev[c], 0));
        launch_diffusion64<kRes, kRec>(S, dt, t0, t1, S.stream2);
    }
    launch_modifiers(S, dt, S.h_tile64_seg0[nt0], d.n);
    if (tm) HIPCHK(hipEventRecord(S.ev[1], S.stream));
    launch_diffusion_rest<kRes, kRec>(S, dt);
    HIPCHK(hipEventRecord(S.fork_ev[C], S.stream2));
    HIPCHK(hipStreamWaitEvent(S.stream, S.fork_ev[C], 0));
}

// The kernel sequence of one tick on S.stream (and S.stream2, joined back; no host
// synchronisation or copies, so it can be captured). With `tm`, events on S.stream bracket the
// phases (with step 1 chunked, the first phase ends when step 1 is done, the overlapped step 2.5
// of its chunks included in either phase).
static void launch_tick(Sim& S, float dt, bool tm) {
    const Net& d = S.d;
    const BchipSchedule& sc = S.sched;
    const uint32_t gy = (d.ns + 255) / 256;
    if (tm) HIPCHK(hipEventRecord(S.ev[0], S.stream));
    if (S.jac_res) launch_step12<true, false>(S, dt, tm);
    else if (d.xrec) launch_step12<false, true>(S, dt, tm);
    else launch_step12<false, false>(S, dt, tm);
    if (tm) HIPCHK(hipEventRecord(S.ev[2], S.stream));
    if (d.ns) {
        k_pre_summary<<<(d.nwatched + 255) / 256, 256, 0, S.stream>>>(d);
        k_syn_pre<<<std::min<uint32_t>(gy, sc.pre_grid), 256, 0, S.stream>>>(d, dt);
    }
    if (tm) HIPCHK(hipEventRecord(S.ev[3], S.stream));
    const uint32_t gs = (d.n + 255) / 256;   // the post pass's windows are 256 segments
    if (S.receptors_only && d.nrec <= 2) k_syn_post_rec<true><<<gs, 256, 0, S.stream>>>(d, dt);
    else if (S.receptors_only) k_syn_post_rec<false><<<gs, 256, 0, S.stream>>>(d, dt);
    else k_syn_post<<<gs, 256, 0, S.stream>>>(d, dt);
    if (d.ord_n) k_syn_ordered<<<1, 64, 0, S.stream>>>(d, dt);
    if (tm) HIPCHK(hipEventRecord(S.ev[4], S.stream));
    if (d.ns) k_syn_plast<<<(d.ns + sc.plast_block - 1) / sc.plast_block, sc.plast_block, 0, S.stream>>>(d, dt);
    if (tm) HIPCHK(hipEventRecord(S.ev[5], S.stream));
    if (d.ns) k_syn_rrp<<<(d.ns + sc.rrp_block - 1) / sc.rrp_block, sc.rrp_block, 0, S.stream>>>(d, dt);
    k_tick_advance<<<1, 64, 0, S.stream>>>(S.tick_dev);
    if (tm) HIPCHK(hipEventRecord(S.ev[6], S.stream));
}

static void before_ticks(Sim& S) {
    if (S.coherence == BCHIP_COHERENCE_COMPAT && !S.host_stale) xfer(S, true);
    if (S.coherence == BCHIP_COHERENCE_COMPAT && S.host_stale) { xfer(S, false); S.host_stale = false; xfer(S, true); }
    if (S.stim_dirty) {   // injections of this upload apply on the tick about to run
        stim_rebuild(S);
        S.stim_dirty = false;
    }
}

static void after_ticks(Sim& S, uint32_t count) {
    S.tick += count;
    S.ticks_run += count;
    stim_after_tick(S);
    S.host_stale = true;
    if (S.coherence == BCHIP_COHERENCE_COMPAT) sync_host(S);
}

static void run_tick(Sim& S, float dt) {
    before_ticks(S);
    plan_ordered(S, dt);
    const bool tm = S.timing != 0;
    launch_tick(S, dt, tm);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(S.stream));
    if (tm)
        for (int k = 0; k < 6; k++) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, S.ev[k], S.ev[k + 1]));
            S.last_ms[k] = ms;
        }
    after_ticks(S, 1);
}

// count ticks as replays of a captured sequence of up to kGraphTicks ticks (recaptured when the
// launch inputs change). Only in FAST coherence without phase timing; otherwise tick by tick.
constexpr uint32_t kGraphTicks = 16;
static void run_ticks(Sim& S, float dt, uint32_t count) {
    if (!count) return;
    if (!S.use_graphs || S.timing || S.coherence != BCHIP_COHERENCE_FAST) {
        for (uint32_t k = 0; k < count; k++) run_tick(S, dt);
        return;
    }
    before_ticks(S);
    plan_ordered(S, dt);
    const uint32_t g = std::min(count, kGraphTicks);
    const bool same = S.graph && S.graph_ticks == g && S.graph_dt == dt &&
                      memcmp(&S.graph_net, &S.d, sizeof(Net)) == 0 &&
                      memcmp(&S.graph_sched, &S.sched, sizeof(BchipSchedule)) == 0;
    if (!same) {
        S.drop_graph();
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
        for (uint32_t k = 0; k < g; k++) launch_tick(S, dt, false);
        HIPCHK(hipStreamEndCapture(S.stream, &gr));
        hipError_t e = hipGraphInstantiate(&S.graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIPCHK(e);
        S.graph_ticks = g;
        S.graph_net = S.d;
        S.graph_dt = dt;
        S.graph_sched = S.sched;
    }
    uint32_t done = 0;
    for (; done + g <= count; done += g) HIPCHK(hipGraphLaunch(S.graph, S.stream));
    for (; done < count; done++) launch_tick(S, dt, false);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(S.stream));
    after_ticks(S, count);
}

// A plain load is one network.
static void plain_networks(Sim& S, const BionData& b) {
    S.net_seg_off = {0, b.num_segments};
    S.net_conn_off = {0, b.num_connections};
    S.net_syn_off = {0, b.num_synapses};
    S.net_stim.resize(1);   // network 0 keeps its stimuli across reloads
}

static int load_into(Sim& S, BionData* src_owned) {
    HIPCHK(hipSetDevice(S.device));
    HIPCHK(hipStreamSynchronize(S.stream));
    S.drop_graph();
    S.free_device();
    S.free_stage();
    free_bion(&S.host);
    S.host = *src_owned;
    memset(src_owned, 0, sizeof(BionData));
    S.tick = 0;
    S.ticks_run = 0;
    S.host_stale = false;
    int rc = build_device(S);
    if (rc != 0) { S.free_device(); free_bion(&S.host); S.loaded = false; return rc; }
    if ((S.unsupported_mod || S.unsupported_syn) && !S.allow_unsupported) {
        fprintf(stderr,
                "biocore_hip: load refused: %llu modifier/receptor instance(s) outside the device set and "
                "%llu unsupported synapse(s) (bchip_allow_unsupported loads them as zero current / "
                "skipped)\n",
                (unsigned long long)S.unsupported_mod, (unsigned long long)S.unsupported_syn);
        HIPCHK(hipStreamSynchronize(S.stream));
        S.free_device();
        free_bion(&S.host);
        S.loaded = false;
        return BCHIP_ERR_UNSUPPORTED;
    }
    xfer(S, true);
    S.st_segs[0].clear();   // the lists named device segments of the previous network
    S.st_segs[1].clear();
    if (S.stim_present) S.stim_dirty = true;
    S.loaded = true;
    return 0;
}

__global__ void k_fill_zero(uint64_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}
__global__ void k_map_u32(uint32_t* p, uint32_t n, const uint32_t* table) {   // p[i] = table[p[i]]
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = table[p[i]];
}
// grad[h][j] (h host segment, j < 32) = gd[j][host2dev[h]] for j < nslot, else 0
__global__ void k_grad_density_out(double* out, const double* gd, uint32_t n, uint32_t P, uint32_t nslot,
                                   const uint32_t* host2dev) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * (uint32_t)BC_MAX_MODIFIERS) return;
    const uint32_t h = e / BC_MAX_MODIFIERS, j = e % BC_MAX_MODIFIERS;
    out[e] = j < nslot ? gd[(size_t)j * P + host2dev[h]] : 0.0;
}
__global__ void k_gather_f64(double* out, const double* in, uint32_t n, const uint32_t* idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}

// Scratch of bchip_adjoint_noalloc: 64 MiB, one copy per GPU the library runs on, held until
// the process ends. In host code `scratch` names a host-side stand-in, not device memory: the
// kernels and copies take the device address hipGetSymbolAddress gives for the current device,
// looked up once per device by scratch_setup (at simulation creation) and kept in g_scratch_ptr.
// Every call on one GPU shares its copy; it keeps the previous call's contents, and calls that
// overlap on one GPU race.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;
static int scratch_setup(int device) {   // `device` is current; 0 or -1
    if (device < 0 || device >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
    g_scratch_ptr[device] = ptr;
    return 0;
}
static void* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Buffers of one adjoint call, every one zeroed when handed out: one hipMalloc each, freed when
// the call ends (bchip_adjoint), or consecutive 256-byte-aligned pieces of the file-scope
// scratch array (bchip_adjoint_noalloc), cleared by a fill kernel.
static size_t adj_round(size_t bytes) { return (std::max<size_t>(bytes, 64) + 255) & ~(size_t)255; }
struct AdjHeap {
    Sim& S;
    std::vector<void*> tmp;
    explicit AdjHeap(Sim& s) : S(s) {}
    bool fits(size_t) const { return true; }
    void* get(size_t bytes) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, adj_round(bytes)));
        tmp.push_back(p);
        HIPCHK(hipMemsetAsync(p, 0, adj_round(bytes), S.stream));
        return p;
    }
    // the optional recorded iterates: within half the free memory and the caller's cap
    bool room(size_t bytes) const {
        size_t mfree = 0, mtotal = 0;
        HIPCHK(hipMemGetInfo(&mfree, &mtotal));
        return bytes <= mfree / 2 && bytes <= S.adj_record_cap;
    }
    ~AdjHeap() {
        for (void* p : tmp) (void)hipFree(p);
    }
};
struct AdjScratch {
    Sim& S;
    char* base;
    size_t cap, used = 0;
    AdjScratch(Sim& s, void* b, size_t c) : S(s), base(static_cast<char*>(b)), cap(c) {}
    bool fits(size_t need) const { return need <= cap; }
    void* get(size_t bytes) {
        const size_t b = adj_round(bytes);
        if (used + b > cap) throw std::runtime_error("scratch");   // excluded by fits() before the run
        char* p = base + used;
        used += b;
        k_fill_zero<<<(uint32_t)std::min<size_t>((b / 8 + 255) / 256, 4096), 256, 0, S.stream>>>(
            reinterpret_cast<uint64_t*>(p), b / 8);
        return p;
    }
    bool room(size_t bytes) const { return adj_round(bytes) <= cap - used && bytes <= S.adj_record_cap; }
};

// Reverse pass over `ticks` cable-mode ticks from the current state (see bchip_adjoint); Mem
// supplies the device buffers (AdjHeap or AdjScratch).
template <class Mem>
static int adjoint_run(Sim& S, Mem& mem, float dt, uint32_t ticks, uint32_t checkpoint_every, const uint32_t* soma,
                       uint32_t ncells, const float* target, double* loss, double* grad_density, double* grad_scale,
                       float* trace) {
    if (!soma || !target || ticks == 0 || ncells == 0) return -1;
    if (!S.loaded || !S.cable || S.d.ns || S.jac_res) return -1;
    Net& d = S.d;
    const uint32_t n = d.n, P = d.npad, nslot = d.nslot;
    if (nslot > 8 || S.big.count) return -2;   // the reverse kernels run on tiles only
    for (uint32_t h = 0; h < n; h++) {
        const Segment& sg = S.host.segments[h];
        for (uint32_t j = 0; j < sg.modifier_count; j++)
            if (!adj::adjoint_supported(sg.modifiers[j].type)) return -2;
    }
    for (uint32_t c = 0; c < ncells; c++)
        if (soma[c] >= n) return -1;
    const bool backward = grad_density || grad_scale;
    const uint32_t K = checkpoint_every ? checkpoint_every : ticks;
    const uint32_t nck = (ticks + K - 1) / K;
    const uint32_t nconn = (uint32_t)S.host.num_connections;
    const size_t vB = (size_t)P * 8, mB = (size_t)nslot * 7 * P * 8, gB = (size_t)nslot * 2 * P * 8;
    const size_t spB = (size_t)P * 4, ntB = (size_t)15 * P * 4;
    const size_t rstride = (size_t)2 * P + (size_t)nslot * 2 * P;
    const size_t xr_tick = (size_t)5 * 51 * P;
    {   // every buffer the run takes apart from the optional recorded iterates, before any effect
        size_t need = adj_round((size_t)(nck + 1) * (vB + mB + 8)) + adj_round((size_t)ncells * 4) +
                      adj_round((size_t)ticks * ncells * 4) + adj_round((size_t)ncells * 8) +
                      (trace ? adj_round((size_t)ticks * ncells * 4) : 0) + adj_round(spB + 2 * ntB + vB);
        if (backward)
            need += adj_round((size_t)(K + 1) * rstride * 8) + 2 * adj_round(vB) + adj_round(gB) +
                    adj_round((size_t)nslot * P * 8) + adj_round((size_t)std::max<uint32_t>(nconn, 1) * 8) +
                    (grad_density ? adj_round((size_t)n * BC_MAX_MODIFIERS * 8) : 0) +
                    (grad_scale ? adj_round((size_t)std::max<uint32_t>(nconn, 1) * 8) : 0);
        if (!mem.fits(need)) return BCHIP_ERR_SCRATCH;
    }
    int rc = 0;
    try {
        (void)hipSetDevice(S.device);
        before_ticks(S);
        char* ck = static_cast<char*>(mem.get((size_t)(nck + 1) * (vB + mB + 8)));
        auto save = [&](uint32_t k) {
            char* b = ck + (size_t)k * (vB + mB + 8);
            HIPCHK(hipMemcpyAsync(b, d.V, vB, hipMemcpyDeviceToDevice, S.stream));
            HIPCHK(hipMemcpyAsync(b + vB, d.mstate, mB, hipMemcpyDeviceToDevice, S.stream));
            HIPCHK(hipMemcpyAsync(b + vB + mB, S.tick_dev, 8, hipMemcpyDeviceToDevice, S.stream));
        };
        auto restore = [&](uint32_t k) {
            const char* b = ck + (size_t)k * (vB + mB + 8);
            HIPCHK(hipMemcpyAsync(d.V, b, vB, hipMemcpyDeviceToDevice, S.stream));
            HIPCHK(hipMemcpyAsync(d.mstate, b + vB, mB, hipMemcpyDeviceToDevice, S.stream));
            HIPCHK(hipMemcpyAsync(S.tick_dev, b + vB + mB, 8, hipMemcpyDeviceToDevice, S.stream));
        };
        // the cells' somas as device segment indices
        uint32_t* dsoma = static_cast<uint32_t*>(mem.get((size_t)ncells * 4));
        HIPCHK(hipMemcpyAsync(dsoma, soma, (size_t)ncells * 4, hipMemcpyHostToDevice, S.stream));
        const uint32_t gc = (ncells + 255) / 256;
        k_map_u32<<<gc, 256, 0, S.stream>>>(dsoma, ncells, S.host2dev_dev);
        float* dtarget = static_cast<float*>(mem.get((size_t)ticks * ncells * 4));
        double* acc = static_cast<double*>(mem.get((size_t)ncells * 8));
        HIPCHK(hipMemcpyAsync(dtarget, target, (size_t)ticks * ncells * 4, hipMemcpyHostToDevice, S.stream));
        // forward: checkpoints every K ticks, loss after every tick
        float* dtrace = trace ? static_cast<float*>(mem.get((size_t)ticks * ncells * 4)) : nullptr;
        for (uint32_t t = 0; t < ticks; t++) {
            if (t % K == 0) save(t / K);
            launch_tick(S, dt, false);
            k_adj_loss<<<gc, 256, 0, S.stream>>>(d.V, dsoma, ncells, dtarget + (size_t)t * ncells, acc,
                                                 dtrace ? dtrace + (size_t)t * ncells : nullptr);
        }
        if (dtrace)
            HIPCHK(hipMemcpyAsync(trace, dtrace, (size_t)ticks * ncells * 4, hipMemcpyDeviceToHost, S.stream));
        save(nck);   // the state after the last tick, restored at the end
        // recomputing a span also counts its spikes again, rewrites I_buffer and re-applies a
        // ligand injection of its first tick: keep those fields as the forward left them
        char* fin = static_cast<char*>(mem.get(spB + 2 * ntB + vB));
        auto keep_final = [&](bool out) {
            void* dev[4] = {d.spikes, d.nt, d.nt_in, d.ibuf};
            const size_t sz[4] = {spB, ntB, ntB, vB};
            size_t o = 0;
            for (int q = 0; q < 4; q++) {
                HIPCHK(hipMemcpyAsync(out ? dev[q] : fin + o, out ? fin + o : dev[q], sz[q],
                                      hipMemcpyDeviceToDevice, S.stream));
                o += sz[q];
            }
        };
        keep_final(false);
        HIPCHK(hipGetLastError());
        double L = 0.0;   // the per-cell sums in cell order, read in pieces
        for (uint32_t c0 = 0; c0 < ncells; c0 += 64) {
            double part[64];
            const uint32_t m = std::min<uint32_t>(64, ncells - c0);
            HIPCHK(hipMemcpyAsync(part, acc + c0, (size_t)m * 8, hipMemcpyDeviceToHost, S.stream));
            HIPCHK(hipStreamSynchronize(S.stream));
            for (uint32_t k = 0; k < m; k++) L += part[k];
        }
        L /= (double)ticks * ncells;
        if (loss) *loss = L;
        if (backward) {
            // per tick of a span: pre-tick V, gate states, and the tick's I_buffer
            double* rec = static_cast<double*>(mem.get((size_t)(K + 1) * rstride * 8));
            AdjNet a{};
            a.lamV = static_cast<double*>(mem.get(vB));
            a.lamG = static_cast<double*>(mem.get(gB));
            a.lamIb = static_cast<double*>(mem.get(vB));
            a.gdens = static_cast<double*>(mem.get((size_t)nslot * P * 8));
            a.gscale = static_cast<double*>(mem.get((size_t)std::max<uint32_t>(nconn, 1) * 8));
            double* gd_out = grad_density ? static_cast<double*>(mem.get((size_t)n * BC_MAX_MODIFIERS * 8)) : nullptr;
            double* gs_out = grad_scale ? static_cast<double*>(mem.get((size_t)std::max<uint32_t>(nconn, 1) * 8)) : nullptr;
            // The forward records each tick's Jacobi iterates (5 x 51 floats per segment) when
            // a span's worth fits (AdjHeap: half the free memory; AdjScratch: the rest of the
            // scratch); otherwise the reverse kernel runs the substeps again from the tick's
            // recorded start state. Both give the same gradients bit for bit.
            const size_t xr_bytes = (size_t)K * xr_tick * 4;
            float* xrec = mem.room(xr_bytes) ? static_cast<float*>(mem.get(xr_bytes)) : nullptr;
            const double coef = 2.0 / ((double)ticks * ncells);
            const uint32_t gs = (n + 255) / 256;
            for (int seg = (int)nck - 1; seg >= 0; seg--) {
                const uint32_t c0 = (uint32_t)seg * K, c1 = std::min(ticks, c0 + K);
                restore((uint32_t)seg);
                // record j: V, gates before tick c0 + j; I_buffer of tick c0 + j (written by
                // the record of the tick after)
                auto recp = [&](uint32_t j) { return rec + (size_t)j * rstride; };
                auto ibp = [&](uint32_t j) { return recp(j) + P + (size_t)nslot * 2 * P; };
                for (uint32_t t = c0; t < c1; t++) {
                    double* r = recp(t - c0);
                    k_adj_record<<<gs, 256, 0, S.stream>>>(d, r, r + P, t > c0 ? ibp(t - c0 - 1) : nullptr);
                    d.xrec = xrec ? xrec + (size_t)(t - c0) * xr_tick : nullptr;
                    launch_tick(S, dt, false);
                    d.xrec = nullptr;
                }
                double* rl = recp(c1 - c0);
                k_adj_record<<<gs, 256, 0, S.stream>>>(d, rl, rl + P, ibp(c1 - c0 - 1));
                for (uint32_t t = c1; t-- > c0;) {
                    const double* post = recp(t + 1 - c0);
                    k_adj_seed<<<gc, 256, 0, S.stream>>>(post, dsoma, ncells, dtarget + (size_t)t * ncells, coef, a.lamV);
                    const double* pre = recp(t - c0);
                    a.recV = pre;
                    a.recG = pre + P;
                    a.recIb = ibp(t - c0);
                    a.xrec = xrec ? xrec + (size_t)(t - c0) * xr_tick : nullptr;
                    for (int c = 0; c < kTileClasses; c++) {
                        if (!S.class_ntiles[c]) continue;
                        if (c == 0) k_adjoint_tick<64><<<S.class_ntiles[0], 64, 0, S.stream>>>(d, S.tiles[0], a, t, dt);
                        if (c == 1) k_adjoint_tick<128><<<S.class_ntiles[1], 128, 0, S.stream>>>(d, S.tiles[1], a, t, dt);
                        if (c == 2) k_adjoint_tick<256><<<S.class_ntiles[2], 256, 0, S.stream>>>(d, S.tiles[2], a, t, dt);
                    }
                    k_adjoint_channels<<<gs, 256, 0, S.stream>>>(d, a, dt);
                }
                HIPCHK(hipGetLastError());
            }
            // gradients in the caller's order: segment h, slot j (0 past nslot); connection i
            if (gd_out) {
                k_grad_density_out<<<(n * BC_MAX_MODIFIERS + 255) / 256, 256, 0, S.stream>>>(
                    gd_out, a.gdens, n, P, nslot, S.host2dev_dev);
                HIPCHK(hipMemcpyAsync(grad_density, gd_out, (size_t)n * BC_MAX_MODIFIERS * 8, hipMemcpyDeviceToHost,
                                      S.stream));
            }
            if (gs_out && nconn) {
                k_gather_f64<<<(nconn + 255) / 256, 256, 0, S.stream>>>(gs_out, a.gscale, nconn, S.conn2dev_dev);
                HIPCHK(hipMemcpyAsync(grad_scale, gs_out, (size_t)nconn * 8, hipMemcpyDeviceToHost, S.stream));
            }
            HIPCHK(hipStreamSynchronize(S.stream));
            restore(nck);
            keep_final(true);
        }
        HIPCHK(hipStreamSynchronize(S.stream));
        after_ticks(S, ticks);
    } catch (...) {
        rc = -1;
    }
    S.d.xrec = nullptr;
    return rc;
}

}  // namespace bc

using bc::Sim;

// ================================================================== C ABI
extern "C" {

int gpu_backend_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void* gpu_backend_create_on_device(int device_index) {
    int ndev = gpu_backend_device_count();
    if (ndev == 0) return nullptr;
    if (device_index < 0 || device_index >= ndev) device_index = 0;
    try {
        Sim* s = new Sim();
        s->device = device_index;
        HIPCHK(hipSetDevice(device_index));
        (void)bc::scratch_setup(device_index);   // bchip_adjoint_noalloc refuses (-1) if this failed
        HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&s->stream2, hipStreamNonBlocking));
        for (auto& e : s->ev) HIPCHK(hipEventCreate(&e));
        for (auto& e : s->fork_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        s->stim_host = new GpuStimulusData();
        memset(s->stim_host, 0, sizeof(GpuStimulusData));
        return s;
    } catch (...) {
        return nullptr;
    }
}

void* gpu_backend_create(void) { return gpu_backend_create_on_device(0); }

void gpu_backend_destroy(void* sim) {
    if (sim) delete static_cast<Sim*>(sim);
}

int gpu_backend_load_bion(void* sim, const char* path) {
    if (!sim || !path) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    BionData b;
    int rc = bc::read_bion(path, &b);
    if (rc != 0) return rc;
    try {
        bc::plain_networks(S, b);
        return bc::load_into(S, &b);
    } catch (...) {
        bc::free_bion(&b);
        return BION_ERR_ALLOC;
    }
}

void gpu_backend_run_tick(void* sim, float dt) {
    if (!sim) return;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded) return;
    try {
        (void)hipSetDevice(S.device);
        bc::run_tick(S, dt);
    } catch (...) {
    }
}

void* gpu_backend_bion_data(void* sim) {
    if (!sim) return nullptr;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded) return nullptr;
    try {
        (void)hipSetDevice(S.device);
        bc::sync_host(S);
    } catch (...) {
    }
    return &S.host;
}

void gpu_backend_upload_stimuli(void* sim, const GpuStimulusData* src) {
    if (!sim || !src) return;
    Sim& S = *static_cast<Sim*>(sim);
    memcpy(S.stim_host, src, sizeof(GpuStimulusData));
    GpuStimulusData& g = *S.stim_host;
    auto cap = [](uint32_t& c) { if (c > BC_GPU_MAX_STIMULI) c = BC_GPU_MAX_STIMULI; };
    cap(g.current_stim_count); cap(g.vclamp_count); cap(g.ramp_count); cap(g.ion_inj_count); cap(g.lig_inj_count);
    Sim::NetStim& ns = S.net_stim[0];
    ns.cur.assign(g.current_stim, g.current_stim + g.current_stim_count);
    ns.ramp.assign(g.ramps, g.ramps + g.ramp_count);
    ns.vc.assign(g.vclamps, g.vclamps + g.vclamp_count);
    ns.ion.assign(g.ion_inj, g.ion_inj + g.ion_inj_count);
    ns.lig.assign(g.lig_inj, g.lig_inj + g.lig_inj_count);
    S.stim_present = true;
    S.stim_dirty = true;
}

int gpu_backend_get_device_info(int index, GpuDeviceInfo* info) {
    if (!info) return -1;
    int ndev = gpu_backend_device_count();
    if (index < 0 || index >= ndev) return -1;
    memset(info, 0, sizeof(GpuDeviceInfo));
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, index) != hipSuccess) return -1;
    strncpy(info->name, p.name, 127);
    info->vendor_id = 0x1002;
    info->is_cpu = 0;
    info->global_mem_bytes = p.totalGlobalMem;
    info->max_compute_units = p.multiProcessorCount;
    info->device_index = index;
    strncpy(info->backend_type, "hip", 15);
    return 0;
}

int gpu_backend_ibuf_offset(void) { return (int)offsetof(Segment, I_buffer); }
int gpu_backend_vm_offset(void) { return (int)offsetof(Segment, V_m); }
int gpu_backend_seg_size(void) { return (int)sizeof(Segment); }

// ---- plugin registration (src/plugin_manager.cpp:37-75) ----
void register_modifier_current(uint8_t type, ModifierFunc func) {
    std::lock_guard<std::mutex> lk(g_plugin_mu);
    if (type < 128) g_mod_table[type] = func;
}
void register_modulator_release(uint8_t type, ReleaseModulatorFunc func) {
    std::lock_guard<std::mutex> lk(g_plugin_mu);
    if (type < 128) g_rel_table[type] = func;
}

// ---- extensions ----
int bchip_load_bion_data(void* sim, const BionData* src) {
    if (!sim || !src || !src->segments || src->num_segments == 0) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    BionData b;
    memset(&b, 0, sizeof(b));
    b.header = src->header;
    b.num_neurons = src->num_neurons;
    b.num_segments = src->num_segments;
    b.num_connections = src->num_connections;
    b.num_synapses = src->num_synapses;
    try {
        b.neurons = new Neuron[b.num_neurons ? b.num_neurons : 1];
        b.segments = new Segment[b.num_segments];
        b.connections = new IntraConnection[b.num_connections ? b.num_connections : 1];
        b.synapses = new Synapse[b.num_synapses ? b.num_synapses : 1];
    } catch (...) {
        bc::free_bion(&b);
        return BION_ERR_ALLOC;
    }
    if (b.num_neurons) memcpy(b.neurons, src->neurons, sizeof(Neuron) * b.num_neurons);
    memcpy(b.segments, src->segments, sizeof(Segment) * b.num_segments);
    if (b.num_connections) memcpy(b.connections, src->connections, sizeof(IntraConnection) * b.num_connections);
    if (b.num_synapses) memcpy(b.synapses, src->synapses, sizeof(Synapse) * b.num_synapses);
    if (src->nuclei && src->num_nuclei) {
        b.nuclei = new (std::nothrow) Nucleus[src->num_nuclei];
        if (b.nuclei) {
            memcpy(b.nuclei, src->nuclei, sizeof(Nucleus) * src->num_nuclei);
            b.num_nuclei = src->num_nuclei;
        }
    }
    for (uint32_t i = 0; i < b.num_connections; i++)
        if (b.connections[i].segment_a >= b.num_segments || b.connections[i].segment_b >= b.num_segments) {
            bc::free_bion(&b); return BION_ERR_BAD_CONN;
        }
    for (uint32_t i = 0; i < b.num_synapses; i++)
        if (b.synapses[i].pre_segment >= b.num_segments || b.synapses[i].post_segment >= b.num_segments ||
            b.synapses[i].type >= 3 || b.synapses[i].post_receptor_count > BC_MAX_POST_RECEPTORS ||
            b.synapses[i].pre_autoreceptor_count > BC_MAX_PRE_AUTORECEPTORS) {
            bc::free_bion(&b); return BION_ERR_BAD_SYNAPSE;
        }
    for (uint32_t i = 0; i < b.num_segments; i++)
        if (b.segments[i].modifier_count > BC_MAX_MODIFIERS) { bc::free_bion(&b); return BION_ERR_BAD_MODCOUNT; }
    try {
        bc::plain_networks(S, b);
        return bc::load_into(S, &b);
    } catch (...) {
        bc::free_bion(&b);
        return BION_ERR_ALLOC;
    }
}

int bchip_read_bion(const char* path, BionData* out) {
    if (!path || !out) return -1;
    try {
        return bc::read_bion(path, out);
    } catch (...) {
        return BION_ERR_ALLOC;
    }
}

void bchip_free_bion(BionData* data) {
    if (data) bc::free_bion(data);
}

int bchip_write_bion(const char* path, const BionData* data) {
    if (!path || !data || !data->segments) return -1;
    return bc::write_bion(path, data);
}

int bchip_save_bion(void* sim, const char* path) {
    if (!sim || !path) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded) return -1;
    try { bc::sync_host(S); } catch (...) { return BION_ERR_WRITE_FAIL; }
    return bc::write_bion(path, &S.host);
}

int bchip_set_coherence(void* sim, int mode) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (mode != BCHIP_COHERENCE_COMPAT && mode != BCHIP_COHERENCE_FAST) return -1;
    try { bc::sync_host(S); } catch (...) { return -1; }
    S.coherence = mode;
    return 0;
}

int bchip_push_host_state(void* sim) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded) return -1;
    try {
        (void)hipSetDevice(S.device);
        bc::sync_host(S);
        bc::xfer(S, true);
    } catch (...) { return -1; }
    return 0;
}

int bchip_run_ticks(void* sim, float dt, uint32_t count) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded) return -1;
    try {
        (void)hipSetDevice(S.device);
        bc::run_ticks(S, dt, count);
    } catch (...) { return -1; }
    return 0;
}

int bchip_load_batch(void* sim, const BionData* const* nets, uint32_t count) {
    if (!sim || !nets || count == 0) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    uint64_t tn = 0, ts = 0, tc = 0, ty = 0;
    for (uint32_t k = 0; k < count; k++) {
        const BionData* x = nets[k];
        if (!x || !x->segments || x->num_segments == 0) return -1;
        tn += x->num_neurons; ts += x->num_segments; tc += x->num_connections; ty += x->num_synapses;
    }
    if (ts >= (1ull << 32) || tc >= (1ull << 32) || ty >= (1ull << 31)) return BION_ERR_ALLOC;
    BionData b;
    memset(&b, 0, sizeof(b));
    b.header = nets[0]->header;
    b.num_neurons = (uint32_t)tn; b.num_segments = (uint32_t)ts;
    b.num_connections = (uint32_t)tc; b.num_synapses = (uint32_t)ty;
    b.header.num_neurons = b.num_neurons; b.header.num_segments = b.num_segments;
    b.header.num_intra_connections = b.num_connections; b.header.num_synapses = b.num_synapses;
    b.header.num_nuclei = 0;
    try {
        b.neurons = new Neuron[tn ? tn : 1];
        b.segments = new Segment[ts];
        b.connections = new IntraConnection[tc ? tc : 1];
        b.synapses = new Synapse[ty ? ty : 1];
    } catch (...) {
        bc::free_bion(&b);
        return BION_ERR_ALLOC;
    }
    std::vector<uint32_t> so{0}, co{0}, yo{0};
    uint32_t on = 0;
    for (uint32_t k = 0; k < count; k++) {
        const BionData* x = nets[k];
        const uint32_t s0 = so.back(), c0 = co.back(), y0 = yo.back(), n = x->num_segments;
        for (uint32_t i = 0; i < x->num_neurons; i++) {
            Neuron r = x->neurons[i];
            r.segment_start += s0;
            r.soma_index += s0;
            b.neurons[on + i] = r;
        }
        memcpy(b.segments + s0, x->segments, sizeof(Segment) * n);
        for (uint32_t i = 0; i < x->num_connections; i++) {
            IntraConnection c = x->connections[i];
            if (c.segment_a >= n || c.segment_b >= n) { bc::free_bion(&b); return BION_ERR_BAD_CONN; }
            c.segment_a += s0;
            c.segment_b += s0;
            b.connections[c0 + i] = c;
        }
        for (uint32_t i = 0; i < x->num_synapses; i++) {
            const Synapse& y = x->synapses[i];
            if (y.pre_segment >= n || y.post_segment >= n || y.type >= 3 ||
                y.post_receptor_count > BC_MAX_POST_RECEPTORS || y.pre_autoreceptor_count > BC_MAX_PRE_AUTORECEPTORS) {
                bc::free_bion(&b);
                return BION_ERR_BAD_SYNAPSE;
            }
            Synapse& z = b.synapses[y0 + i];
            z = y;
            z.pre_segment += s0;
            z.post_segment += s0;
        }
        for (uint32_t i = 0; i < n; i++)
            if (x->segments[i].modifier_count > BC_MAX_MODIFIERS) { bc::free_bion(&b); return BION_ERR_BAD_MODCOUNT; }
        on += x->num_neurons;
        so.push_back(s0 + n);
        co.push_back(c0 + x->num_connections);
        yo.push_back(y0 + x->num_synapses);
    }
    try {
        S.net_seg_off = so;
        S.net_conn_off = co;
        S.net_syn_off = yo;
        S.net_stim.assign(count, Sim::NetStim{});
        S.stim_present = false;
        return bc::load_into(S, &b);
    } catch (...) {
        bc::free_bion(&b);
        return BION_ERR_ALLOC;
    }
}

uint32_t bchip_network_count(void* sim) {
    return sim ? (uint32_t)static_cast<Sim*>(sim)->net_stim.size() : 0;
}

int bchip_set_stimuli(void* sim, uint32_t net, const CurrentStimulus* cur, uint32_t ncur,
                      const RampStimulus* ramp, uint32_t nramp, const VoltageClamp* vc, uint32_t nvc,
                      const IonInjection* ion, uint32_t nion, const LigandInjection* lig, uint32_t nlig) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (net >= S.net_stim.size()) return -1;
    if ((ncur && !cur) || (nramp && !ramp) || (nvc && !vc) || (nion && !ion) || (nlig && !lig)) return -1;
    Sim::NetStim& ns = S.net_stim[net];
    ns.cur.assign(cur, cur + ncur);
    ns.ramp.assign(ramp, ramp + nramp);
    ns.vc.assign(vc, vc + nvc);
    ns.ion.assign(ion, ion + nion);
    ns.lig.assign(lig, lig + nlig);
    S.stim_present = true;
    S.stim_dirty = true;
    return 0;
}

int bchip_batch_read(void* sim, uint32_t net, Segment* segs, Synapse* syns, uint32_t* spikes) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded || net + 1 >= S.net_seg_off.size()) return -1;
    try {
        (void)hipSetDevice(S.device);
        bc::sync_host(S);
    } catch (...) { return -1; }
    const uint32_t s0 = S.net_seg_off[net], s1 = S.net_seg_off[net + 1];
    const uint32_t y0 = S.net_syn_off[net], y1 = S.net_syn_off[net + 1];
    if (segs) memcpy(segs, S.host.segments + s0, sizeof(Segment) * (s1 - s0));
    if (syns)
        for (uint32_t i = y0; i < y1; i++) {
            syns[i - y0] = S.host.synapses[i];
            syns[i - y0].pre_segment -= s0;
            syns[i - y0].post_segment -= s0;
        }
    if (spikes) {
        std::vector<uint32_t> dev(S.d.n);
        if (hipMemcpy(dev.data(), S.d.spikes, (size_t)S.d.n * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        for (uint32_t h = s0; h < s1; h++) spikes[h - s0] = dev[S.host2dev[h]];
    }
    return 0;
}

int bchip_get_schedule(void* sim, BchipSchedule* out) {
    if (!sim || !out) return -1;
    *out = static_cast<Sim*>(sim)->sched;
    return 0;
}

int bchip_set_schedule(void* sim, const BchipSchedule* in) {
    if (!sim || !in) return -1;
    auto blk = [](uint32_t b) { return b == 64 || b == 128 || b == 256; };
    if (!blk(in->modifiers_block) || !blk(in->plast_block) || !blk(in->rrp_block)) return -1;
    if (in->diffusion_tiles != 1 && in->diffusion_tiles != 2 && in->diffusion_tiles != 4) return -1;
    if (in->pre_grid < 1 || in->pre_grid > (1u << 20)) return -1;
    if (in->step1_chunks < 1 || in->step1_chunks > 16) return -1;
    static_cast<Sim*>(sim)->sched = *in;
    return 0;
}

int bchip_autotune(void* sim, float dt, uint32_t ticks_per_candidate) {
    if (!sim || ticks_per_candidate == 0) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded) return -1;
    // field, candidates, the phase its kernel is timed in (last_phase_ms index)
    struct Field { uint32_t BchipSchedule::*f; std::vector<uint32_t> cand; int phase; };
    const std::vector<Field> fields = {
        {&BchipSchedule::modifiers_block, {64, 128, 256}, 0},
        {&BchipSchedule::diffusion_tiles, {1, 2, 4}, 1},
        {&BchipSchedule::pre_grid, {1024, 4096, 16384, 65536}, 2},
        {&BchipSchedule::plast_block, {64, 128, 256}, 4},
        {&BchipSchedule::rrp_block, {64, 128, 256}, 5},
        {&BchipSchedule::step1_chunks, {1, 2, 4, 8}, -1},   // timed as steps 1 and 2.5 together
    };
    int ran = 0;
    const int timing = S.timing;
    try {
        (void)hipSetDevice(S.device);
        S.timing = 1;
        for (const Field& fd : fields) {
            if ((fd.phase == 1 || fd.phase == -1) && (!S.d.has_conn || !S.class_ntiles[0])) continue;
            if (fd.phase >= 2 && !S.d.ns) continue;
            uint32_t best = S.sched.*(fd.f);
            float best_ms = 1e30f;
            for (uint32_t c : fd.cand) {
                S.sched.*(fd.f) = c;
                float acc = 0.0f;
                for (uint32_t t = 0; t < ticks_per_candidate; t++) {
                    bc::run_tick(S, dt);
                    acc += fd.phase >= 0 ? S.last_ms[fd.phase] : S.last_ms[0] + S.last_ms[1];
                    ran++;
                }
                if (acc < best_ms) { best_ms = acc; best = c; }
            }
            S.sched.*(fd.f) = best;
        }
    } catch (...) {
        S.timing = timing;
        return -1;
    }
    S.timing = timing;
    return ran;
}

int bchip_set_jacobi(void* sim, double tol, int max_sweeps) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (tol > 0.0 && (max_sweeps < 1 || max_sweeps > 100000)) return -1;
    S.jac_res = tol > 0.0;
    S.jac_tol = S.jac_res ? tol : 0.0;
    S.jac_max = S.jac_res ? max_sweeps : 50;
    S.d.jac_tol = S.jac_tol;
    S.d.jac_max = S.jac_max;
    return 0;
}

// Device evaluation of the tick's exact-arithmetic helpers (see bchip_eval_math).
__global__ void k_eval_math(int fn, const double* a, const double* b, double* out, uint32_t n) {
    bc::exp_table_stage();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double r = 0.0;
        switch (fn) {
            case 0: r = bc::exp(a[i]); break;
            case 1: r = bc::exp_t(a[i], bc::ExpTabGlobal{}); break;
            case 2: r = bc::div_rcp(a[i], b[i], bc::rcp_prep(b[i])); break;
            case 3: {
                const float x = (float)a[i], y = (float)b[i];
                r = bc::div_wide(x, y, 1.0 / (double)y);
                break;
            }
            case 4: {
                const double y = (double)(float)b[i];
                r = bc::div_f24(a[i], y, 1.0 / y);
                break;
            }
            case 5: r = bc::bernoulli((float)a[i], bc::ExpTabGlobal{}); break;
            default: break;
        }
        out[i] = r;
    }
}

int bchip_eval_math(int fn, const double* a, const double* b, double* out, uint32_t n) {
    if (fn < 0 || fn > 5 || !a || !out || ((fn >= 2 && fn <= 4) && !b)) return -1;
    if (!n) return 0;
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    int rc = 0;
    try {
        const size_t B = (size_t)n * 8;
        HIPCHK(hipMalloc(&da, B));
        HIPCHK(hipMalloc(&dout, B));
        HIPCHK(hipMemcpy(da, a, B, hipMemcpyHostToDevice));
        if (b) {
            HIPCHK(hipMalloc(&db, B));
            HIPCHK(hipMemcpy(db, b, B, hipMemcpyHostToDevice));
        }
        k_eval_math<<<std::min<uint32_t>((n + 255) / 256, 8192), 256>>>(fn, da, db, dout, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dout, B, hipMemcpyDeviceToHost));
    } catch (...) {
        rc = -1;
    }
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dout);
    return rc;
}

int bchip_allow_unsupported(void* sim, int on) {
    if (!sim) return -1;
    static_cast<Sim*>(sim)->allow_unsupported = on != 0;
    return 0;
}

int bchip_set_protein_mapping(void* sim, int on) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    S.protein_map = on != 0;
    if (!S.loaded || !S.d.ns) return 0;
    HIPCHK(hipSetDevice(S.device));
    std::vector<float> rmax, prel, prim0, rrecs;
    release_params(S, rmax, prel, prim0, rrecs);
    const size_t b = sizeof(float) * rmax.size();
    if (S.d.y_rrec)
        HIPCHK(hipMemcpyAsync(const_cast<float*>(S.d.y_rrec), rrecs.data(), b, hipMemcpyHostToDevice, S.stream));
    HIPCHK(hipMemcpyAsync(const_cast<float*>(S.d.y_rrp_max), rmax.data(), b, hipMemcpyHostToDevice, S.stream));
    HIPCHK(hipMemcpyAsync(const_cast<float*>(S.d.y_prel), prel.data(), b, hipMemcpyHostToDevice, S.stream));
    HIPCHK(hipMemcpyAsync(const_cast<float*>(S.d.y_prim0), prim0.data(), b, hipMemcpyHostToDevice, S.stream));
    HIPCHK(hipStreamSynchronize(S.stream));
    S.h_rmax = std::move(rmax);   // the hazards depend on the effective RRP_max
    S.ord_dt = NAN;
    return 0;
}

int bchip_set_cable_mode(void* sim, int on) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (on && S.loaded && S.d.ns) return -1;   // receptors would move ions: not a cable
    S.cable = on != 0;
    S.d.cable = S.cable ? 1 : 0;
    return 0;
}

int bchip_set_coupling_scale(void* sim, const float* scale, uint32_t n) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded) return -1;
    if (!scale) { S.d.conn_scale = nullptr; return 0; }
    if (n != S.host.num_connections) return -1;
    try {
        (void)hipSetDevice(S.device);
        std::vector<float> dev(std::max<uint32_t>(n, 1), 1.0f);
        for (uint32_t i = 0; i < n; i++) dev[S.conn2dev[i]] = scale[i];
        if (!S.conn_scale_dev) S.conn_scale_dev = S.dalloc<float>(dev.size());
        HIPCHK(hipMemcpyAsync(S.conn_scale_dev, dev.data(), dev.size() * 4, hipMemcpyHostToDevice, S.stream));
        HIPCHK(hipStreamSynchronize(S.stream));
        S.d.conn_scale = S.conn_scale_dev;
    } catch (...) { return -1; }
    return 0;
}

int bchip_adjoint(void* sim, float dt, uint32_t ticks, uint32_t checkpoint_every, const uint32_t* soma,
                  uint32_t ncells, const float* target, double* loss, double* grad_density, double* grad_scale,
                  float* trace) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    bc::AdjHeap mem(S);
    return bc::adjoint_run(S, mem, dt, ticks, checkpoint_every, soma, ncells, target, loss, grad_density,
                           grad_scale, trace);
}

int bchip_adjoint_noalloc(void* sim, float dt, uint32_t ticks, uint32_t checkpoint_every, const uint32_t* soma,
                          uint32_t ncells, const float* target, double* loss, double* grad_density,
                          double* grad_scale, float* trace) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    // building stimulus lists (a pending upload) and refreshing the host records (COMPAT)
    // allocate: refused here
    if (S.coherence != BCHIP_COHERENCE_FAST || S.stim_dirty) return -1;
    void* base = bc::scratch_of(S.device);
    if (!base) return -1;
    bc::AdjScratch mem(S, base, bc::kScratchBytes);
    return bc::adjoint_run(S, mem, dt, ticks, checkpoint_every, soma, ncells, target, loss, grad_density,
                           grad_scale, trace);
}

int bchip_set_stimulus_convention(void* sim, int convention) {
    if (!sim || convention < 0 || convention > 1) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    S.stim_tick_shift = (uint32_t)convention;
    S.d.stim.tick_shift = S.stim_tick_shift;
    return 0;
}

int bchip_set_adjoint_memory(void* sim, uint64_t max_bytes) {
    if (!sim) return -1;
    static_cast<Sim*>(sim)->adj_record_cap = max_bytes;
    return 0;
}

int bchip_set_graphs(void* sim, int on) {
    if (!sim) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    S.use_graphs = on != 0;
    if (!S.use_graphs) S.drop_graph();
    return 0;
}

uint64_t bchip_tick_count(void* sim) { return sim ? static_cast<Sim*>(sim)->tick : 0; }

int bchip_spike_counts(void* sim, uint32_t* out, uint32_t n) {
    if (!sim || !out) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    if (!S.loaded || n > S.d.n) return -1;
    std::vector<uint32_t> dev(S.d.n);
    if (hipMemcpy(dev.data(), S.d.spikes, (size_t)S.d.n * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    for (uint32_t h = 0; h < n; h++) out[h] = dev[S.host2dev[h]];
    return 0;
}

int bchip_delay_slots(void* sim, uint8_t* out, uint64_t cap) {
    // DelayedRelease layout (delay_buffer.h:15-26): u64 arrival, f32 amount, u8 nt, pad[3];
    // 4 slots per synapse in original synapse order.
    if (!sim || !out) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    const uint32_t ns = S.d.ns;
    if (cap < (uint64_t)ns * 64) return -1;
    std::vector<uint64_t> arr((size_t)4 * ns);
    std::vector<float> amt((size_t)4 * ns);
    std::vector<uint8_t> nt((size_t)4 * ns), mask(ns);
    if (ns) {
        (void)hipMemcpy(arr.data(), S.d.d_arr, arr.size() * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(amt.data(), S.d.d_amt, amt.size() * 4, hipMemcpyDeviceToHost);
        (void)hipMemcpy(nt.data(), S.d.d_nt, nt.size(), hipMemcpyDeviceToHost);
        (void)hipMemcpy(mask.data(), S.d.d_mask, ns, hipMemcpyDeviceToHost);
    }
    memset(out, 0, (size_t)ns * 64);
    for (uint32_t k = 0; k < ns; k++) {
        uint8_t* rec = out + (size_t)S.slot2syn[k] * 64;
        for (int s = 0; s < 4; s++) {
            uint8_t* e = rec + s * 16;
            uint64_t a = (mask[k] & (1u << s)) ? arr[(size_t)s * ns + k] : 0;
            memcpy(e, &a, 8);
            memcpy(e + 8, &amt[(size_t)s * ns + k], 4);
            e[12] = nt[(size_t)s * ns + k];
        }
    }
    return 0;
}

int bchip_stats(void* sim, BchipStats* out) {
    if (!sim || !out) return -1;
    Sim& S = *static_cast<Sim*>(sim);
    memset(out, 0, sizeof(*out));
    out->num_segments = S.d.n;
    out->num_connections = S.host.num_connections;
    out->num_synapses = S.d.ns;
    out->num_tiles = S.ntiles;
    out->modifier_slots = S.d.nslot;
    out->receptor_slots = S.d.nrec;
    out->unsupported_modifiers = S.unsupported_mod;
    out->unsupported_synapses = S.unsupported_syn;
    out->order_hazard_synapses = S.hazard_syn;
    out->lds_bytes_per_tile = (uint32_t)S.lds_bytes;
    for (int k = 0; k < 6; k++) out->last_phase_ms[k] = S.last_ms[k];
    out->ordered_events = S.ord_events;
    out->ordered_segments = S.ord_segments;
    out->big_islands = S.big.count;
    return 0;
}

int bchip_set_phase_timing(void* sim, int on) {
    if (!sim) return -1;
    static_cast<Sim*>(sim)->timing = on;
    return 0;
}

void* bchip_stream(void* sim) { return sim ? static_cast<Sim*>(sim)->stream : nullptr; }

int bchip_plugin_load(const char* path) {
    // Mirrors plugin_load (src/plugin_manager.cpp:208-243): RTLD_GLOBAL so the plugin binds to
    // this library's register_* symbols, then biocore_plugin_init(). Registered CPU functions
    // are recorded; device ticks keep the built-in device behaviour for those IDs.
    if (!path) return -1;
    void* h = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
    if (!h) return -2;
    using InitFn = int (*)(void);
    InitFn init = reinterpret_cast<InitFn>(dlsym(h, "biocore_plugin_init"));
    if (!init) return -3;
    return init();
}

int bchip_registered_modifier(uint8_t type) {
    std::lock_guard<std::mutex> lk(g_plugin_mu);
    return type < 128 && g_mod_table[type] != nullptr;
}

int bchip_modifier_supported(uint8_t type) { return bc::modifier_supported(type) ? 1 : 0; }

}  // extern "C"

// C++-mangled registration entry points plugins link against
// (_Z22register_user_modifierhPFdP8ModifierP7SegmentfE, _Z23register_user_modulatorhPFfP8ModifierPffE).
void register_user_modifier(uint8_t type, ModifierFunc func) {
    std::lock_guard<std::mutex> lk(g_plugin_mu);
    if (type <= 127) g_mod_table[type] = func;
}
void register_user_modulator(uint8_t type, ReleaseModulatorFunc func) {
    std::lock_guard<std::mutex> lk(g_plugin_mu);
    if (type <= 127) g_rel_table[type] = func;
}
