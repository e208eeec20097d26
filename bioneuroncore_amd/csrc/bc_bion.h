// bc_bion.h — BION file reader (format versions 3-14) and writer (v14), host-only.
//
// The reader follows the reference's host loader (src/bion_loader_host.cpp:26-425 with the
// record migrations of include/bion_format.h:30-450): the header is 68 bytes up to v4, 84 up to
// v9 and 100 from v10; records are read in the layout each version (and, from v10, each
// header file_sz_* field) names and converted field by field into the v14 records. Error codes,
// validation and the intra_connection_ids fill are the reference's. Where the reference reads
// a legacy record through today's structure, this reader does the same, so a file gives the
// same records in both:
//   - v4-v6 segments are 1692-byte records whose 16 modifiers have the current 88-byte layout,
//     and whose first 196 bytes (ids .. transmitters, float fields) land byte for byte on the
//     first 196 bytes of the v14 segment;
//   - v7-v9 segments (and v10 files with 1692/1696-byte segments) carry 5 ions, but 7 are
//     converted: ions 5 and 6 are the two floats after the 5-element array (the first two
//     extracellular ions for the intracellular side, the first two intracellular transmitters
//     for the extracellular side);
//   - v3-v6 synapses are the first 824 bytes of a v14 synapse;
//   - v6 nuclei are the first 332 bytes of a v14 nucleus.
// Bytes the reference leaves unwritten in those cases (the rest of a v3-v6 synapse beyond the
// vesicle pool fields, of a v6 nucleus, and the reserved bytes of v3-v8 connections) are zero
// here.
#pragma once
#include <cstdio>
#include <cstring>
#include <ctime>
#include <new>
#include <vector>

namespace bc {
namespace legacy {

#pragma pack(push, 1)
struct ModV3 {          // v3 modifier, 41 bytes
    uint8_t type;
    float density;
    float state[8];
    uint8_t active_vars, flags, res[2];
};
struct ModF {           // v4-v12 float modifier, 44 bytes
    uint8_t type, pad[3];
    float density;
    float state[8];
    uint8_t active_vars, flags, res[2];
};
struct SegV3 {          // 852 bytes
    uint32_t id, neuron_id;
    Vec3 position;
    float length, radius, V_m;
    float intra_ions[4], extra_ions[4];
    float neurotransmitter[12];
    ModV3 modifiers[16];
    uint8_t modifier_count, res[3];
    float ATP_level, temperature, I_buffer;
    uint32_t intra_connection_ids[16];
    uint8_t intra_connection_count, res2[3];
};
struct SegV4 {          // 1692 bytes as the reference reads v4-v6 segments
    uint32_t id, neuron_id;
    Vec3 position, direction;
    float length, radius, V_m;
    float intra_ions[5], extra_ions[5];
    float intra_nt[14], neurotransmitter[14];
    Modifier modifiers[16];
    uint8_t modifier_count, segment_type, res1[2];
    float ATP_level, temperature, I_buffer, transcription_signal;
    uint32_t intra_connection_ids[16];
    uint8_t intra_connection_count, res2[3];
};
struct SegV7 {          // 1692 bytes: 5 ions, 14 transmitters, 32 float modifiers
    uint32_t id, neuron_id;
    Vec3 position, direction;
    float length, radius, V_m;
    float intra_ions[5], extra_ions[5];
    float intra_nt[14], neurotransmitter[14];
    ModF modifiers[32];
    uint8_t modifier_count, segment_type, res1[2];
    float ATP_level, temperature, I_buffer, transcription_signal;
    uint32_t intra_connection_ids[16];
    uint8_t intra_connection_count, res2[3];
};
struct SegV10 {         // 1704 bytes: 5 ions, 15 transmitters, ER Ca
    uint32_t id, neuron_id;
    Vec3 position, direction;
    float length, radius, V_m;
    float intra_ions[5], extra_ions[5];
    float intra_nt[15], neurotransmitter[15];
    ModF modifiers[32];
    uint8_t modifier_count, segment_type, res1[2];
    float ATP_level, temperature, I_buffer, transcription_signal;
    uint32_t intra_connection_ids[16];
    uint8_t intra_connection_count, res2[3];
    float ER_Ca2;
};
struct SegV13 {         // 3208 bytes: double state, I_buffer not 8-aligned
    uint32_t id, neuron_id;
    Vec3 position, direction;
    float length, radius;
    double V_m;
    double intra_ions[8], extra_ions[8];
    float intra_nt[15], neurotransmitter[15];
    Modifier modifiers[32];
    uint8_t modifier_count, segment_type, res1[2];
    float ATP_level, temperature;
    double I_buffer;
    float transcription_signal;
    uint32_t intra_connection_ids[16];
    uint8_t intra_connection_count, res2[3];
    float ER_Ca2;
};
struct ConnV8 {         // 12 bytes
    uint32_t segment_a, segment_b;
    float coupling_resistance;
};
struct SynV7 {          // 836 bytes: float modifiers, 14 cleft species, vesicle pool
    uint32_t id, pre_segment, post_segment;
    uint8_t type, res[3];
    float weight, delay, pre_area, post_area;
    ModF pre_autoreceptors[8];
    uint8_t pre_autoreceptor_count, res2[3];
    ModF post_receptors[8];
    uint8_t post_receptor_count, res3[3];
    float release_probability, release_amount;
    float cleft_neurotransmitter[14];
    float facilitation, depression, tau_clearance;
    uint8_t refractory, res4[3];
    float RRP, RRP_max, RRP_recovery;
};
struct SynV13 {         // 1560 bytes: double modifiers, receptors not 8-aligned
    uint32_t id, pre_segment, post_segment;
    uint8_t type, res[3];
    float weight, delay, pre_area, post_area;
    Modifier pre_autoreceptors[8];
    uint8_t pre_autoreceptor_count, res2[3];
    Modifier post_receptors[8];
    uint8_t post_receptor_count, res3[3];
    float release_probability, release_amount;
    float cleft_neurotransmitter[15];
    float facilitation, depression, tau_clearance;
    uint8_t refractory, res4[3];
    float RRP, RRP_max, RRP_recovery, bouton_Ca, spine_Ca, spine_camkii, spine_calcineurin;
};
struct NucV5 {          // 204 bytes: 20 genes
    uint32_t neuron_id;
    uint8_t enabled, consolidation_mode, pad[2];
    uint32_t consolidation_interval_ticks, last_consolidation_tick;
    float gene_activity[20], protein_level[20];
    float protein_synthesis_rate, pcreb_level, erk_level, nfkb_level, mtorc1_level, amino_acid_level;
    uint8_t res[4];
};
struct NucV46 {         // 412 bytes: 46 genes
    uint32_t neuron_id;
    uint8_t enabled, consolidation_mode, pad[2];
    uint32_t consolidation_interval_ticks, last_consolidation_tick;
    float gene_activity[46], protein_level[46];
    float protein_synthesis_rate, pcreb_level, erk_level, nfkb_level, mtorc1_level, amino_acid_level;
    uint8_t res[4];
};
#pragma pack(pop)
static_assert(sizeof(ModV3) == 41 && sizeof(ModF) == 44, "legacy modifier sizes");
static_assert(sizeof(SegV3) == 852 && sizeof(SegV4) == 1692 && sizeof(SegV7) == 1692, "legacy segment sizes");
static_assert(sizeof(SegV10) == 1704 && sizeof(SegV13) == 3208, "legacy segment sizes");
static_assert(sizeof(ConnV8) == 12 && sizeof(SynV7) == 836 && sizeof(SynV13) == 1560, "legacy sizes");
static_assert(sizeof(NucV5) == 204 && sizeof(NucV46) == 412, "legacy nucleus sizes");
static_assert(offsetof(SegV4, modifiers) == 196, "v4 segment prefix");

constexpr uint32_t kSegV7 = 1692, kSegV10Er = 1696, kSegV10 = 1704, kSegV13 = 3208;
constexpr uint32_t kSynV6 = 824, kSynV7 = 836, kSynV13 = 1560, kNucV6 = 332, kNucV46 = 412;

inline void mod_from(Modifier& m, const ModF& o) {
    m.type = o.type;
    m.density = o.density;
    for (int k = 0; k < 8; k++) m.state[k] = o.state[k];
    m.active_vars = o.active_vars;
    m.flags = o.flags;
}
// Fields every segment layout from v7 on has under the same names.
template <class L>
inline void seg_common(Segment& s, const L& o) {
    s.id = o.id;
    s.neuron_id = o.neuron_id;
    s.position = o.position;
    s.direction = o.direction;
    s.length = o.length;
    s.radius = o.radius;
    s.V_m = o.V_m;
    s.segment_type = o.segment_type;
    s.ATP_level = o.ATP_level;
    s.temperature = o.temperature;
    s.I_buffer = o.I_buffer;
    s.transcription_signal = o.transcription_signal;
    for (int j = 0; j < 16; j++) s.intra_connection_ids[j] = o.intra_connection_ids[j];
    s.intra_connection_count = o.intra_connection_count;
    s.modifier_count = o.modifier_count > 32 ? 32 : o.modifier_count;
}
inline void seg_from_v7(Segment& s, const SegV7& o, const uint8_t* raw, float er) {
    memset(&s, 0, sizeof s);
    seg_common(s, o);
    // seven ions read from 5-element arrays: the two floats after each array follow
    float f;
    for (int j = 0; j < 7; j++) {
        memcpy(&f, raw + offsetof(SegV7, intra_ions) + 4 * j, 4);
        s.intra_ions[j] = f;
        memcpy(&f, raw + offsetof(SegV7, extra_ions) + 4 * j, 4);
        s.extra_ions[j] = f;
    }
    for (int j = 0; j < 14; j++) {
        s.intra_nt[j] = o.intra_nt[j];
        s.neurotransmitter[j] = o.neurotransmitter[j];
    }
    for (int j = 0; j < s.modifier_count; j++) mod_from(s.modifiers[j], o.modifiers[j]);
    s.ER_Ca2 = er;
}
inline void seg_from_v10(Segment& s, const SegV10& o) {
    memset(&s, 0, sizeof s);
    seg_common(s, o);
    for (int j = 0; j < 5; j++) {
        s.intra_ions[j] = o.intra_ions[j];
        s.extra_ions[j] = o.extra_ions[j];
    }
    for (int j = 0; j < 15; j++) {
        s.intra_nt[j] = o.intra_nt[j];
        s.neurotransmitter[j] = o.neurotransmitter[j];
    }
    for (int j = 0; j < s.modifier_count; j++) mod_from(s.modifiers[j], o.modifiers[j]);
    s.ER_Ca2 = o.ER_Ca2;
}
inline void seg_from_v13(Segment& s, const SegV13& o) {
    memset(&s, 0, sizeof s);
    seg_common(s, o);
    for (int j = 0; j < 8; j++) {
        s.intra_ions[j] = o.intra_ions[j];
        s.extra_ions[j] = o.extra_ions[j];
    }
    for (int j = 0; j < 15; j++) {
        s.intra_nt[j] = o.intra_nt[j];
        s.neurotransmitter[j] = o.neurotransmitter[j];
    }
    for (int j = 0; j < s.modifier_count; j++) s.modifiers[j] = o.modifiers[j];
    s.ER_Ca2 = o.ER_Ca2;
}
inline void seg_from_v4(Segment& s, const SegV4& o) {
    memset(&s, 0, sizeof s);
    memcpy(&s, &o, offsetof(SegV4, modifiers));
    s.modifier_count = o.modifier_count > 16 ? 16 : o.modifier_count;
    for (int j = 0; j < s.modifier_count; j++) s.modifiers[j] = o.modifiers[j];
    s.segment_type = o.segment_type;
    s.ATP_level = o.ATP_level;
    s.temperature = o.temperature;
    s.I_buffer = o.I_buffer;
    s.transcription_signal = o.transcription_signal;
    for (int j = 0; j < 16; j++) s.intra_connection_ids[j] = o.intra_connection_ids[j];
    s.intra_connection_count = o.intra_connection_count;
}
inline void seg_from_v3(Segment& s, const SegV3& o) {
    memset(&s, 0, sizeof s);
    s.id = o.id;
    s.neuron_id = o.neuron_id;
    s.position = o.position;
    s.direction = Vec3{0.0f, 1.0f, 0.0f};
    s.length = o.length;
    s.radius = o.radius;
    s.V_m = o.V_m;
    for (int j = 0; j < 4; j++) {
        s.intra_ions[j] = o.intra_ions[j];
        s.extra_ions[j] = o.extra_ions[j];
    }
    for (int j = 0; j < 12; j++) s.neurotransmitter[j] = o.neurotransmitter[j];
    s.modifier_count = o.modifier_count;
    s.ATP_level = o.ATP_level;
    s.temperature = o.temperature;
    s.I_buffer = o.I_buffer;
    for (int j = 0; j < 16; j++) s.intra_connection_ids[j] = o.intra_connection_ids[j];
    s.intra_connection_count = o.intra_connection_count;
    for (int j = 0; j < s.modifier_count && j < 16; j++) {
        Modifier& m = s.modifiers[j];
        m.type = o.modifiers[j].type;
        m.density = o.modifiers[j].density;
        for (int k = 0; k < 8; k++) m.state[k] = o.modifiers[j].state[k];
        m.active_vars = o.modifiers[j].active_vars;
        m.flags = o.modifiers[j].flags;
    }
}
template <class L, class ModCopy>
inline void syn_common(Synapse& y, const L& o, ModCopy mc) {
    memset(&y, 0, sizeof y);
    y.id = o.id;
    y.pre_segment = o.pre_segment;
    y.post_segment = o.post_segment;
    y.type = o.type;
    y.weight = o.weight;
    y.delay = o.delay;
    y.pre_area = o.pre_area;
    y.post_area = o.post_area;
    y.pre_autoreceptor_count = o.pre_autoreceptor_count > 8 ? 8 : o.pre_autoreceptor_count;
    for (int j = 0; j < y.pre_autoreceptor_count; j++) mc(y.pre_autoreceptors[j], o.pre_autoreceptors[j]);
    y.post_receptor_count = o.post_receptor_count > 8 ? 8 : o.post_receptor_count;
    for (int j = 0; j < y.post_receptor_count; j++) mc(y.post_receptors[j], o.post_receptors[j]);
    y.release_probability = o.release_probability;
    y.release_amount = o.release_amount;
    y.facilitation = o.facilitation;
    y.depression = o.depression;
    y.tau_clearance = o.tau_clearance;
    y.refractory = o.refractory;
    y.RRP = o.RRP;
    y.RRP_max = o.RRP_max;
    y.RRP_recovery = o.RRP_recovery;
}
inline void syn_from_v7(Synapse& y, const SynV7& o) {
    syn_common(y, o, [](Modifier& m, const ModF& f) { mod_from(m, f); });
    for (int j = 0; j < 14; j++) y.cleft_neurotransmitter[j] = o.cleft_neurotransmitter[j];
}
inline void syn_from_v13(Synapse& y, const SynV13& o) {
    syn_common(y, o, [](Modifier& m, const Modifier& f) { m = f; });
    for (int j = 0; j < 15; j++) y.cleft_neurotransmitter[j] = o.cleft_neurotransmitter[j];
    y.bouton_Ca = o.bouton_Ca;
    y.spine_Ca = o.spine_Ca;
    y.spine_camkii = o.spine_camkii;
    y.spine_calcineurin = o.spine_calcineurin;
}
template <class L>
inline void nuc_from(Nucleus& n, const L& o, int genes) {
    memset(&n, 0, sizeof n);
    n.neuron_id = o.neuron_id;
    n.enabled = o.enabled;
    n.consolidation_mode = o.consolidation_mode;
    n.consolidation_interval_ticks = o.consolidation_interval_ticks;
    n.last_consolidation_tick = o.last_consolidation_tick;
    for (int g = 0; g < genes; g++) {
        n.gene_activity[g] = o.gene_activity[g];
        n.protein_level[g] = o.protein_level[g];
    }
    n.protein_synthesis_rate = o.protein_synthesis_rate;
    n.pcreb_level = o.pcreb_level;
    n.erk_level = o.erk_level;
    n.nfkb_level = o.nfkb_level;
    n.mtorc1_level = o.mtorc1_level;
    n.amino_acid_level = o.amino_acid_level;
}

}  // namespace legacy

static void free_bion(BionData* b) {
    delete[] b->neurons;
    delete[] b->segments;
    delete[] b->connections;
    delete[] b->synapses;
    delete[] b->nuclei;
    memset(b, 0, sizeof(BionData));
}

// Reads n records of `disk` bytes each and converts each with conv(dst_index, raw bytes).
template <class Conv>
static bool read_records(FILE* f, uint32_t n, size_t disk, Conv conv) {
    std::vector<uint8_t> raw(disk);
    for (uint32_t i = 0; i < n; i++) {
        if (fread(raw.data(), disk, 1, f) != 1) return false;
        conv(i, raw.data());
    }
    return true;
}
template <class T>
static T rec_as(const uint8_t* p) {
    T t;
    memcpy(&t, p, sizeof t);
    return t;
}

static int read_bion(const char* path, BionData* b) {
    using namespace legacy;
    memset(b, 0, sizeof(BionData));
    FILE* f = fopen(path, "rb");
    if (!f) return BION_ERR_FILE_OPEN;
    NetworkHeader& h = b->header;
    if (fread(&h, 68, 1, f) != 1) { fclose(f); return BION_ERR_FILE_READ; }
    if (h.version >= 5) {
        const size_t ext = h.version >= 10 ? sizeof(NetworkHeader) - 68 : 16;
        if (fread(reinterpret_cast<uint8_t*>(&h) + 68, ext, 1, f) != 1) { fclose(f); return BION_ERR_FILE_READ; }
    }
    if (h.magic != BION_MAGIC_U32) { fclose(f); return BION_ERR_BAD_MAGIC; }
    if (h.version < 3 || h.version > BION_VERSION_CURRENT) { fclose(f); return BION_ERR_BAD_VERSION; }
    if (h.num_segments == 0) { fclose(f); return BION_ERR_ZERO_SEGS; }
    const uint32_t v = h.version;
    b->num_neurons = h.num_neurons;
    b->num_segments = h.num_segments;
    b->num_connections = h.num_intra_connections;
    b->num_synapses = h.num_synapses;
    b->neurons = new (std::nothrow) Neuron[b->num_neurons ? b->num_neurons : 1]();
    b->segments = new (std::nothrow) Segment[b->num_segments]();
    b->connections = new (std::nothrow) IntraConnection[b->num_connections ? b->num_connections : 1]();
    b->synapses = new (std::nothrow) Synapse[b->num_synapses ? b->num_synapses : 1]();
    if (!b->neurons || !b->segments || !b->connections || !b->synapses) {
        free_bion(b); fclose(f); return BION_ERR_ALLOC;
    }
    auto fail = [&](int rc) { free_bion(b); fclose(f); return rc; };
    if (b->num_neurons && fread(b->neurons, sizeof(Neuron), b->num_neurons, f) != b->num_neurons)
        return fail(BION_ERR_FILE_READ);

    // ---- segments ----
    Segment* sg = b->segments;
    const uint32_t ns = b->num_segments;
    bool ok;
    if (v <= 3) {
        ok = read_records(f, ns, sizeof(SegV3), [&](uint32_t i, const uint8_t* r) { seg_from_v3(sg[i], rec_as<SegV3>(r)); });
    } else if (v <= 6) {
        ok = read_records(f, ns, sizeof(SegV4), [&](uint32_t i, const uint8_t* r) { seg_from_v4(sg[i], rec_as<SegV4>(r)); });
    } else {
        // v7-v9 always 1692; from v10 the header's size selects the layout
        const uint32_t fs = v < 10 ? kSegV7 : h.file_sz_segment;
        if (fs == 0 || fs == sizeof(Segment)) {
            ok = fread(sg, sizeof(Segment), ns, f) == ns;
        } else if (fs == kSegV7 || fs == kSegV10Er) {
            ok = read_records(f, ns, fs, [&](uint32_t i, const uint8_t* r) {
                float er = 0.0f;
                if (fs == kSegV10Er) memcpy(&er, r + kSegV7, 4);
                seg_from_v7(sg[i], rec_as<SegV7>(r), r, er);
            });
        } else if (fs == kSegV10) {
            ok = read_records(f, ns, fs, [&](uint32_t i, const uint8_t* r) { seg_from_v10(sg[i], rec_as<SegV10>(r)); });
        } else if (fs == kSegV13) {
            ok = read_records(f, ns, fs, [&](uint32_t i, const uint8_t* r) { seg_from_v13(sg[i], rec_as<SegV13>(r)); });
        } else {   // another size: the common prefix, the rest zero
            ok = read_records(f, ns, fs, [&](uint32_t i, const uint8_t* r) {
                memset(&sg[i], 0, sizeof(Segment));
                memcpy(&sg[i], r, fs < sizeof(Segment) ? fs : sizeof(Segment));
            });
        }
    }
    if (!ok) return fail(BION_ERR_FILE_READ);

    // ---- connections ----
    if (b->num_connections) {
        if (v <= 8) {
            ok = read_records(f, b->num_connections, sizeof(ConnV8), [&](uint32_t i, const uint8_t* r) {
                const ConnV8 o = rec_as<ConnV8>(r);
                IntraConnection& c = b->connections[i];
                memset(&c, 0, sizeof c);
                c.segment_a = o.segment_a;
                c.segment_b = o.segment_b;
                c.is_intracellular = o.coupling_resistance != 0.0f ? 1 : 0;   // resistance becomes a flag
            });
        } else {
            ok = fread(b->connections, sizeof(IntraConnection), b->num_connections, f) == b->num_connections;
        }
        if (!ok) return fail(BION_ERR_FILE_READ);
    }

    // ---- synapses ----
    if (b->num_synapses) {
        Synapse* sy = b->synapses;
        const uint32_t ny = b->num_synapses;
        if (v <= 6) {
            ok = read_records(f, ny, kSynV6, [&](uint32_t i, const uint8_t* r) {
                memset(&sy[i], 0, sizeof(Synapse));
                memcpy(&sy[i], r, kSynV6);
                sy[i].RRP = 10.0f;
                sy[i].RRP_max = 10.0f;
                sy[i].RRP_recovery = 0.02f;
            });
        } else {
            const uint32_t fs = v < 10 ? kSynV7 : h.file_sz_synapse;
            if (fs == 0 || fs == sizeof(Synapse)) {
                ok = fread(sy, sizeof(Synapse), ny, f) == ny;
            } else if (fs == kSynV7) {
                ok = read_records(f, ny, fs, [&](uint32_t i, const uint8_t* r) { syn_from_v7(sy[i], rec_as<SynV7>(r)); });
            } else if (fs == kSynV13) {
                ok = read_records(f, ny, fs, [&](uint32_t i, const uint8_t* r) { syn_from_v13(sy[i], rec_as<SynV13>(r)); });
            } else {
                ok = read_records(f, ny, fs, [&](uint32_t i, const uint8_t* r) {
                    memset(&sy[i], 0, sizeof(Synapse));
                    memcpy(&sy[i], r, fs < sizeof(Synapse) ? fs : sizeof(Synapse));
                });
            }
        }
        if (!ok) return fail(BION_ERR_FILE_READ);
    }

    // ---- nuclei: stored from v5, else one default per neuron ----
    if (v >= 5 && h.num_nuclei > 0) {
        const uint32_t nn = h.num_nuclei;
        b->nuclei = new (std::nothrow) Nucleus[nn]();
        if (b->nuclei) {
            b->num_nuclei = nn;
            Nucleus* nu = b->nuclei;
            if (v == 5) {
                ok = read_records(f, nn, sizeof(NucV5), [&](uint32_t i, const uint8_t* r) { nuc_from(nu[i], rec_as<NucV5>(r), 20); });
            } else if (v == 6) {
                ok = read_records(f, nn, kNucV6, [&](uint32_t i, const uint8_t* r) {
                    memset(&nu[i], 0, sizeof(Nucleus));
                    memcpy(&nu[i], r, kNucV6);
                    for (int g = 36; g < 46; g++) nu[i].gene_activity[g] = nu[i].protein_level[g] = 0.0f;
                });
            } else if (h.file_sz_nucleus == kNucV46) {
                ok = read_records(f, nn, kNucV46, [&](uint32_t i, const uint8_t* r) { nuc_from(nu[i], rec_as<NucV46>(r), 46); });
            } else {
                ok = fread(nu, sizeof(Nucleus), nn, f) == nn;
            }
            if (!ok) return fail(BION_ERR_FILE_READ);
        }
    } else {
        b->nuclei = new (std::nothrow) Nucleus[b->num_neurons ? b->num_neurons : 1]();
        if (b->nuclei) {
            b->num_nuclei = b->num_neurons;
            for (uint32_t i = 0; i < b->num_neurons; i++) {
                Nucleus& n = b->nuclei[i];
                n.neuron_id = b->neurons[i].id;
                n.enabled = 1;
                n.consolidation_mode = 3;
                n.amino_acid_level = 0.5f;
            }
        }
    }
    fclose(f);

    // ---- validation and the per-segment connection lists ----
    for (uint32_t i = 0; i < b->num_connections; i++)
        if (b->connections[i].segment_a >= ns || b->connections[i].segment_b >= ns) {
            free_bion(b); return BION_ERR_BAD_CONN;
        }
    for (uint32_t i = 0; i < b->num_synapses; i++) {
        const Synapse& y = b->synapses[i];
        if (y.pre_segment >= ns || y.post_segment >= ns) { free_bion(b); return BION_ERR_BAD_SYNAPSE; }
        if (y.pre_autoreceptor_count > BC_MAX_PRE_AUTORECEPTORS || y.post_receptor_count > BC_MAX_POST_RECEPTORS) {
            free_bion(b); return BION_ERR_BAD_MODCOUNT;
        }
        if (y.type >= 3) { free_bion(b); return BION_ERR_BAD_SYNAPSE; }
    }
    for (uint32_t i = 0; i < b->num_neurons; i++)
        if ((uint64_t)b->neurons[i].segment_start + b->neurons[i].segment_count > ns) {
            free_bion(b); return BION_ERR_BAD_NEURON;
        }
    for (uint32_t i = 0; i < ns; i++)
        if (sg[i].modifier_count > BC_MAX_MODIFIERS) { free_bion(b); return BION_ERR_BAD_MODCOUNT; }
    for (uint32_t i = 0; i < ns; i++) sg[i].intra_connection_count = 0;
    for (uint32_t i = 0; i < b->num_connections; i++) {
        const uint32_t e[2] = {b->connections[i].segment_a, b->connections[i].segment_b};
        for (uint32_t s : e)
            if (sg[s].intra_connection_count < BC_MAX_INTRACONN_PER_SEG)
                sg[s].intra_connection_ids[sg[s].intra_connection_count++] = i;
    }
    return 0;
}

// bion_save (src/bion_loader_host.cpp:449-504): v14 records; the header's mode, diffusion type
// and grid are cleared, voxel size 1, the timestamp is the time of the call, build id "v5.0.0".
static int write_bion(const char* path, const BionData* b) {
    FILE* f = fopen(path, "wb");
    if (!f) return BION_ERR_FILE_OPEN;
    NetworkHeader h = b->header;
    h.magic = BION_MAGIC_U32;
    h.version = BION_VERSION_CURRENT;
    h.mode = 0;
    h.diffusion_type = 0;
    h.num_neurons = b->num_neurons;
    h.num_segments = b->num_segments;
    h.num_intra_connections = b->num_connections;
    h.num_synapses = b->num_synapses;
    h.grid_dim_x = h.grid_dim_y = h.grid_dim_z = 0;
    h.voxel_size = 1.0f;
    h.timestamp = (uint64_t)time(nullptr);
    strncpy(h.build_id, "v5.0.0", sizeof(h.build_id) - 1);
    h.num_nuclei = b->num_nuclei;
    h.nucleus_version = 1;
    h.file_sz_segment = sizeof(Segment);
    h.file_sz_synapse = sizeof(Synapse);
    h.file_sz_nucleus = sizeof(Nucleus);
    h.file_sz_intraconnection = sizeof(IntraConnection);
    h.file_sz_neuron = sizeof(Neuron);
    h.file_sz_modifier = sizeof(Modifier);
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1;
    auto wr = [&](const void* p, size_t sz, size_t n) { if (ok && n) ok = fwrite(p, sz, n, f) == n; };
    wr(b->neurons, sizeof(Neuron), b->num_neurons);
    if (ok) ok = fwrite(b->segments, sizeof(Segment), b->num_segments, f) == b->num_segments;
    wr(b->connections, sizeof(IntraConnection), b->num_connections);
    wr(b->synapses, sizeof(Synapse), b->num_synapses);
    if (b->nuclei) wr(b->nuclei, sizeof(Nucleus), b->num_nuclei);
    fclose(f);
    return ok ? 0 : BION_ERR_WRITE_FAIL;
}

}  // namespace bc
