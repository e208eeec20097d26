// biocore_abi.h — binary layouts shared with BioNeuronCore callers (BION v14 records,
// stimulus structs, device info, BionData). Field order, widths and packing are fixed by
// the reference's on-disk format and its dlopen ABI (include/types.h:446-710,
// include/bion_loader.h:16-28); every size and the offsets callers rely on are asserted below.
// These are plain C structs: no behaviour lives here.
#pragma once
#include <stddef.h>
#include <stdint.h>

#define BC_ION_COUNT 7
#define BC_ION_MAX 8
#define BC_NT_COUNT 15
#define BC_MAX_MODIFIERS 32
#define BC_MAX_STATE 8
#define BC_MAX_PRE_AUTORECEPTORS 8
#define BC_MAX_POST_RECEPTORS 8
#define BC_MAX_INTRACONN_PER_SEG 16
#define BC_GPU_MAX_STIMULI 256

#define BC_ION_NA 0
#define BC_ION_K 1
#define BC_ION_CA 2
#define BC_ION_CL 3
#define BC_ION_MG 4
#define BC_ION_A 5
#define BC_ION_H 6

#define BC_NT_GLUTAMATE 0
#define BC_NT_GABA 1
#define BC_NT_DOPAMINE 2
#define BC_NT_SEROTONIN 3
#define BC_NT_ACETYLCHOLINE 4
#define BC_NT_GLYCINE 7
#define BC_NT_ENDOCANNABINOID 9

#define BION_MAGIC_U32 0x42494F4Eu
#define BION_VERSION_CURRENT 14u

// Loader error codes (reference src/bion_loader_host.cpp:14-24).
#define BION_ERR_FILE_OPEN -1
#define BION_ERR_BAD_MAGIC -2
#define BION_ERR_BAD_VERSION -3
#define BION_ERR_ZERO_SEGS -4
#define BION_ERR_ALLOC -5
#define BION_ERR_BAD_CONN -6
#define BION_ERR_BAD_SYNAPSE -7
#define BION_ERR_BAD_NEURON -8
#define BION_ERR_BAD_MODCOUNT -9
#define BION_ERR_WRITE_FAIL -10
#define BION_ERR_FILE_READ -11

#pragma pack(push, 1)

typedef struct { float x, y, z; } Vec3;

typedef struct {
    uint8_t type;
    uint8_t _pad[7];
    double density;
    double state[BC_MAX_STATE];
    uint8_t active_vars;
    uint8_t flags;
    uint8_t _reserved[6];
} Modifier;

typedef struct {
    uint32_t id;
    uint32_t neuron_id;
    Vec3 position;
    Vec3 direction;
    float length;
    float radius;
    double V_m;
    double intra_ions[BC_ION_MAX];
    double extra_ions[BC_ION_MAX];
    float intra_nt[BC_NT_COUNT];
    float neurotransmitter[BC_NT_COUNT];
    Modifier modifiers[BC_MAX_MODIFIERS];
    uint8_t modifier_count;
    uint8_t segment_type;
    uint8_t reserved1[6];
    float ATP_level;
    float temperature;
    double I_buffer;
    float transcription_signal;
    uint32_t intra_connection_ids[BC_MAX_INTRACONN_PER_SEG];
    uint8_t intra_connection_count;
    uint8_t reserved2[7];
    float ER_Ca2;
} Segment;

typedef struct {
    uint32_t id;
    uint32_t segment_start;
    uint32_t segment_count;
    uint32_t soma_index;
    float threshold;
    uint8_t type;
    uint8_t reserved[3];
} Neuron;

#define BC_GENE_COUNT 57
typedef struct {
    uint32_t neuron_id;
    uint8_t enabled;
    uint8_t consolidation_mode;
    uint8_t _pad0[2];
    uint32_t consolidation_interval_ticks;
    uint32_t last_consolidation_tick;
    float gene_activity[BC_GENE_COUNT];
    float protein_level[BC_GENE_COUNT];
    float protein_synthesis_rate;
    float pcreb_level;
    float erk_level;
    float nfkb_level;
    float mtorc1_level;
    float amino_acid_level;
    uint8_t _reserved[4];
} Nucleus;

typedef struct {
    uint32_t segment_a;
    uint32_t segment_b;
    uint8_t is_intracellular;
    uint8_t reserved[3];
    uint32_t _pad;
} IntraConnection;

typedef struct {
    uint32_t id;
    uint32_t pre_segment;
    uint32_t post_segment;
    uint8_t type;
    uint8_t reserved[3];
    float weight;
    float delay;
    float pre_area;
    float post_area;
    Modifier pre_autoreceptors[BC_MAX_PRE_AUTORECEPTORS];
    uint8_t pre_autoreceptor_count;
    uint8_t reserved2[7];
    Modifier post_receptors[BC_MAX_POST_RECEPTORS];
    uint8_t post_receptor_count;
    uint8_t reserved3[3];
    float release_probability;
    float release_amount;
    float cleft_neurotransmitter[BC_NT_COUNT];
    float facilitation;
    float depression;
    float tau_clearance;
    uint8_t refractory;
    uint8_t reserved4[7];
    float RRP;
    float RRP_max;
    float RRP_recovery;
    float bouton_Ca;
    float spine_Ca;
    float spine_camkii;
    float spine_calcineurin;
} Synapse;

typedef struct {
    uint32_t magic;
    uint32_t version;
    uint8_t mode;
    uint8_t diffusion_type;
    uint16_t reserved;
    uint32_t num_neurons;
    uint32_t num_segments;
    uint32_t num_intra_connections;
    uint32_t num_synapses;
    uint32_t grid_dim_x;
    uint32_t grid_dim_y;
    uint32_t grid_dim_z;
    float voxel_size;
    uint64_t timestamp;
    char build_id[16];
    uint32_t num_nuclei;
    uint32_t nucleus_version;
    uint32_t file_sz_segment;
    uint32_t file_sz_synapse;
    uint32_t file_sz_nucleus;
    uint32_t file_sz_intraconnection;
    uint32_t file_sz_neuron;
    uint32_t file_sz_modifier;
} NetworkHeader;

typedef struct {
    uint32_t segment_id;
    float amplitude_pA;
    float duration_ms;
    uint64_t start_tick;
} CurrentStimulus;

typedef struct {
    uint32_t segment_id;
    float target_mV;
    uint64_t start_tick;
    uint64_t end_tick;
} VoltageClamp;

typedef struct {
    uint32_t segment_id;
    float start_pA;
    float end_pA;
    uint64_t start_tick;
    uint64_t end_tick;
} RampStimulus;

typedef struct {
    uint32_t segment_id;
    int32_t ion_idx;
    float amount_mM;
    uint8_t extracellular;
    uint8_t _pad[3];
} IonInjection;

typedef struct {
    uint32_t segment_id;
    int32_t nt_idx;
    float amount;
    uint8_t intracellular;
    uint8_t _pad2[3];
} LigandInjection;

typedef struct {
    uint32_t current_stim_count;
    uint32_t vclamp_count;
    uint32_t ramp_count;
    uint32_t ion_inj_count;
    uint32_t lig_inj_count;
    uint64_t current_tick;
    CurrentStimulus current_stim[BC_GPU_MAX_STIMULI];
    VoltageClamp vclamps[BC_GPU_MAX_STIMULI];
    RampStimulus ramps[BC_GPU_MAX_STIMULI];
    IonInjection ion_inj[BC_GPU_MAX_STIMULI];
    LigandInjection lig_inj[BC_GPU_MAX_STIMULI];
} GpuStimulusData;

typedef struct GpuDeviceInfo {
    char name[128];
    uint32_t vendor_id;
    uint8_t is_cpu;
    uint64_t global_mem_bytes;
    uint32_t max_compute_units;
    int32_t device_index;
    char backend_type[16];
} GpuDeviceInfo;

#pragma pack(pop)

// Host view handed out by gpu_backend_bion_data (natural alignment, not packed).
typedef struct BionData {
    NetworkHeader header;
    Neuron* neurons;
    Segment* segments;
    IntraConnection* connections;
    Synapse* synapses;
    Nucleus* nuclei;
    uint32_t num_neurons;
    uint32_t num_segments;
    uint32_t num_connections;
    uint32_t num_synapses;
    uint32_t num_nuclei;
} BionData;

#ifdef __cplusplus
static_assert(sizeof(Modifier) == 88, "Modifier layout");
static_assert(sizeof(Segment) == 3216, "Segment layout");
static_assert(sizeof(Synapse) == 1568, "Synapse layout");
static_assert(sizeof(IntraConnection) == 16, "IntraConnection layout");
static_assert(sizeof(Neuron) == 24, "Neuron layout");
static_assert(sizeof(Nucleus) == 500, "Nucleus layout");
static_assert(sizeof(NetworkHeader) == 100, "NetworkHeader layout");
static_assert(sizeof(CurrentStimulus) == 20, "CurrentStimulus layout");
static_assert(sizeof(VoltageClamp) == 24, "VoltageClamp layout");
static_assert(sizeof(RampStimulus) == 28, "RampStimulus layout");
static_assert(sizeof(IonInjection) == 16, "IonInjection layout");
static_assert(sizeof(LigandInjection) == 16, "LigandInjection layout");
static_assert(sizeof(GpuDeviceInfo) == 165, "GpuDeviceInfo layout");
static_assert(offsetof(Segment, V_m) == 40, "V_m offset");
static_assert(offsetof(Segment, intra_ions) == 48, "intra_ions offset");
static_assert(offsetof(Segment, extra_ions) == 112, "extra_ions offset");
static_assert(offsetof(Segment, modifier_count) == 3112, "modifier_count offset");
static_assert(offsetof(Segment, I_buffer) == 3128, "I_buffer offset");
static_assert(offsetof(Synapse, post_receptors) == 744, "post_receptors offset");
static_assert(offsetof(BionData, neurons) == 104, "BionData layout");
static_assert(sizeof(BionData) == 168, "BionData layout");
#endif
