// biocore_hip_backend.h — C ABI of libbiocore_hip_backend.so, the MI355X (gfx950) backend for
// BioNeuronCore's per-tick simulation step.
//
// The first block is the drop-in contract: the same eleven symbols, argument meanings and error
// behaviour as BioNeuronCore's GPU backends (src/gpu_backend_cuda.cu:316-390,
// src/gpu_backend.cpp:40-144), resolved by callers with dlopen/dlsym
// (src/daemon_server.cpp:198-206, tests/test_cpu_gpu_parity_cuda.cpp:45-70).
// The second block is the plugin-registration surface plugins bind to
// (src/plugin_manager.cpp:37-75, include/modifier_interface.h:38-39).
// The third block ("bchip_*") is new: it has no counterpart in the reference.
#pragma once
#include <stdint.h>
#include "biocore_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef double (*ModifierFunc)(Modifier* mod, Segment* seg, float dt);
typedef float (*ReleaseModulatorFunc)(Modifier* mod, float* cleft_nt, float dt);

// ---------------------------------------------------------------- drop-in (reference ABI)
// gpu_backend_create -> create_on_device(0)                     (gpu_backend_cuda.cu:330-332)
void* gpu_backend_create(void);
// Out-of-range index selects device 0; NULL when no device      (gpu_backend_cuda.cu:318-328)
void* gpu_backend_create_on_device(int device_index);
// NULL-safe                                                      (gpu_backend_cuda.cu:334-336)
void gpu_backend_destroy(void* sim);
// 0 ok; -1 NULL args; BION_ERR_* loader codes (bion_loader_host.cpp:14-24); versions 3-14
// (older records migrated as bion_loader_host.cpp:99-340 does);
// islands of any size (one beyond a 256-segment tile runs in a workgroup of its own);
// BCHIP_ERR_UNSUPPORTED when the network needs work the device does not do (see below)
//                                                                (gpu_backend_cuda.cu:338-341)
int gpu_backend_load_bion(void* sim, const char* path);
// One tick (steps 1, 2.5, 3, 4); synchronous on return; NULL-safe (gpu_backend_cuda.cu:343-346)
void gpu_backend_run_tick(void* sim, float dt);
// Host-readable BionData (AoS), owned by sim; current as of the last tick (gpu_backend_cuda.cu:348-351)
void* gpu_backend_bion_data(void* sim);
// Deep copy; persists until the next upload; ion/ligand injections are one-shot
//                                                                (gpu_backend_cuda.cu:353-356)
void gpu_backend_upload_stimuli(void* sim, const GpuStimulusData* src);
int gpu_backend_device_count(void);                             // (gpu_backend_cuda.cu:358-362)
// 0 ok / -1 bad index or NULL info                               (gpu_backend_cuda.cu:364-382)
int gpu_backend_get_device_info(int index, GpuDeviceInfo* info);
int gpu_backend_ibuf_offset(void);   // 3128                      (gpu_backend_cuda.cu:386)
int gpu_backend_vm_offset(void);     // 40                        (gpu_backend_cuda.cu:387)
int gpu_backend_seg_size(void);      // 3216                      (gpu_backend_cuda.cu:388)

// ---------------------------------------------------------------- plugin registration
// IDs < 128; recorded for host-side lookups. Device ticks never call host function pointers
// (as in the reference's device dispatch, kernel_dispatch.h:340-344).
void register_modifier_current(uint8_t type, ModifierFunc func);     // plugin_manager.cpp:37-41
void register_modulator_release(uint8_t type, ReleaseModulatorFunc func);   // plugin_manager.cpp:43-47

// ---------------------------------------------------------------- extensions (no reference counterpart)
#define BCHIP_COHERENCE_COMPAT 0   // AoS uploaded before and refreshed after every tick
#define BCHIP_COHERENCE_FAST 1     // AoS refreshed only by gpu_backend_bion_data / bchip_save_bion;
                                   // host edits reach the device only via bchip_push_host_state
#define BCHIP_ERR_ISLAND_TOO_LARGE -20   // no longer returned: islands of any size load
// A load (gpu_backend_load_bion, bchip_load_bion_data, bchip_load_batch) returns this when the
// network needs work the device does not do - modifier or receptor types outside the device set
// (BchipStats.unsupported_modifiers) - instead of running those parts as zero current; a warning
// names the counts on stderr and bchip_stats reports them. bchip_allow_unsupported(sim, 1) loads
// such networks anyway (those parts then do nothing on the device).
#define BCHIP_ERR_UNSUPPORTED -21
#define BCHIP_ERR_SCRATCH -22   // bchip_adjoint_noalloc: the call needs more than its fixed array
int bchip_allow_unsupported(void* sim, int on);
// Device evaluation, on the current device, of the exact-arithmetic helpers the tick kernels use
// (for checking them against host glibc and IEEE division): out[i] = fn(a[i], b[i]) with fn
// 0 exp (table staged in LDS), 1 exp (table read from global memory), 2 a / b in double through
// the reciprocal (div_rcp), 3 float(a) / float(b) through the double reciprocal (div_wide),
// 4 a / double(float(b)) (div_f24), 5 the Scharfetter-Gummel Bernoulli function of float(a)
// (b unused for 0, 1 and 5). Returns 0 or -1.
int bchip_eval_math(int fn, const double* a, const double* b, double* out, uint32_t n);

typedef struct BchipStats {
    uint32_t num_segments, num_connections, num_synapses, num_tiles;
    uint32_t modifier_slots, receptor_slots, lds_bytes_per_tile, _pad;
    uint64_t unsupported_modifiers;   // modifier/receptor instances outside the device set (run as 0 current)
    uint64_t unsupported_synapses;    // 0: every synapse kind runs on the device (field kept for layout)
    uint64_t order_hazard_synapses;   // synapses whose step-3 effect depends on another post chain's
                                      // progress (immediate releases behind a synapse onto their
                                      // presynaptic segment, gap junctions behind a Ca-changing synapse),
                                      // completed exactly by the ordered pass; before the first tick: those
                                      // that are hazards at any dt
    float last_phase_ms[6];           // step 1, 2.5, 3-pre, 3-post+4, 3-plasticity, 3-rrp of the last
                                      // timed tick
    uint32_t ordered_events;          // synapse events the ordered pass walks (last planned dt)
    uint32_t ordered_segments;        // segments it walks (the parallel post pass skips them)
    uint32_t big_islands;             // islands beyond one tile (one workgroup each in step 2.5)
    uint32_t _pad2;
} BchipStats;

int bchip_load_bion_data(void* sim, const BionData* src);   // load from in-memory AoS (copied)
int bchip_save_bion(void* sim, const char* path);            // BION v14 save (bion_save layout)
// Host-only BION file access (no device involved): bion_load_host (src/bion_loader_host.cpp:26),
// reading versions 3-14 with the reference's record migrations; bion_unload_host (:428);
// bion_save (:449, v14). Same error codes.
int bchip_read_bion(const char* path, BionData* out);
void bchip_free_bion(BionData* data);
int bchip_write_bion(const char* path, const BionData* data);
int bchip_set_coherence(void* sim, int mode);
int bchip_push_host_state(void* sim);                        // upload AoS edits (FAST mode)
int bchip_run_ticks(void* sim, float dt, uint32_t count);
uint64_t bchip_tick_count(void* sim);
int bchip_spike_counts(void* sim, uint32_t* out, uint32_t n);   // step-4 upward crossings of -40 mV
int bchip_delay_slots(void* sim, uint8_t* out, uint64_t cap);   // 4 x DelayedRelease per synapse
int bchip_stats(void* sim, BchipStats* out);
int bchip_set_phase_timing(void* sim, int on);
void* bchip_stream(void* sim);                                 // the sim's hipStream_t
int bchip_plugin_load(const char* path);                       // dlopen + biocore_plugin_init
int bchip_registered_modifier(uint8_t type);
int bchip_modifier_supported(uint8_t type);

// ---- batches: independent networks advanced together by one tick launch sequence.
// The networks are copied into one device state (a disjoint union; every record keeps its own
// network's values, including Segment.id and Synapse.id); network k's results equal a
// separate load of network k bit for bit. A plain load is a batch of one.
int bchip_load_batch(void* sim, const BionData* const* nets, uint32_t count);
uint32_t bchip_network_count(void* sim);
// Stimuli of network `net` (arrays copied, no size limit, Segment.id matched within that
// network), replacing its previous ones; injections apply on the next tick only, as the CPU
// tick's one-shot injections. gpu_backend_upload_stimuli sets those of network 0.
int bchip_set_stimuli(void* sim, uint32_t net, const CurrentStimulus* cur, uint32_t ncur,
                      const RampStimulus* ramp, uint32_t nramp, const VoltageClamp* vc, uint32_t nvc,
                      const IonInjection* ion, uint32_t nion, const LigandInjection* lig, uint32_t nlig);
// Copies network `net`'s current records into caller arrays sized for that network (its own
// segment indices in Synapse pre/post), and its spike counts (may be NULL).
int bchip_batch_read(void* sim, uint32_t net, Segment* segs, Synapse* syns, uint32_t* spikes);

// ---- launch schedule: workgroup shapes of the tick's kernels. Results do not depend on it.
typedef struct BchipSchedule {
    uint32_t modifiers_block;   // step-1 workgroup size: 64, 128 or 256
    uint32_t diffusion_tiles;   // 64-segment tiles per step-2.5 workgroup: 1, 2 or 4
    uint32_t pre_grid;          // workgroups of the grid-stride pre-side pass
    uint32_t plast_block;       // plasticity-pass workgroup size: 64, 128 or 256
    uint32_t rrp_block;         // vesicle-pool-pass workgroup size: 64, 128 or 256
    uint32_t step1_chunks;      // 1-16: the 64-segment tiles in this many chunks, step 2.5 of a
                                // chunk on a second stream overlapping step 1 of the next (1: none)
} BchipSchedule;
int bchip_get_schedule(void* sim, BchipSchedule* out);
int bchip_set_schedule(void* sim, const BchipSchedule* in);   // -1 if a field is out of range
// Times each candidate of each schedule field on ticks_per_candidate live ticks (coordinate
// descent; the ticks advance the simulation) and keeps the fastest. Returns the ticks run.
int bchip_autotune(void* sim, float dt, uint32_t ticks_per_candidate);
// bchip_run_ticks replays a captured hipGraph of up to 16 ticks in FAST coherence (on by default).
int bchip_set_graphs(void* sim, int on);
// Residual-terminated Jacobi for both potential solves of step 2.5 (an extension; the reference
// always sweeps 50 times): each connected cell stops after the first sweep whose largest update
// is <= tol (mV), or after max_sweeps. tol <= 0 restores the reference's fixed 50 sweeps.
int bchip_set_jacobi(void* sim, double tol, int max_sweeps);
// Cable mode (an extension; the tick bchip_adjoint differentiates): ion concentrations and ATP
// are held at their values and phi_e = 0; the axial current is the intracellular SG flux and
// chord term at those concentrations; ion injections are ignored, ligand injections apply.
// Networks with synapses are refused (-1), and loading one ends the mode. The mode and the
// scales below stay set across ticks; a load resets the scales to 1.
int bchip_set_cable_mode(void* sim, int on);
// Per-connection scale (connection order, n = num_connections) of the axial conductance and the
// intracellular ion fluxes; NULL restores 1 (the reference). Scale 1 changes no bit.
int bchip_set_coupling_scale(void* sim, const float* scale, uint32_t n);
// Gradient of the cable-mode tick run for `ticks` ticks from the current state, for the loss
// mean over ticks and cells of (V_m of host segment soma[c] after the tick - target[tick][c])^2
// (target: ticks x ncells). State is checkpointed every checkpoint_every ticks and each span
// is recomputed once for the reverse pass. grad_density: num_segments x 32 (d loss / d density
// of modifier slot j of segment h); grad_scale: num_connections (d loss / d coupling scale).
// With both NULL only the forward pass and the loss run. The simulation ends `ticks` ticks
// later, as after bchip_run_ticks. trace (ticks x ncells, may be NULL) receives the somatic V_m
// after each tick. The reverse pass reads the Jacobi iterates the forward recorded for the span
// being reversed (5 x 51 floats per segment and tick) when checkpoint_every ticks of them fit in
// half the free device memory, and otherwise runs each tick's substeps again. Returns 0, -1
// (arguments or state: cable mode off, synapses present, residual-terminated Jacobi on), -2 (a
// modifier type outside Nav1.2/Kv-DR/K2P/NALCN/HCN1, more than 8 modifier slots, or an island
// beyond one 256-segment tile).
int bchip_adjoint(void* sim, float dt, uint32_t ticks, uint32_t checkpoint_every, const uint32_t* soma,
                  uint32_t ncells, const float* target, double* loss, double* grad_density, double* grad_scale,
                  float* trace);
// No counterpart in the reference: for callers that cannot allocate device memory (code under
// HIP graph capture, serving and training loops, harnesses that forbid hipMalloc). Same as
// bchip_adjoint, with every buffer of the call (checkpoints, recordings, adjoint state, staging)
// taken from a fixed 64 MiB device array the library holds per GPU instead of allocated; it
// calls no hipMalloc/hipFree and allocates no host memory. The recorded iterates are used when
// they fit in what the call leaves of the array. Requires FAST coherence and no stimulus upload
// pending since the last tick (building the stimulus lists allocates): -1 otherwise.
// BCHIP_ERR_SCRATCH (nothing run, state unchanged) when the call needs more than the array
// holds: checkpoints ((ticks / checkpoint_every + 2) x (8 + 56 x slots) bytes per segment), the
// targets (4 x ticks x cells bytes), the span record (16 x (checkpoint_every + 1) x (1 + slots)
// bytes per segment) and the gradients (256 bytes per segment, 8 per connection). Calls that
// overlap on one GPU share the array and race; each call clears what it uses.
int bchip_adjoint_noalloc(void* sim, float dt, uint32_t ticks, uint32_t checkpoint_every, const uint32_t* soma,
                          uint32_t ncells, const float* target, double* loss, double* grad_density,
                          double* grad_scale, float* trace);
// Tick at which stimulus time windows (current pulses, ramps, voltage clamps) are matched.
// 0 (default): the tick being run, as the CPU tick (src/simulation_host.cpp:186,247-300) and the
// CUDA backend (src/gpu_backend_cuda.cu:252-255) match them. 1: that tick plus one, as the SYCL
// backend does (src/simulation.cpp:201,234), for callers that were built against it. One-shot
// ion and ligand injections apply on the next tick either way.
int bchip_set_stimulus_convention(void* sim, int convention);
// Protein -> release-parameter mapping (on != 0): every chemical synapse whose presynaptic
// neuron has a nucleus in the loaded BionData takes RRP_max = 10 + 40 Munc13-1, release
// probability = 0.3 + 0.7 synaptotagmin-1 and RRP_recovery = 0.02 + 0.08 Munc13-1 - 0.01
// complexin (clamped to [0.005, 0.2]) from that nucleus's protein levels, as the CPU tick does
// while its consolidation runs (src/simulation_host.cpp:407-414, include/kernel_dispatch.h:
// 651-663). Nucleus state is not advanced on the device. Takes effect from the next tick; the
// records show the mapped values once a tick has run. Returns 0, or -1 (no simulation).
int bchip_set_protein_mapping(void* sim, int on);
// Upper bound on the device memory bchip_adjoint may use for recorded Jacobi iterates (bytes;
// default unbounded, i.e. half the free memory). 0 makes the reverse pass run the substeps again.
int bchip_set_adjoint_memory(void* sim, uint64_t max_bytes);

#ifdef __cplusplus
}
// C++ registration entry points (plugin_manager.cpp:49-75).
void register_user_modifier(uint8_t type, ModifierFunc func);
void register_user_modulator(uint8_t type, ReleaseModulatorFunc func);
#endif
