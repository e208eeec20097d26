"""BION v14 records as numpy structured dtypes, file read/write, and the synthetic networks
the hot path is measured on.

The record layouts are the fixed BioNeuronCore ABI (reference include/types.h:446-603,
sizes asserted at :703-710); the file layout is header, neurons, segments, connections,
synapses, nuclei (reference src/bion_loader_host.cpp:26-107, bion_save :449-504).
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

MODIFIER_DT = np.dtype([
    ("type", "u1"), ("_pad", "u1", 7), ("density", "<f8"), ("state", "<f8", 8),
    ("active_vars", "u1"), ("flags", "u1"), ("_reserved", "u1", 6)])
SEGMENT_DT = np.dtype([
    ("id", "<u4"), ("neuron_id", "<u4"), ("position", "<f4", 3), ("direction", "<f4", 3),
    ("length", "<f4"), ("radius", "<f4"), ("V_m", "<f8"), ("intra_ions", "<f8", 8),
    ("extra_ions", "<f8", 8), ("intra_nt", "<f4", 15), ("neurotransmitter", "<f4", 15),
    ("modifiers", MODIFIER_DT, 32), ("modifier_count", "u1"), ("segment_type", "u1"),
    ("reserved1", "u1", 6), ("ATP_level", "<f4"), ("temperature", "<f4"), ("I_buffer", "<f8"),
    ("transcription_signal", "<f4"), ("intra_connection_ids", "<u4", 16),
    ("intra_connection_count", "u1"), ("reserved2", "u1", 7), ("ER_Ca2", "<f4")])
NEURON_DT = np.dtype([
    ("id", "<u4"), ("segment_start", "<u4"), ("segment_count", "<u4"), ("soma_index", "<u4"),
    ("threshold", "<f4"), ("type", "u1"), ("reserved", "u1", 3)])
CONN_DT = np.dtype([
    ("segment_a", "<u4"), ("segment_b", "<u4"), ("is_intracellular", "u1"), ("reserved", "u1", 3),
    ("_pad", "<u4")])
SYNAPSE_DT = np.dtype([
    ("id", "<u4"), ("pre_segment", "<u4"), ("post_segment", "<u4"), ("type", "u1"),
    ("reserved", "u1", 3), ("weight", "<f4"), ("delay", "<f4"), ("pre_area", "<f4"),
    ("post_area", "<f4"), ("pre_autoreceptors", MODIFIER_DT, 8), ("pre_autoreceptor_count", "u1"),
    ("reserved2", "u1", 7), ("post_receptors", MODIFIER_DT, 8), ("post_receptor_count", "u1"),
    ("reserved3", "u1", 3), ("release_probability", "<f4"), ("release_amount", "<f4"),
    ("cleft_neurotransmitter", "<f4", 15), ("facilitation", "<f4"), ("depression", "<f4"),
    ("tau_clearance", "<f4"), ("refractory", "u1"), ("reserved4", "u1", 7), ("RRP", "<f4"),
    ("RRP_max", "<f4"), ("RRP_recovery", "<f4"), ("bouton_Ca", "<f4"), ("spine_Ca", "<f4"),
    ("spine_camkii", "<f4"), ("spine_calcineurin", "<f4")])
HEADER_DT = np.dtype([
    ("magic", "<u4"), ("version", "<u4"), ("mode", "u1"), ("diffusion_type", "u1"),
    ("reserved", "<u2"), ("num_neurons", "<u4"), ("num_segments", "<u4"),
    ("num_intra_connections", "<u4"), ("num_synapses", "<u4"), ("grid_dim_x", "<u4"),
    ("grid_dim_y", "<u4"), ("grid_dim_z", "<u4"), ("voxel_size", "<f4"), ("timestamp", "<u8"),
    ("build_id", "S16"), ("num_nuclei", "<u4"), ("nucleus_version", "<u4"),
    ("file_sz_segment", "<u4"), ("file_sz_synapse", "<u4"), ("file_sz_nucleus", "<u4"),
    ("file_sz_intraconnection", "<u4"), ("file_sz_neuron", "<u4"), ("file_sz_modifier", "<u4")])
NUCLEUS_SIZE = 500
# Nucleus record (reference include/types.h: gene activity and protein levels of 57 genes).
NUCLEUS_DT = np.dtype([("neuron_id", "<u4"), ("enabled", "u1"), ("consolidation_mode", "u1"), ("pad", "u1", 2),
                       ("consolidation_interval_ticks", "<u4"), ("last_consolidation_tick", "<u4"),
                       ("gene_activity", "<f4", 57), ("protein_level", "<f4", 57),
                       ("protein_synthesis_rate", "<f4"), ("pcreb_level", "<f4"), ("erk_level", "<f4"),
                       ("nfkb_level", "<f4"), ("mtorc1_level", "<f4"), ("amino_acid_level", "<f4"),
                       ("res", "u1", 4)])
assert NUCLEUS_DT.itemsize == NUCLEUS_SIZE
CURRENT_STIM_DT = np.dtype([("segment_id", "<u4"), ("amplitude_pA", "<f4"), ("duration_ms", "<f4"),
                            ("start_tick", "<u8")])
VCLAMP_DT = np.dtype([("segment_id", "<u4"), ("target_mV", "<f4"), ("start_tick", "<u8"),
                      ("end_tick", "<u8")])
RAMP_DT = np.dtype([("segment_id", "<u4"), ("start_pA", "<f4"), ("end_pA", "<f4"),
                    ("start_tick", "<u8"), ("end_tick", "<u8")])
ION_INJ_DT = np.dtype([("segment_id", "<u4"), ("ion_idx", "<i4"), ("amount_mM", "<f4"),
                       ("extracellular", "u1"), ("_pad", "u1", 3)])
LIG_INJ_DT = np.dtype([("segment_id", "<u4"), ("nt_idx", "<i4"), ("amount", "<f4"),
                       ("intracellular", "u1"), ("_pad2", "u1", 3)])

assert MODIFIER_DT.itemsize == 88 and SEGMENT_DT.itemsize == 3216
assert SYNAPSE_DT.itemsize == 1568 and CONN_DT.itemsize == 16 and NEURON_DT.itemsize == 24
assert HEADER_DT.itemsize == 100

BION_MAGIC = 0x42494F4E
BION_VERSION = 14

# Modifier IDs (reference include/types.h:22-263) used by the hot-path workloads.
PUMP, AMPA, NMDA, GABAA = 7, 8, 9, 11
SK, BK, K2P, KV7, KVDR, HCN1, NALCN = 103, 104, 107, 112, 123, 129, 133
CAV21, NAV16, NAV12, KV31, KV11 = 140, 141, 142, 150, 156
CX36, CX43 = 114, 198   # connexins (gap junctions read their connexon count and open probability)
# Na+-coupled transmitter transporters: glial (GLAST marks a segment as glia for step 3's
# capture) and neuronal.
GLAST, GLYT1, GAT3, EAAT3, GLYT2, GAT1 = 147, 148, 149, 183, 184, 185
TRANSPORTERS = (GLAST, GLYT1, GAT3, EAAT3, GLYT2, GAT1)
# Ion homeostasis: pump isoforms, Ca2+ extrusion and buffering, cation-chloride cotransport.
HCN2, PUMP_A3, PUMP_GLIA, PMCA, NCX, KCC2, NKCC1, CA_BUFFER = 130, 197, 110, 98, 99, 100, 122, 102
HOMEOSTASIS = (HCN2, PUMP_A3, PUMP_GLIA, PMCA, NCX, KCC2, NKCC1, CA_BUFFER)
# Further voltage-gated and leak channels of the tadpole spinal network's cell types.
CAV13, CLC2, KV42, KV21, NAP, NAP_TAD, KIR, NAV12DEV = 1, 2, 85, 196, 105, 195, 106, 193
CHANNELS2 = (CAV13, CLC2, KV42, KV21, NAP, NAP_TAD, KIR, NAV12DEV)
# Receptors: further ionotropic types, and the Gi/o- (17-40) and Gs-coupled (41-51) families.
KAINATE, GLYCINE, NICOTINIC, HT3, ALPHA7, NMDA2B, GABAB = 10, 12, 13, 14, 97, 194, 17
RECEPTORS2 = (KAINATE, GLYCINE, NICOTINIC, HT3, ALPHA7, NMDA2B) + tuple(range(17, 52))
# Gq-coupled receptors (52-62, 70), P2X (15) and P2X2 (134), the network-builder receptor
# variants (63, 64, 67, 68, 71-74), HCN3 (131) and HCN4 (132).
P2X, P2X2, HCN3, HCN4 = 15, 134, 131, 132
RECEPTORS3 = tuple(range(52, 63)) + (70, P2X, P2X2, 63, 64, 67, 68, 71, 72, 73, 74, HCN3, HCN4)
NAV11, NAV13, NAV17, NAV18, NAV19 = 199, 177, 169, 170, 171
NAV_VARIANTS = (NAV11, NAV13, NAV17, NAV18, NAV19)
KV12, KV13, KV14, KV16, KV22 = 157, 154, 206, 207, 205
KV_VARIANTS = (KV12, KV13, KV14, KV16, KV22)
K_MORE = (203, 204, 202, 160, 161, 166, 167, 208, 178, 176)   # Kv3.2/3.3/4.3, GIRK, KATP, THIK-1/2, Kir2.2/7.1, KCNQ4
CAV_MORE = (168, 200, 201, 191)   # Cav2.2, Cav3.1, Cav3.2, Cav3.3
CATION_MORE = (210, 162, 211, 175, 182, 172, 173, 174, 179, 163, 164, 180)   # TRPC3/4/5/6-7, TRPM4/7/8,
# TRPV1/V3, TRPA1, ENaC, CALHM1/3
MISC_MORE = (159, 101, 96)   # Kv3.4, Ca2+ leak, post-inhibitory-rebound T-type Ca2+
CA_ACT = (84, 144, 145, 146, 152, 186, 190)   # L-type Cav1.2, Cav1.4, CaBP4, CaCC, CAN, AQP4, TMC1
MYELIN = 4   # no current; blocks the ionotropic receptors and the listed channel types
ECB_SYNTHESIS = 86   # 2-AG synthesis; step 3 deposits it into the clefts of incoming synapses
SERCA, IP3R, RYR, PLC = 126, 127, 128, 136   # the ER Ca2+ store and the Gq -> PLC -> IP3/DAG pool
ER_STORE = (SERCA, IP3R, RYR, PLC)
METABOTROPIC, RELEASE_SITE, SOMA_ATP, SOMA_THERMAL, CREB, MRGPR = 3, 5, 77, 78, 79, 188
SIGNALLING = (METABOTROPIC, RELEASE_SITE, SOMA_ATP, SOMA_THERMAL, CREB, MRGPR)
TRPV1_R, MAPK, NFKB, MTORC1, PROTEIN_SYNTH, COFILIN, MMP9, ROD = 16, 87, 88, 89, 90, 91, 92, 192
CASCADES = (TRPV1_R, MAPK, NFKB, MTORC1, PROTEIN_SYNTH, COFILIN, MMP9, ROD)
PHYS_SENSOR, HUNGER, LIGHT, OLFACTORY, THERMAL_NOISE = 6, 94, 95, 108, 111
VIBRATION, OTOLITH, CUPULA, CRYPTOCHROME = 115, 116, 117, 120
SENSORS = (PHYS_SENSOR, HUNGER, LIGHT, OLFACTORY, THERMAL_NOISE, VIBRATION, OTOLITH, CUPULA, CRYPTOCHROME)
CAMKII, CALCINEURIN, PP1, PKC = 82, 83, 113, 137   # the kinase/phosphatase activities others read
KINASES = (CAMKII, CALCINEURIN, PP1, PKC)
MSN_NMDA, MSN_AMPA = 75, 76   # striatal NMDA/AMPA scaled by the segment's D1/D2 modulator
# IDs without code in the reference: reserved, the plugin stubs 124/125 (no plugin loaded), and
# 212 (the type count) and above
RESERVED = (0, 80, 81, 93, 124, 125, 138, 139, 143)
NO_CODE = RESERVED + tuple(range(212, 256))
SUPPORTED_MODIFIERS = (MYELIN, PUMP, AMPA, NMDA, GABAA, SK, BK, K2P, KV7, KVDR, HCN1, NALCN, CAV21,
                       NAV16, NAV12, KV31, KV11, CX36, CX43, ECB_SYNTHESIS) + TRANSPORTERS + HOMEOSTASIS + CHANNELS2 + RECEPTORS2 + RECEPTORS3 + NAV_VARIANTS + KV_VARIANTS + K_MORE + CAV_MORE + CATION_MORE + MISC_MORE + CA_ACT + ER_STORE + SIGNALLING + CASCADES + SENSORS + KINASES + (MSN_NMDA, MSN_AMPA) + NO_CODE

# Physiological ion means (reference include/cpg_config.h:400-412).
INTRA_IONS = np.array([15.0, 140.0, 0.0001, 5.0, 0.5, 151.0, 6.3e-5, 0.0], dtype=np.float32)
EXTRA_IONS = np.array([140.0, 3.0, 1.2, 130.0, 2.0, 0.0, 4.0e-5, 0.0], dtype=np.float32)

# Canonical densities (reference include/cpg_config.h:16-76, tests/test_cpu_gpu_parity_cuda.cpp:246-247).
CANON_DENSITY = {NAV12: 15.6, NAV16: 15.6, KVDR: 3.84, K2P: 4.17, NALCN: 0.11, PUMP: 1500.0,
                 CAV21: 100.0, HCN1: 120.0, SK: 6.0, BK: 0.15, KV11: 0.346, KV31: 2.0, KV7: 2.0,
                 CX36: 400.0, CX43: 60.0, GLAST: 1.0, GLYT1: 1.0, GAT3: 1.0, EAAT3: 1.0,
                 GLYT2: 1.0, GAT1: 1.0, HCN2: 120.0, PUMP_A3: 1500.0, PUMP_GLIA: 1500.0, PMCA: 300.0,
                 NCX: 30.0, KCC2: 0.093, NKCC1: 0.1, CA_BUFFER: 0.5, CAV13: 2.0, CLC2: 0.8,
                 KV42: 41.51, KV21: 2.595, NAP: 3.0, NAP_TAD: 10.6, KIR: 0.35, NAV12DEV: 60.0}
CANON_DENSITY.update({t: 40.0 for t in RECEPTORS2 + RECEPTORS3})
CANON_DENSITY.update({HCN3: 120.0, HCN4: 120.0, NAV11: 8.0, NAV13: 8.0, NAV17: 8.0, NAV18: 8.0, NAV19: 2.0,
                      KV12: 2.0, KV13: 2.0, KV14: 2.0, KV16: 2.0, KV22: 2.0})
CANON_DENSITY.update({t: 0.1 for t in K_MORE})
CANON_DENSITY.update({t: 2.0 for t in CAV_MORE})
CANON_DENSITY.update({t: 0.05 for t in CATION_MORE})
CANON_DENSITY.update({159: 1.0, 101: 0.005, 96: 2.0})
CANON_DENSITY.update({84: 2.0, 144: 5.0, 145: 1.0, 146: 1.0, 152: 0.05, 186: 1e11, 190: 0.05})
CANON_DENSITY.update({SERCA: 1e5, IP3R: 100.0, RYR: 1.0, PLC: 100.0})
CANON_DENSITY.update({METABOTROPIC: 2.0, RELEASE_SITE: 1.0, SOMA_ATP: 1.0, SOMA_THERMAL: 1.0, CREB: 1.0, MRGPR: 1.0})
CANON_DENSITY.update({TRPV1_R: 0.05, MAPK: 1.0, NFKB: 1.0, MTORC1: 1.0, PROTEIN_SYNTH: 1.0, COFILIN: 1.0, MMP9: 1.0,
                      ROD: 0.05})
CANON_DENSITY.update({t: 1.0 for t in SENSORS})
CANON_DENSITY[PHYS_SENSOR] = 0.8   # cpg_config.h:85
CANON_DENSITY.update({CAMKII: 40.0, CALCINEURIN: 4.0, PP1: 40.0, PKC: 1.0, MSN_NMDA: 40.0, MSN_AMPA: 40.0})
CANON_DENSITY.update({t: 1.0 for t in NO_CODE})
# Gate initial values near rest (-65 mV).
GATE_INIT = {NAV12: {0: 0.02, 1: 0.75}, NAV16: {0: 0.01, 1: 0.6}, KVDR: {0: 0.05},
             CAV21: {0: 0.0, 1: 0.5}, HCN1: {0: 0.3}, KV11: {0: 0.03, 1: 1.0},
             KV31: {0: 0.0, 1: 1.0}, KV7: {0: 0.1, 6: 1.0}, SK: {0: 0.0}, BK: {0: 0.0},
             CX36: {0: 0.8, 1: 0.1, 4: 1.0, 5: 7.2, 6: 0.2}, CX43: {0: 0.6, 4: 1.0, 5: 7.1}}


@dataclasses.dataclass
class Network:
    """A BION network held as structured arrays (the file's AoS records)."""
    neurons: np.ndarray
    segments: np.ndarray
    connections: np.ndarray
    synapses: np.ndarray
    mode: int = 0

    @property
    def num_segments(self) -> int:
        return int(self.segments.shape[0])

    def copy(self) -> "Network":
        return Network(self.neurons.copy(), self.segments.copy(), self.connections.copy(),
                       self.synapses.copy(), self.mode)


def make_header(net: Network) -> np.ndarray:
    h = np.zeros(1, HEADER_DT)
    h["magic"] = BION_MAGIC
    h["version"] = BION_VERSION
    h["mode"] = net.mode
    h["num_neurons"] = len(net.neurons)
    h["num_segments"] = len(net.segments)
    h["num_intra_connections"] = len(net.connections)
    h["num_synapses"] = len(net.synapses)
    h["voxel_size"] = 1.0
    h["timestamp"] = int(time.time())
    h["build_id"] = b"bchip"
    h["nucleus_version"] = 1
    h["file_sz_segment"] = SEGMENT_DT.itemsize
    h["file_sz_synapse"] = SYNAPSE_DT.itemsize
    h["file_sz_nucleus"] = NUCLEUS_SIZE
    h["file_sz_intraconnection"] = CONN_DT.itemsize
    h["file_sz_neuron"] = NEURON_DT.itemsize
    h["file_sz_modifier"] = MODIFIER_DT.itemsize
    return h


def write_bion(path: str, net: Network) -> None:
    with open(path, "wb") as f:
        f.write(make_header(net).tobytes())
        for arr in (net.neurons, net.segments, net.connections, net.synapses):
            f.write(np.ascontiguousarray(arr).tobytes())


def read_bion(path: str) -> Network:
    with open(path, "rb") as f:
        raw = f.read()
    h = np.frombuffer(raw[:100], HEADER_DT)[0]
    if int(h["magic"]) != BION_MAGIC:
        raise ValueError("bad BION magic")
    if int(h["version"]) != BION_VERSION:
        raise ValueError(f"BION version {int(h['version'])} not supported (v14 only)")
    off = 100
    out = []
    for dt, n in ((NEURON_DT, h["num_neurons"]), (SEGMENT_DT, h["num_segments"]),
                  (CONN_DT, h["num_intra_connections"]), (SYNAPSE_DT, h["num_synapses"])):
        n = int(n)
        out.append(np.frombuffer(raw[off:off + n * dt.itemsize], dt).copy())
        off += n * dt.itemsize
    return Network(*out, mode=int(h["mode"]))


# ------------------------------------------------------------------ synthetic networks
def _set_modifiers(segs: np.ndarray, types: list[int], densities: np.ndarray) -> None:
    """types: per-slot IDs; densities: [n_seg, n_slot]."""
    k = len(types)
    mods = segs["modifiers"]
    for j, t in enumerate(types):
        mods["type"][:, j] = t
        mods["density"][:, j] = densities[:, j]
        for idx, val in GATE_INIT.get(t, {}).items():
            mods["state"][:, j, idx] = val
    segs["modifier_count"] = k


def _init_segment_state(segs: np.ndarray, rng: np.random.Generator) -> None:
    n = len(segs)
    jit = rng.uniform(0.95, 1.05, size=(n, 8)).astype(np.float32)
    segs["intra_ions"] = (INTRA_IONS[None, :] * jit).astype(np.float64)
    segs["extra_ions"] = EXTRA_IONS[None, :].astype(np.float64)
    segs["V_m"] = -65.0
    segs["ATP_level"] = 1.0
    segs["temperature"] = 37.0
    segs["ER_Ca2"] = 0.5
    segs["direction"] = (0.0, 1.0, 0.0)


def _tree_parents(n_nodes: int) -> np.ndarray:
    """Parent of each dendrite node in a binary tree laid out breadth-first; node 0 hangs
    off the soma (parent -1 means soma)."""
    return np.array([-1] + [(i - 1) // 2 for i in range(1, n_nodes)], dtype=np.int64)


def build_network(n_neurons: int, seg_per_neuron: int, morphology: str, seg_types: list[int],
                  soma_extra: list[int], synapses_per_neuron: float, receptor_mix: str,
                  max_delay_ticks: int, seed: int, dt: float = 0.025, lognormal_sigma: float = 0.0,
                  density_scale: dict | None = None, pre_pool: str = "soma") -> Network:
    """Soma + dendrite (chain or breadth-first binary tree) neurons with a fixed modifier set
    per segment, random chemical synapses from presynaptic somas onto uniform post segments."""
    rng = np.random.default_rng(seed)
    ns = n_neurons * seg_per_neuron
    segs = np.zeros(ns, SEGMENT_DT)
    nid = np.repeat(np.arange(n_neurons, dtype=np.uint32), seg_per_neuron)
    local = np.tile(np.arange(seg_per_neuron), n_neurons)
    segs["id"] = np.arange(ns, dtype=np.uint32)
    segs["neuron_id"] = nid
    is_soma = local == 0
    segs["segment_type"] = np.where(is_soma, 1, 2)
    segs["radius"] = np.where(is_soma, 5.0, rng.uniform(0.5, 2.0, ns)).astype(np.float32)
    segs["length"] = np.where(is_soma, 10.0, rng.uniform(10.0, 50.0, ns)).astype(np.float32)
    segs["position"][:, 0] = nid.astype(np.float32) * 100.0
    segs["position"][:, 1] = local.astype(np.float32) * 20.0
    _init_segment_state(segs, rng)

    types = list(seg_types)
    dens = np.empty((ns, len(types) + len(soma_extra)), dtype=np.float64)
    for j, t in enumerate(types + list(soma_extra)):
        base = CANON_DENSITY[t] * (density_scale or {}).get(t, 1.0)
        if lognormal_sigma > 0:
            dens[:, j] = base * rng.lognormal(0.0, lognormal_sigma, ns)
        else:
            dens[:, j] = base
    _set_modifiers(segs, types, dens[:, :len(types)])
    if soma_extra:
        soma_idx = np.nonzero(is_soma)[0]
        k = len(types)
        for j, t in enumerate(soma_extra):
            segs["modifiers"]["type"][soma_idx, k + j] = t
            segs["modifiers"]["density"][soma_idx, k + j] = dens[soma_idx, k + j]
            for idx, val in GATE_INIT.get(t, {}).items():
                segs["modifiers"]["state"][soma_idx, k + j, idx] = val
        segs["modifier_count"][soma_idx] = k + len(soma_extra)

    neurons = np.zeros(n_neurons, NEURON_DT)
    neurons["id"] = np.arange(n_neurons)
    neurons["segment_start"] = np.arange(n_neurons) * seg_per_neuron
    neurons["segment_count"] = seg_per_neuron
    neurons["soma_index"] = np.arange(n_neurons) * seg_per_neuron
    neurons["threshold"] = -40.0

    # intracellular connections: chain (i -> i+1) or binary tree (child -> parent)
    m = seg_per_neuron - 1
    if morphology == "chain":
        par_local = np.arange(0, m)            # node k+1 hangs off node k
    elif morphology == "tree":
        par_local = _tree_parents(m) + 1       # node k+1's parent (0 = soma)
    else:
        raise ValueError(morphology)
    base = (np.arange(n_neurons, dtype=np.int64) * seg_per_neuron)[:, None]
    a = (base + par_local[None, :]).reshape(-1)
    b = (base + np.arange(1, seg_per_neuron)[None, :]).reshape(-1)
    conns = np.zeros(len(a), CONN_DT)
    conns["segment_a"] = a
    conns["segment_b"] = b
    conns["is_intracellular"] = 1

    nsyn = int(round(synapses_per_neuron * n_neurons))
    syns = np.zeros(nsyn, SYNAPSE_DT)
    if nsyn:
        post_neuron = np.repeat(np.arange(n_neurons), int(np.ceil(synapses_per_neuron)))[:nsyn] \
            if synapses_per_neuron >= 1 else rng.integers(0, n_neurons, nsyn)
        if synapses_per_neuron >= 1:
            post = post_neuron * seg_per_neuron + rng.integers(0, seg_per_neuron, nsyn)
        else:
            post = rng.integers(0, ns, nsyn)
        pre = rng.integers(0, n_neurons, nsyn) * seg_per_neuron
        syns["id"] = np.arange(1, nsyn + 1, dtype=np.uint32)
        syns["pre_segment"] = pre
        syns["post_segment"] = post
        syns["weight"] = 1.0
        syns["delay"] = (rng.integers(1, max_delay_ticks + 1, nsyn) * np.float32(dt)).astype(np.float32)
        syns["pre_area"] = 0.1
        syns["post_area"] = 0.1
        syns["release_probability"] = 0.5
        syns["release_amount"] = 1.0
        syns["tau_clearance"] = 5.0
        syns["RRP"] = 10.0
        syns["RRP_max"] = 10.0
        syns["RRP_recovery"] = 0.02
        syns["bouton_Ca"] = 0.0001
        syns["spine_Ca"] = 0.0001
        syns["depression"] = 0.01
        pr = syns["post_receptors"]
        if receptor_mix == "ampa":
            pr["type"][:, 0] = AMPA
            pr["density"][:, 0] = 500.0
            syns["post_receptor_count"] = 1
        elif receptor_mix == "ampa_nmda_gaba":
            exc = rng.random(nsyn) < 0.7
            pr["type"][:, 0] = np.where(exc, AMPA, GABAA)
            pr["density"][:, 0] = np.where(exc, 500.0, 300.0)
            pr["type"][:, 1] = np.where(exc, NMDA, 0)
            pr["density"][:, 1] = np.where(exc, 50.0, 0.0)
            pr["state"][:, 1, 2] = np.where(exc, 1.0, 0.0)   # NMDA CDI gate starts open
            syns["post_receptor_count"] = np.where(exc, 2, 1)
        else:
            raise ValueError(receptor_mix)
    return Network(neurons, segs, conns, syns)


def gap_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6) -> Network:
    """Chain neurons with connexins (Cx36 and Cx43 on somas) and AMPA synapses, plus electrical
    synapses: soma-to-soma gap junctions between neighbouring neurons (type 1; every other one
    with a delay, i.e. filtered) and dendritic bridges (type 2), interleaved in index order with
    the chemical synapses."""
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN], [CX36, CX43],
                        3.0, "ampa", 4, seed)
    rng = np.random.default_rng(seed + 1)
    e = []
    for k in range(n_neurons - 1):
        e.append((k * seg_per_neuron, (k + 1) * seg_per_neuron, 1, 0.5 if k % 2 else 0.0))
        d = 1 + k % (seg_per_neuron - 1)
        e.append(((k + 1) * seg_per_neuron + d, k * seg_per_neuron + d, 2, 0.0))
    gj = np.zeros(len(e), SYNAPSE_DT)
    gj["pre_segment"] = [x[0] for x in e]
    gj["post_segment"] = [x[1] for x in e]
    gj["type"] = [x[2] for x in e]
    gj["delay"] = [x[3] for x in e]
    gj["depression"] = rng.uniform(0.0, 0.3, len(e)).astype(np.float32)
    gj["weight"] = 1.0
    syn = np.concatenate([net.synapses, gj])
    order = rng.permutation(len(syn))
    syn = syn[order]
    syn["id"] = np.arange(1, len(syn) + 1, dtype=np.uint32)
    net.synapses = syn
    return net


def spinal_glia_network(seed: int, n_spinal: int = 4, per_spinal: int = 12, seg_per_neuron: int = 6,
                        glia_per_spinal: int = 2, synapses: bool = True) -> Network:
    """One connected island far larger than a tile, built as the reference's spinal network
    couples its glia (include/cpg_config.h:657-713): chain neurons in n_spinal spinal segments,
    single-segment glial cells (GLAST, pump, K2P) each coupled extracellularly to every segment of
    the neurons of its own and the neighbouring spinal segments, and each soma coupled
    extracellularly to the soma of the same index in the next spinal segment. With `synapses`,
    AMPA+NMDA / GABA-A synapses between somas (delays 1-4 ticks); without, the extracellular
    transmitters diffuse over the coupling."""
    nn = n_spinal * per_spinal
    net = build_network(nn, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], [CAV21],
                        4.0 if synapses else 0.0, "ampa_nmda_gaba", 4, seed)
    rng = np.random.default_rng(seed + 11)
    ng = n_spinal * glia_per_spinal
    g = np.zeros(ng, SEGMENT_DT)
    _init_segment_state(g, rng)
    g["segment_type"] = 1
    g["radius"], g["length"] = 4.0, 8.0
    g["neuron_id"] = nn + np.arange(ng)
    gt = [GLAST, PUMP, K2P]
    _set_modifiers(g, gt, np.array([[CANON_DENSITY[t] for t in gt]] * ng, np.float64))
    n0 = len(net.segments)
    net.segments = np.concatenate([net.segments, g])
    net.segments["id"] = np.arange(len(net.segments), dtype=np.uint32)
    gneu = np.zeros(ng, NEURON_DT)
    gneu["id"] = nn + np.arange(ng)
    gneu["segment_start"] = gneu["soma_index"] = n0 + np.arange(ng)
    gneu["segment_count"] = 1
    gneu["threshold"] = -40.0
    net.neurons = np.concatenate([net.neurons, gneu])
    e = []
    span = per_spinal * seg_per_neuron
    for k in range(ng):
        s = k // glia_per_spinal
        for ds in (-1, 0, 1):
            if 0 <= s + ds < n_spinal:
                e += [(n0 + k, x) for x in range((s + ds) * span, (s + ds + 1) * span)]
    for s in range(n_spinal - 1):
        e += [(s * span + j * seg_per_neuron, (s + 1) * span + j * seg_per_neuron) for j in range(per_spinal)]
    ext = np.zeros(len(e), CONN_DT)
    ext["segment_a"] = [x[0] for x in e]
    ext["segment_b"] = [x[1] for x in e]
    ext["is_intracellular"] = 0
    net.connections = np.concatenate([net.connections, ext])
    return net


def long_cell(seed: int, n_seg: int = 1000, n_small: int = 6) -> Network:
    """A 1000-compartment cell (soma + binary-branching dendrite; Nav1.2, Kv-DR, K2P, NALCN, pump)
    followed by n_small 8-segment chain neurons, no synapses: one island beyond any tile next to
    islands that tile."""
    big = build_network(1, n_seg, "tree", [NAV12, KVDR, K2P, NALCN, PUMP], [], 0.0, "ampa", 1, seed)
    small = build_network(n_small, 8, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], [], 0.0, "ampa", 1, seed + 1)
    small.segments["id"] += n_seg
    small.segments["neuron_id"] += 1
    small.neurons["id"] += 1
    small.neurons["segment_start"] += n_seg
    small.neurons["soma_index"] += n_seg
    small.connections["segment_a"] += n_seg
    small.connections["segment_b"] += n_seg
    return Network(np.concatenate([big.neurons, small.neurons]), np.concatenate([big.segments, small.segments]),
                   np.concatenate([big.connections, small.connections]), np.zeros(0, SYNAPSE_DT))


def hazard_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6, glia: bool = False) -> Network:
    """A recurrent network whose step-3 results depend on synapse order across post chains:
    somas (the presynaptic segments) receive AMPA+NMDA / GABA-A synapses; a third of all
    chemical synapses release immediately (delay 0, or 0.01 ms, which rounds to 0 ticks at
    dt = 0.025 ms), so their vesicle recovery reads a presynaptic Ca that earlier synapses onto
    that soma changed; soma-to-soma gap junctions (filtered and not) are interleaved in index
    order behind those synapses. With `glia`, two glial (GLAST) segments capture cleft
    glutamate and GABA."""
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN], [CAV21, CX36, CX43],
                        8.0, "ampa_nmda_gaba", 3, seed)
    rng = np.random.default_rng(seed + 7)
    y = net.synapses
    onto_soma = rng.random(len(y)) < 0.6
    y["post_segment"][onto_soma] = rng.integers(0, n_neurons, int(onto_soma.sum())) * seg_per_neuron
    now = rng.random(len(y)) < 0.35
    y["delay"][now] = np.where(rng.random(int(now.sum())) < 0.7, 0.0, 0.01).astype(np.float32)
    e = []
    for k in range(n_neurons - 1):
        e.append((k * seg_per_neuron, (k + 1) * seg_per_neuron, 1, 0.5 if k % 2 else 0.0))
    gj = np.zeros(len(e), SYNAPSE_DT)
    gj["pre_segment"] = [x[0] for x in e]
    gj["post_segment"] = [x[1] for x in e]
    gj["type"] = [x[2] for x in e]
    gj["delay"] = [x[3] for x in e]
    gj["depression"] = rng.uniform(0.0, 0.3, len(e)).astype(np.float32)
    gj["weight"] = 1.0
    syn = np.concatenate([y, gj])
    syn = syn[rng.permutation(len(syn))]
    syn["id"] = np.arange(1, len(syn) + 1, dtype=np.uint32)
    net.synapses = syn
    if glia:   # two dendritic segments become glial capture sites, segments spread over depth
        s = net.segments
        s["position"][:, 2] = rng.uniform(0.0, 400.0, len(s)).astype(np.float32)
        for g in (2, seg_per_neuron + 3):
            c = s["modifier_count"][g]
            s["modifiers"][g, c]["type"] = GLAST
            s["modifiers"][g, c]["density"] = CANON_DENSITY.get(GLAST, 1.0)
            s["modifier_count"][g] = c + 1
    return net


def glia_network(seed: int, n_neurons: int = 16, seg_per_neuron: int = 6, n_glia: int = 5,
                 glia_first: bool = False) -> Network:
    """Chain neurons with AMPA+NMDA / GABA-A synapses and neuronal transporters (EAAT3, GlyT2,
    GAT1 on somas), spread over depth z, plus n_glia single-segment glial cells carrying GLAST,
    GlyT1, GAT3, the pump and K2P (glia_first: the glial cells come first in segment order).
    Glial depths repeat in pairs, so some synapses have two equally near glia (the first in
    segment order takes the capture); every fourth glial cell also receives synapses."""
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP],
                        [CAV21, EAAT3, GLYT2, GAT1], 3.0, "ampa_nmda_gaba", 4, seed)
    rng = np.random.default_rng(seed + 7)
    segs = net.segments
    segs["position"][:, 2] = np.repeat(rng.uniform(0.0, 400.0, n_neurons), seg_per_neuron) \
        + np.tile(np.arange(seg_per_neuron) * 3.0, n_neurons)
    g = np.zeros(n_glia, SEGMENT_DT)
    _init_segment_state(g, rng)
    g["segment_type"] = 1
    g["radius"], g["length"] = 4.0, 8.0
    g["neuron_id"] = n_neurons + np.arange(n_glia)
    g["position"][:, 0] = -50.0
    g["position"][:, 2] = np.repeat(rng.uniform(0.0, 400.0, (n_glia + 1) // 2), 2)[:n_glia]
    gt = [GLAST, GLYT1, GAT3, PUMP, K2P]
    _set_modifiers(g, gt, np.array([[CANON_DENSITY[t] for t in gt]] * n_glia, np.float64))
    ng = n_neurons * seg_per_neuron
    gneu = np.zeros(n_glia, NEURON_DT)
    gneu["id"] = n_neurons + np.arange(n_glia)
    gneu["segment_count"] = 1
    gneu["threshold"] = -40.0
    syn = net.synapses
    gsyn = syn[rng.integers(0, len(syn), (n_glia + 3) // 4)].copy()
    if glia_first:   # glia at 0..n_glia-1, neurons after them
        net.segments = np.concatenate([g, segs])
        gneu["segment_start"] = gneu["soma_index"] = np.arange(n_glia)
        net.neurons["segment_start"] += n_glia
        net.neurons["soma_index"] += n_glia
        net.connections["segment_a"] += n_glia
        net.connections["segment_b"] += n_glia
        syn["pre_segment"] += n_glia
        syn["post_segment"] += n_glia
        gsyn["pre_segment"] += n_glia
        gsyn["post_segment"] = np.arange(0, n_glia, 4)
        net.neurons = np.concatenate([gneu, net.neurons])
    else:
        net.segments = np.concatenate([segs, g])
        gneu["segment_start"] = gneu["soma_index"] = ng + np.arange(n_glia)
        gsyn["post_segment"] = ng + np.arange(0, n_glia, 4)
        net.neurons = np.concatenate([net.neurons, gneu])
    net.segments["id"] = np.arange(len(net.segments), dtype=np.uint32)
    syn = np.concatenate([syn, gsyn])
    syn = syn[rng.permutation(len(syn))]
    syn["id"] = np.arange(1, len(syn) + 1, dtype=np.uint32)
    net.synapses = syn
    return net


def homeostasis_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6, n_glia: int = 4) -> Network:
    """Chain neurons whose somas carry the ion-homeostasis set (alpha3 pump, PMCA, NCX, KCC2,
    NKCC1, Ca2+ buffer, HCN2) next to Cav2.1, dendrites with PMCA and the Ca2+ buffer, AMPA+NMDA /
    GABA-A synapses (NMDA and Cav2.1 bring Ca2+ in), plus glial cells with the glial pump, K2P,
    NKCC1, PMCA and the buffer."""
    soma = [CAV21, PUMP_A3, PMCA, NCX, KCC2, NKCC1, CA_BUFFER, HCN2]
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], soma,
                        3.0, "ampa_nmda_gaba", 4, seed)
    segs = net.segments
    rng = np.random.default_rng(seed + 3)
    for i in np.nonzero(segs["segment_type"] == 2)[0][::2]:
        k = int(segs["modifier_count"][i])
        for j, (t, d) in enumerate(((PMCA, 300.0), (CA_BUFFER, 0.5))):
            segs["modifiers"]["type"][i, k + j] = t
            segs["modifiers"]["density"][i, k + j] = d
        segs["modifier_count"][i] = k + 2
    segs["intra_ions"][:, 2] *= rng.uniform(1.0, 20.0, len(segs))   # some Ca loads to remove
    g = np.zeros(n_glia, SEGMENT_DT)
    _init_segment_state(g, rng)
    g["segment_type"], g["radius"], g["length"] = 1, 4.0, 8.0
    g["neuron_id"] = n_neurons + np.arange(n_glia)
    g["intra_ions"][:, 3] = 25.0
    gt = [PUMP_GLIA, K2P, NKCC1, PMCA, CA_BUFFER]
    _set_modifiers(g, gt, np.array([[CANON_DENSITY[t] for t in gt]] * n_glia, np.float64))
    gneu = np.zeros(n_glia, NEURON_DT)
    gneu["id"] = n_neurons + np.arange(n_glia)
    gneu["segment_start"] = gneu["soma_index"] = len(segs) + np.arange(n_glia)
    gneu["segment_count"] = 1
    gneu["threshold"] = -40.0
    net.segments = np.concatenate([segs, g])
    net.segments["id"] = np.arange(len(net.segments), dtype=np.uint32)
    net.neurons = np.concatenate([net.neurons, gneu])
    return net


def channel_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 5, myelin: bool = False) -> Network:
    """Chain neurons built on the developmental Nav1.2 with Kv2.1, Kv4.2, Kv1.1, K2P, NALCN and
    the pump on every segment; somas add Cav1.3, Cav2.1, NaP (Nav1.6 and tadpole), Kir2.1, the
    ClC-2 leak and HCN2; AMPA+NMDA / GABA-A synapses. myelin: every third segment also carries
    myelin, which blocks the listed channel types and the ionotropic receptors."""
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12DEV, KV21, KV42, KV11, K2P, NALCN, PUMP],
                        [CAV13, CAV21, NAP, NAP_TAD, KIR, CLC2, HCN2] + list(NAV_VARIANTS) + list(KV_VARIANTS)
                        + list(K_MORE[:6]), 3.0,
                        "ampa_nmda_gaba", 4, seed,
                        lognormal_sigma=0.2)
    mods = net.segments["modifiers"]
    for t, init in ((NAV12DEV, {0: 0.02, 1: 0.7, 3: 0.7}), (KV21, {0: 0.05, 1: 0.9}), (KV42, {0: 0.1, 1: 0.5})):
        i, j = np.nonzero(mods["type"] == t)
        for k, v in init.items():
            mods["state"][i, j, k] = v
    segs = net.segments
    for i in np.nonzero(segs["segment_type"] == 2)[0][::2]:   # the rest of the K+ set on dendrites
        k = int(segs["modifier_count"][i])
        extra = K_MORE[6:] + CAV_MORE + CATION_MORE + MISC_MORE
        for j, t in enumerate(extra):
            mods["type"][i, k + j], mods["density"][i, k + j] = t, CANON_DENSITY[t]
        segs["modifier_count"][i] = k + len(extra)
    for n, i in enumerate(np.nonzero(segs["segment_type"] == 2)[0][1::2]):   # Ca-activated set
        k = int(segs["modifier_count"][i])
        extra = CA_ACT if n % 2 else (84, 145, 144, 146, 152, 186, 190)   # CaBP4 before/after Cav1.4
        for j, t in enumerate(extra):
            mods["type"][i, k + j], mods["density"][i, k + j] = t, CANON_DENSITY[t]
        segs["modifier_count"][i] = k + len(extra)
    for t, k, v in ((210, 4, 0.4), (162, 4, 0.3), (175, 4, 30.0), (172, 4, 0.5), (173, 4, 20.0),
                    (174, 4, 20.0), (179, 4, 5.0), (163, 4, 0.3)):   # DAG, Gq, ligands
        i, j = np.nonzero(mods["type"] == t)
        mods["state"][i, j, k] = v
    i, j = np.nonzero(mods["type"] == 160)
    mods["state"][i, j, 4] = 0.3    # GIRK: some G-beta-gamma
    i, j = np.nonzero(mods["type"] == 190)
    mods["state"][i[::2], j[::2], 4] = 0.5   # TMC1: half the bundles deflected
    i, j = np.nonzero(mods["type"] == 161)
    mods["state"][i, j, 4] = 5.0    # KATP: partly closed by ATP
    if myelin:
        segs = net.segments
        for i in range(0, len(segs), 3):
            k = int(segs["modifier_count"][i])
            mods["type"][i, k], mods["density"][i, k] = 4, 1.0
            segs["modifier_count"][i] = k + 1
    return net


def receptor_network(seed: int, n_neurons: int = 14, seg_per_neuron: int = 5) -> Network:
    """Chain neurons whose synapses carry, by index, glycine, nicotinic, 5-HT3+kainate, GluN2B
    NMDA, GABA-B, D1 and alpha7+D2 receptors besides AMPA+NMDA and GABA-A (so glycine,
    acetylcholine, serotonin, GABA and dopamine are released); somas carry Gi/o (GABA-B, 5-HT1A,
    D2, M2, mGluR2, CB1, H3, MT1) and Gs (D1, 5-HT7, beta1, A2A, H2) receptors as segment
    modifiers, dendrites glycine and GluN2B receptors."""
    soma = [17, 18, 22, 25, 29, 36, 38, 39, 44, 43, 48, 46, 50, 52, 55, 57, 60, 62, 70, 63, 64, 67, 74, HCN3, HCN4]
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], [CAV21] + soma,
                        4.0, "ampa_nmda_gaba", 4, seed)
    segs = net.segments
    mods = segs["modifiers"]
    soma_idx = np.nonzero(segs["segment_type"] == 1)[0]
    for j in range(6, 6 + len(soma)):
        mods["density"][soma_idx, j] = 40.0
    mods["active_vars"][soma_idx[::2], 6] = 3            # EC50/gamma from state[3]/state[4]
    mods["state"][soma_idx[::2], 6, 3] = 0.4
    for i in np.nonzero(segs["segment_type"] == 2)[0][::2]:
        k = int(segs["modifier_count"][i])
        mods["type"][i, k], mods["density"][i, k] = GLYCINE, 100.0
        mods["type"][i, k + 1], mods["density"][i, k + 1] = NMDA2B, 20.0
        mods["state"][i, k + 1, 2] = 1.0
        segs["modifier_count"][i] = k + 2
    syn = net.synapses
    pr = syn["post_receptors"]
    kinds = [[GLYCINE], [NICOTINIC], [HT3, KAINATE], [NMDA2B], [GABAB], [44], [ALPHA7, 22], [P2X2, 71, 72],
             [73, 68, P2X, 57]]
    for y in range(len(syn)):
        c = y % 11
        if c >= len(kinds):
            continue
        ts = kinds[c]
        pr["type"][y, :] = 0
        pr["density"][y, :] = 0.0
        pr["state"][y, :, :] = 0.0
        for j, t in enumerate(ts):
            pr["type"][y, j] = t
            pr["density"][y, j] = 300.0 if t not in (GABAB, 44, 22, 57) else 30.0
        if NMDA2B in ts:
            pr["state"][y, ts.index(NMDA2B), 2] = 1.0
        syn["post_receptor_count"][y] = len(ts)
    return net


GPCR_REST = (19, 20, 21, 23, 24, 26, 27, 28, 30, 31, 32, 33, 34, 35, 37, 40, 41, 42, 45, 47, 49, 51, 53, 54,
             56, 58, 59, 61)   # the G-protein-coupled IDs receptor_network does not place


def er_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6) -> Network:
    """Chain neurons whose somas carry Cav2.1 (Ca2+ in), PMCA, the ER store (SERCA, IP3R, RyR) and
    PLC driven by mGluR5 (a Gq receptor of glutamate), with AMPA+NMDA / GABA-A synapses that put
    glutamate into the clefts; every other dendrite segment carries SERCA and RyR. ER Ca2+ starts
    in 0.05-1 mM, cytosolic Ca2+ at 1-20x rest, IP3 in 0-0.5 uM; RyR and IP3R gates start at 0
    (steady-state initialisation) on every other soma."""
    soma = [CAV21, PMCA, SERCA, IP3R, RYR, PLC, 61]
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], soma,
                        3.0, "ampa_nmda_gaba", 4, seed)
    segs = net.segments
    mods = segs["modifiers"]
    rng = np.random.default_rng(seed + 5)
    n = len(segs)
    segs["ER_Ca2"] = rng.uniform(0.05, 1.0, n).astype(np.float32)
    segs["intra_ions"][:, 2] *= rng.uniform(1.0, 20.0, n)
    soma_idx = np.nonzero(segs["segment_type"] == 1)[0]
    base = 5   # the dendrite set's five slots come first
    mods["state"][soma_idx, base + 5, 0] = rng.uniform(0.0, 0.0005, len(soma_idx))   # PLC IP3 (mM)
    mods["state"][soma_idx[1::2], base + 3, 1] = rng.uniform(0.2, 0.9, len(soma_idx[1::2]))   # IP3R h
    mods["state"][soma_idx[1::2], base + 4, 0] = rng.uniform(0.2, 0.9, len(soma_idx[1::2]))   # RyR w
    mods["state"][soma_idx[::3], base + 2, 1] = 0.0005   # a SERCA K_m override
    for i in np.nonzero(segs["segment_type"] == 2)[0][::2]:
        k = int(segs["modifier_count"][i])
        for j, t in enumerate((SERCA, RYR)):
            mods["type"][i, k + j] = t
            mods["density"][i, k + j] = CANON_DENSITY[t] * rng.uniform(0.5, 2.0)
        segs["modifier_count"][i] = k + 2
    return net


def signalling_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6) -> Network:
    """Chain neurons whose somas carry Cav2.1, the metabotropic G protein (3) on the summed
    transmitters, the SK-like release site (5), somatic ATP synthesis (77) and thermoregulation
    (78), CREB transcription (79), and MRGPR (188) with PLC, with AMPA+NMDA / GABA-A synapses;
    every other dendrite segment carries the metabotropic G protein and the release site.
    Temperatures start in 35-41 C, ATP in 0.5-4, cytosolic Ca2+ at 1-20x rest, MRGPR ligand in
    0-60 uM on every other soma."""
    soma = [CAV21, METABOTROPIC, RELEASE_SITE, SOMA_ATP, SOMA_THERMAL, CREB, MRGPR, PLC]
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], soma,
                        3.0, "ampa_nmda_gaba", 4, seed)
    segs = net.segments
    mods = segs["modifiers"]
    rng = np.random.default_rng(seed + 7)
    n = len(segs)
    segs["temperature"] = rng.uniform(35.0, 41.0, n).astype(np.float32)
    segs["ATP_level"] = rng.uniform(0.5, 4.0, n).astype(np.float32)
    segs["intra_ions"][:, 2] *= rng.uniform(1.0, 20.0, n)
    soma_idx = np.nonzero(segs["segment_type"] == 1)[0]
    base = 5
    mods["state"][soma_idx[::2], base + 6, 4] = rng.uniform(0.0, 60.0, len(soma_idx[::2]))   # MRGPR ligand
    mods["state"][soma_idx[1::3], base + 1, 2] = 0.5    # metabotropic coupling override
    mods["state"][soma_idx[1::3], base + 2, 4] = 40.0   # release-site facilitation tau
    mods["state"][soma_idx[::3], base + 4, 1] = 800.0   # faster thermoregulation
    mods["state"][soma_idx[::4], base + 5, 3] = 0.0002  # CREB Ca2+ K_d override
    for i in np.nonzero(segs["segment_type"] == 2)[0][::2]:
        k = int(segs["modifier_count"][i])
        for j, t in enumerate((METABOTROPIC, RELEASE_SITE)):
            mods["type"][i, k + j] = t
            mods["density"][i, k + j] = CANON_DENSITY[t] * rng.uniform(0.5, 2.0)
        segs["modifier_count"][i] = k + 2
    return net


def cascade_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6) -> Network:
    """Chain neurons whose somas carry Cav2.1, the MAPK/ERK, NF-kB, mTORC1, protein-synthesis,
    cofilin and MMP-9 cascades with CREB (MAPK raises every CREB's pCREB; protein synthesis,
    cofilin and MMP-9 read the last mTORC1, MAPK and NF-kB), thermoregulation, the TRPV1
    transmitter receptor (heat and capsaicin, transmitter 14) and rod phototransduction, with
    AMPA+NMDA / GABA-A synapses. Temperatures in 36-46 C, amino-acid inputs and light levels
    vary per soma; every other dendrite segment carries MAPK and CREB."""
    soma = [CAV21, MAPK, NFKB, MTORC1, PROTEIN_SYNTH, COFILIN, MMP9, CREB, SOMA_THERMAL, TRPV1_R, ROD]
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], soma,
                        3.0, "ampa_nmda_gaba", 4, seed)
    segs = net.segments
    mods = segs["modifiers"]
    rng = np.random.default_rng(seed + 9)
    n = len(segs)
    segs["temperature"] = rng.uniform(36.0, 46.0, n).astype(np.float32)
    segs["intra_ions"][:, 2] *= rng.uniform(1.0, 50.0, n)
    segs["neurotransmitter"][:, 14] = rng.uniform(0.0, 2.0, n).astype(np.float32)   # capsaicin
    soma_idx = np.nonzero(segs["segment_type"] == 1)[0]
    base = 5
    mods["state"][soma_idx, base + 3, 3] = rng.uniform(0.0, 1.0, len(soma_idx))   # mTORC1 amino acids
    mods["state"][soma_idx[::2], base + 10, 4] = rng.uniform(0.0, 1.0, len(soma_idx[::2]))   # rod light
    mods["state"][soma_idx, base + 10, 3] = 0.8   # rod cGMP in the dark
    mods["density"][soma_idx, base + 1:base + 7] *= rng.uniform(20.0, 200.0, (len(soma_idx), 6))   # faster cascades
    mods["state"][soma_idx[1::3], base + 1, 4] = 0.2   # a MAPK K_d override
    for i in np.nonzero(segs["segment_type"] == 2)[0][::2]:
        k = int(segs["modifier_count"][i])
        for j, t in enumerate((MAPK, CREB)):
            mods["type"][i, k + j] = t
            mods["density"][i, k + j] = 50.0 * rng.uniform(0.5, 2.0)
        segs["modifier_count"][i] = k + 2
    return net


def sensor_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6) -> Network:
    """Chain neurons whose somas carry Cav2.1 and every sensor: physical (temperature and stretch),
    hunger (ATP), light, olfactory, thermal noise (seeded by Segment.id and V), vibration, otolith
    hair cell, semicircular cupula and cryptochrome (which boosts the first light sensor, here
    before it in slot order), with AMPA+NMDA / GABA-A synapses; every other dendrite segment
    carries a cryptochrome ahead of a light sensor. Temperatures in 30-41 C, lengths in 20-120 um,
    ATP in 0.2-4; stimulus states (light, ligand, vibration, shear and its direction, angular
    acceleration, sin^2 of the field angle) vary per segment."""
    soma = [CAV21, PHYS_SENSOR, HUNGER, LIGHT, OLFACTORY, THERMAL_NOISE, VIBRATION, OTOLITH, CUPULA,
            CRYPTOCHROME]
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], soma,
                        3.0, "ampa_nmda_gaba", 4, seed)
    segs = net.segments
    mods = segs["modifiers"]
    rng = np.random.default_rng(seed + 11)
    n = len(segs)
    segs["temperature"] = rng.uniform(30.0, 41.0, n).astype(np.float32)
    segs["length"] = rng.uniform(20.0, 120.0, n).astype(np.float32)
    segs["ATP_level"] = rng.uniform(0.2, 4.0, n).astype(np.float32)
    segs["id"] = rng.permutation(n).astype(np.uint32) * 3 + 17
    soma_idx = np.nonzero(segs["segment_type"] == 1)[0]
    ns = len(soma_idx)
    base = 5
    st = mods["state"]
    st[soma_idx[::2], base + 1, 1] = rng.uniform(0.5, 2.0, len(soma_idx[::2]))   # physical sensitivity
    st[soma_idx[1::2], base + 1, 6] = 500.0                                       # faster habituation
    st[soma_idx, base + 3, 1] = rng.uniform(0.0, 0.5, ns)                         # light intensity
    st[soma_idx, base + 4, 1] = rng.uniform(0.0, 0.03, ns)                        # odorant
    st[soma_idx[::2], base + 5, 0] = rng.uniform(0.02, 0.2, len(soma_idx[::2]))   # noise sigma
    st[soma_idx, base + 5, 1] = rng.uniform(0.0, 50.0, ns)                        # noise seed offset
    st[soma_idx, base + 6, 1] = rng.uniform(0.0, 1.0, ns)                         # vibration
    st[soma_idx, base + 7, 1] = rng.uniform(0.0, 1.0, ns)                         # shear
    st[soma_idx, base + 7, 5] = rng.uniform(-1.0, 1.0, ns)                        # shear direction
    st[soma_idx[::3], base + 7, 5] = -100.0                                       # unset: along it
    st[soma_idx, base + 8, 1] = rng.uniform(-1.0, 1.0, ns)                        # angular acceleration
    st[soma_idx[1::3], base + 8, 4] = 2.0                                         # stiffer cupula
    st[soma_idx, base + 9, 1] = rng.uniform(0.0, 1.0, ns)                         # light at the cryptochrome
    st[soma_idx, base + 9, 2] = rng.uniform(0.0, 1.0, ns)                         # sin^2 of the field angle
    for i in np.nonzero(segs["segment_type"] == 2)[0][::2]:
        k = int(segs["modifier_count"][i])
        for j, t in enumerate((CRYPTOCHROME, LIGHT)):
            mods["type"][i, k + j] = t
            mods["density"][i, k + j] = CANON_DENSITY[t] * rng.uniform(0.5, 2.0)
        st[i, k, 1:3] = rng.uniform(0.2, 1.0, 2)
        st[i, k + 1, 0] = rng.uniform(0.3, 0.9)
        st[i, k + 1, 1] = rng.uniform(0.0, 0.3)
        segs["modifier_count"][i] = k + 2
    return net


def kinase_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6) -> Network:
    """Chain neurons whose somas carry Cav2.1, CaMKII, calcineurin, PP1 / DARPP-32, PLC/IP3/DAG
    (DAG from its initial state), PKC, Cx36, Kv4.2 and the MSN NMDA/AMPA receptors (75/76, scale
    1 without D1/D2 modulators) beside Nav1.2 (a PKC reader), with AMPA+NMDA / GABA-A synapses
    (every third excitatory one with the MSN pair as its receptors): vesicle recovery reads the presynaptic CaMKII and calcineurin,
    plasticity the post segment's PP1. Every other dendrite segment appends PLC, PKC, CaMKII, Cx36,
    PP1, Kv4.2, a second CaMKII and PKC (the readers take the last CaMKII and the first PKC),
    calcineurin and the reserved IDs 80, 124 and 143 with non-zero states. Intracellular Ca
    varies 1-50x per segment."""
    soma = [CAV21, CAMKII, CALCINEURIN, PP1, PLC, PKC, CX36, KV42, MSN_NMDA, MSN_AMPA]
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], soma,
                        3.0, "ampa_nmda_gaba", 4, seed)
    pr = net.synapses["post_receptors"]
    msn = np.nonzero(net.synapses["post_receptor_count"] == 2)[0][::3]   # MSN receptors on some
    pr["type"][msn, 0] = MSN_AMPA
    pr["type"][msn, 1] = MSN_NMDA
    segs = net.segments
    mods = segs["modifiers"]
    rng = np.random.default_rng(seed + 11)
    n = len(segs)
    segs["intra_ions"][:, 2] *= rng.uniform(1.0, 50.0, n)
    base = 5
    soma_idx = np.nonzero(segs["segment_type"] == 1)[0]
    mods["state"][soma_idx, base + 4, 1] = rng.uniform(0.05, 1.0, len(soma_idx))   # DAG
    mods["state"][soma_idx, base + 1, 0] = rng.uniform(0.0, 0.5, len(soma_idx))    # CaMKII
    mods["state"][soma_idx, base + 3, 0] = rng.uniform(0.0, 1.0, len(soma_idx))    # p-DARPP-32
    mods["state"][soma_idx, base + 3, 1] = rng.uniform(0.0, 1.0, len(soma_idx))    # p-inhibitor-1
    mods["state"][soma_idx[::3], base + 1, 2] = 0.6   # a CaMKII K_d override
    for i in np.nonzero(segs["segment_type"] == 2)[0][::2]:
        k = int(segs["modifier_count"][i])
        extra = (PLC, PKC, CAMKII, CX36, PP1, KV42, CAMKII, PKC, CALCINEURIN, 80, 124, 143)
        for j, t in enumerate(extra):
            mods["type"][i, k + j] = t
            mods["density"][i, k + j] = CANON_DENSITY[t] * rng.uniform(0.5, 2.0)
            for q, v in GATE_INIT.get(t, {}).items():
                mods["state"][i, k + j, q] = v
        mods["state"][i, k, 1] = rng.uniform(0.05, 1.0)        # DAG
        mods["state"][i, k + 2, 0] = rng.uniform(0.0, 0.5)     # the first CaMKII
        mods["state"][i, k + 4, 0] = rng.uniform(0.0, 1.0)     # p-DARPP-32
        for j in (9, 10, 11):
            mods["state"][i, k + j, :7] = rng.uniform(0.1, 1.0, 7)
            mods["active_vars"][i, k + j] = 2
        segs["modifier_count"][i] = k + len(extra)
    return net


def gpcr_network(seed: int, n_neurons: int = 8, seg_per_neuron: int = 5) -> Network:
    """Chain neurons without synapses whose somas carry the first half of GPCR_REST and whose
    dendrites (local segments 1 and 3) carry the second half, as segment modifiers; with no
    synapse the segments' transmitter levels persist and diffuse, so ligand injections
    (gpcr_ligands) reach every receptor's transmitter. Every other soma's first receptor takes
    parameter states (active_vars 3)."""
    half = len(GPCR_REST) // 2
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP],
                        list(GPCR_REST[:half]), 0.0, "ampa", 1, seed)
    rng = np.random.default_rng(seed + 3)
    segs = net.segments
    mods = segs["modifiers"]
    soma_idx = np.nonzero(segs["segment_type"] == 1)[0]
    for j in range(5, 5 + half):
        mods["density"][soma_idx, j] = rng.uniform(20.0, 60.0, len(soma_idx))
    mods["active_vars"][soma_idx[::2], 5] = 3
    mods["state"][soma_idx[::2], 5, 3] = 0.4
    local = np.arange(len(segs)) % seg_per_neuron
    for i in np.nonzero((local == 1) | (local == 3))[0]:
        k = int(segs["modifier_count"][i])
        for q, t in enumerate(GPCR_REST[half:]):
            mods["type"][i, k + q] = t
            mods["density"][i, k + q] = rng.uniform(20.0, 60.0)
        segs["modifier_count"][i] = k + len(GPCR_REST) - half
    return net


# The reference's built-in presynaptic release modulators (receptors_biological.cpp:455-491):
# 5-HT1A, 5-HT1B, D2, A1, A2A, D1, mGluR2, mGluR3, CB1, H3, nicotinic, alpha2.
RELEASE_MODULATORS = (18, 19, 22, 27, 46, 44, 29, 30, 36, 38, 13, 35)


def modulated_network(seed: int, n_neurons: int = 16, seg_per_neuron: int = 6, autoreceptors: bool = True,
                      ecb: bool = True) -> Network:
    """Chain neurons with somatic Cav2.1 and 8 synapses per neuron whose first post receptor is
    AMPA, 5-HT3, nicotinic, D1 or D2 (so releases put glutamate, serotonin, acetylcholine or
    dopamine into the cleft). With `autoreceptors`, each synapse carries 0-4 presynaptic
    autoreceptors drawn from RELEASE_MODULATORS and one ID without a modulator (125), with
    densities 0.5-15 (some factors reach zero or below), random state[0] and state[7]; clefts hold
    adenosine, noradrenaline, histamine and eCB at load and pre areas vary, so every modulator
    sees its transmitter. With `ecb`, local segments 2 and 5 of every neuron and the first soma
    carry eCB synthesis (86) with random 2-AG levels and rates, so their incoming synapses take
    the retrograde deposits in synapse order. A few synapses start refractory."""
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], [CAV21], 8.0,
                        "ampa", 5, seed)
    rng = np.random.default_rng(seed + 11)
    y = net.synapses
    ny = len(y)
    rt = np.array([AMPA, 14, 13, 44, 22], np.uint8)[rng.integers(0, 5, ny)]
    pr = y["post_receptors"]
    pr["type"][:, 0] = rt
    pr["density"][:, 0] = np.where((rt == 44) | (rt == 22), 40.0, 500.0)
    y["pre_area"] = rng.uniform(0.05, 0.3, ny).astype(np.float32)
    refr = rng.random(ny) < 0.1
    y["refractory"][refr] = rng.integers(1, 40, int(refr.sum()))
    if autoreceptors:
        pool = np.array(RELEASE_MODULATORS + (125,), np.uint8)
        na = rng.integers(0, 5, ny)
        ar = y["pre_autoreceptors"]
        for k in range(ny):
            for j in range(int(na[k])):
                ar["type"][k, j] = pool[rng.integers(0, len(pool))]
                ar["density"][k, j] = rng.uniform(0.5, 15.0)
                ar["state"][k, j, 0] = rng.uniform(0.0, 0.6)
                ar["state"][k, j, 7] = rng.uniform(-0.1, 0.4)
        y["pre_autoreceptor_count"] = na
        some = rng.random(ny) < 0.5
        for nt in (5, 6, 8, 9):
            y["cleft_neurotransmitter"][some, nt] = rng.uniform(0.0, 1.5, int(some.sum())).astype(np.float32)
    if ecb:
        segs = net.segments
        local = np.arange(len(segs)) % seg_per_neuron
        carriers = np.nonzero((local == 2) | (local == 5))[0].tolist() + [0]
        mods = segs["modifiers"]
        for i in carriers:
            k = int(segs["modifier_count"][i])
            mods["type"][i, k] = ECB_SYNTHESIS
            mods["density"][i, k] = rng.uniform(0.5, 2.0)
            mods["state"][i, k, 0] = rng.uniform(0.0, 3.0)
            if rng.random() < 0.5:
                mods["state"][i, k, 1] = rng.uniform(0.005, 0.05)
                mods["state"][i, k, 2] = rng.uniform(0.001, 0.02)
            segs["modifier_count"][i] = k + 1
    return net


def gpcr_ligands(net: Network, tick: int, seed: int = 5) -> np.ndarray:
    """The tick's extracellular ligand injections of gpcr_network: every 4th tick each of the 15
    transmitters onto a few segments (amounts 0.05-0.5)."""
    if tick % 4:
        return np.zeros(0, LIG_INJ_DT)
    rng = np.random.default_rng(seed * 100003 + tick)
    n = len(net.segments)
    out = np.zeros(15 * 3, LIG_INJ_DT)
    out["nt_idx"] = np.repeat(np.arange(15), 3)
    out["segment_id"] = net.segments["id"][rng.integers(0, n, len(out))]
    out["amount"] = rng.uniform(0.05, 0.5, len(out)).astype(np.float32)
    return out


def segment_receptor_network(seed: int, n_neurons: int = 12, seg_per_neuron: int = 6) -> Network:
    """Chain neurons with AMPA+NMDA/GABA-A synapses whose dendrites also carry ionotropic
    receptors as segment modifiers (AMPA and NMDA on every third segment, GABA-A on the next):
    step 1 runs them on the segment's own transmitters (the NMDA carriers' glutamate bath and
    the previous tick's synaptic deposits)."""
    net = build_network(n_neurons, seg_per_neuron, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], [CAV21],
                        3.0, "ampa_nmda_gaba", 4, seed)
    segs = net.segments
    local = np.arange(len(segs)) % seg_per_neuron
    mods = segs["modifiers"]
    for i in np.nonzero(local > 0)[0]:
        k = int(segs["modifier_count"][i])
        if local[i] % 3 == 1:
            mods["type"][i, k], mods["density"][i, k] = AMPA, 200.0
            mods["type"][i, k + 1], mods["density"][i, k + 1] = NMDA, 40.0
            mods["state"][i, k + 1, 2] = 1.0   # CDI gate open
            segs["modifier_count"][i] = k + 2
        elif local[i] % 3 == 2:
            mods["type"][i, k], mods["density"][i, k] = GABAA, 150.0
            segs["modifier_count"][i] = k + 1
    return net


def config1(seed: int = 0xB10C0001) -> Network:
    """64 neurons x (soma + 7-segment chain); Nav1.2/Kv-DR/K2P/NALCN/pump everywhere,
    Cav2.1 on somas; 256 AMPA synapses (soma -> uniform segment), delays 1-5 ticks."""
    return build_network(64, 8, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], [CAV21], 4.0, "ampa", 5, seed)


CONFIG4_TYPES = [NAV12, KVDR, K2P, NALCN, PUMP, CAV21]


def config4_neurons(n_networks: int = 512, seed: int = 0xB10C0004, lo: float = 2_000.0,
                    hi: float = 200_000.0) -> np.ndarray:
    """Neuron counts of the batched sweep: segment counts log-uniform in [lo, hi], 8 segments per
    neuron (mean ~43k segments, ~22M in total at 512 networks)."""
    rng = np.random.default_rng(seed)
    seg = np.exp(rng.uniform(np.log(lo), np.log(hi), n_networks))
    return np.maximum(1, np.round(seg / 8.0)).astype(np.int64)


def config4_network(k: int, n_neurons: int, seed: int = 0xB10C0004) -> Network:
    """Network k of the batched sweep: built like the 64-neuron chain network (soma + 7-segment
    chain, 0.5 AMPA synapses per segment, delays 1-5 ticks) with its own seed, and its own
    parameter set: every channel type's g_max scaled by one U(0.5, 2) factor."""
    s = seed + 1 + k
    net = build_network(int(n_neurons), 8, "chain", [NAV12, KVDR, K2P, NALCN, PUMP], [CAV21], 4.0, "ampa", 5, s)
    rng = np.random.default_rng(s ^ 0x5EED)
    dens = net.segments["modifiers"]["density"]
    types = net.segments["modifiers"]["type"]
    for t in CONFIG4_TYPES:
        dens[types == t] *= rng.uniform(0.5, 2.0)
    return net


def soma_current(net: Network, amp_pA: float) -> np.ndarray:
    """A constant current clamp on every soma for the whole run."""
    somas = np.nonzero(net.segments["segment_type"] == 1)[0]
    cur = np.zeros(len(somas), CURRENT_STIM_DT)
    cur["segment_id"] = net.segments["id"][somas]
    cur["amplitude_pA"] = amp_pA
    cur["duration_ms"] = 1e9
    return cur


CONFIG2_TYPES = [NAV12, NAV16, KV11, KV31, KV7, CAV21, HCN1, SK, BK, K2P, PUMP, NALCN]


def config2(n_neurons: int = 100_000, seed: int = 0xB10C0002, syn_per_neuron: float = 100.0) -> Network:
    """n_neurons x (soma + 31-node binary tree) = 32 segments each, the 12-modifier set with
    log-normal (sigma 0.3) densities, fan-in syn_per_neuron from uniform presynaptic somas,
    70% AMPA+NMDA / 30% GABA-A, delays 1-20 ticks."""
    return build_network(n_neurons, 32, "tree", CONFIG2_TYPES, [], syn_per_neuron, "ampa_nmda_gaba", 20,
                         seed, lognormal_sigma=0.3)


def config3(n_cables: int = 10_000, seg_per_cable: int = 100, seed: int = 0xB10C0003) -> Network:
    """Passive unbranched cables with linear +-20% concentration ramps (electrodiffusion stress)."""
    net = build_network(n_cables, seg_per_cable, "chain", [], [], 0.0, "ampa", 1, seed)
    segs = net.segments
    segs["modifier_count"] = 0
    frac = np.tile(np.linspace(-0.2, 0.2, seg_per_cable), n_cables)
    segs["intra_ions"][:, :7] = INTRA_IONS[None, :7] * (1.0 + frac[:, None])
    segs["extra_ions"][:, :7] = EXTRA_IONS[None, :7] * (1.0 + frac[:, None])
    segs["ATP_level"] = (1.0 + frac).astype(np.float32)
    segs["temperature"] = (37.0 * (1.0 + 0.01 * frac)).astype(np.float32)
    return net


def config5_cell(seed: int, n_seg: int = 200, gmax_scale: np.ndarray | None = None) -> Network:
    """One deterministic cell: soma + branched dendrite (binary tree), Nav1.2/Kv-DR/K2P/NALCN/HCN1."""
    net = build_network(1, n_seg, "tree", [NAV12, KVDR, K2P, NALCN, HCN1], [], 0.0, "ampa", 1, seed)
    if gmax_scale is not None:
        net.segments["modifiers"]["density"][:, :5] *= gmax_scale
    return net


def poisson_soma_pulses(net: Network, rate_hz: float, ticks: int, dt: float, amp_pA: float,
                        seed: int) -> list[np.ndarray]:
    """Per-tick lists of 1-tick current pulses (CURRENT_STIM_DT) at rate_hz per soma."""
    rng = np.random.default_rng(seed)
    somas = np.nonzero(net.segments["segment_type"] == 1)[0]
    p = rate_hz * dt * 1e-3
    out = []
    for t in range(ticks):
        hit = somas[rng.random(len(somas)) < p]
        st = np.zeros(len(hit), CURRENT_STIM_DT)
        st["segment_id"] = net.segments["id"][hit]
        st["amplitude_pA"] = amp_pA
        st["duration_ms"] = dt
        st["start_tick"] = t
        out.append(st)
    return out
