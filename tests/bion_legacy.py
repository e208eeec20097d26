"""Record layouts of BION format versions 3-13 (include/bion_format.h:30-450,
src/bion_loader_host.cpp:107-340) as numpy dtypes, and a writer of synthetic files in them.

Each layout is the one the reference loader reads for that version (for v4-v6 segments that is
a 1692-byte record whose 16 modifiers have today's 88-byte layout). Field values are random
but valid: finite floats, counts within their limits, indices within the network.
"""
from __future__ import annotations

import os

import numpy as np

from bioneuroncore_amd import bion

MOD_V3 = np.dtype([("type", "u1"), ("density", "<f4"), ("state", "<f4", 8), ("active_vars", "u1"),
                   ("flags", "u1"), ("res", "u1", 2)])
MOD_F = np.dtype([("type", "u1"), ("pad", "u1", 3), ("density", "<f4"), ("state", "<f4", 8),
                  ("active_vars", "u1"), ("flags", "u1"), ("res", "u1", 2)])
_HEAD = [("id", "<u4"), ("neuron_id", "<u4"), ("position", "<f4", 3), ("direction", "<f4", 3),
         ("length", "<f4"), ("radius", "<f4")]
_TAIL = [("modifier_count", "u1"), ("segment_type", "u1"), ("res1", "u1", 2), ("ATP_level", "<f4"),
         ("temperature", "<f4"), ("I_buffer", "<f4"), ("transcription_signal", "<f4"),
         ("intra_connection_ids", "<u4", 16), ("intra_connection_count", "u1"), ("res2", "u1", 3)]
SEG_V3 = np.dtype([("id", "<u4"), ("neuron_id", "<u4"), ("position", "<f4", 3), ("length", "<f4"),
                   ("radius", "<f4"), ("V_m", "<f4"), ("intra_ions", "<f4", 4), ("extra_ions", "<f4", 4),
                   ("neurotransmitter", "<f4", 12), ("modifiers", MOD_V3, 16), ("modifier_count", "u1"),
                   ("res", "u1", 3), ("ATP_level", "<f4"), ("temperature", "<f4"), ("I_buffer", "<f4"),
                   ("intra_connection_ids", "<u4", 16), ("intra_connection_count", "u1"), ("res2", "u1", 3)])
SEG_V4 = np.dtype(_HEAD + [("V_m", "<f4"), ("intra_ions", "<f4", 5), ("extra_ions", "<f4", 5),
                           ("intra_nt", "<f4", 14), ("neurotransmitter", "<f4", 14),
                           ("modifiers", bion.MODIFIER_DT, 16)] + _TAIL)
SEG_V7 = np.dtype(_HEAD + [("V_m", "<f4"), ("intra_ions", "<f4", 5), ("extra_ions", "<f4", 5),
                           ("intra_nt", "<f4", 14), ("neurotransmitter", "<f4", 14),
                           ("modifiers", MOD_F, 32)] + _TAIL)
SEG_V10ER = np.dtype(SEG_V7.descr + [("ER_Ca2", "<f4")])
SEG_V10 = np.dtype(_HEAD + [("V_m", "<f4"), ("intra_ions", "<f4", 5), ("extra_ions", "<f4", 5),
                            ("intra_nt", "<f4", 15), ("neurotransmitter", "<f4", 15),
                            ("modifiers", MOD_F, 32)] + _TAIL + [("ER_Ca2", "<f4")])
SEG_V13 = np.dtype(_HEAD + [("V_m", "<f8"), ("intra_ions", "<f8", 8), ("extra_ions", "<f8", 8),
                            ("intra_nt", "<f4", 15), ("neurotransmitter", "<f4", 15),
                            ("modifiers", bion.MODIFIER_DT, 32), ("modifier_count", "u1"),
                            ("segment_type", "u1"), ("res1", "u1", 2), ("ATP_level", "<f4"),
                            ("temperature", "<f4"), ("I_buffer", "<f8"), ("transcription_signal", "<f4"),
                            ("intra_connection_ids", "<u4", 16), ("intra_connection_count", "u1"),
                            ("res2", "u1", 3), ("ER_Ca2", "<f4")])
CONN_V8 = np.dtype([("segment_a", "<u4"), ("segment_b", "<u4"), ("coupling_resistance", "<f4")])
_SYN_HEAD = [("id", "<u4"), ("pre_segment", "<u4"), ("post_segment", "<u4"), ("type", "u1"),
             ("res", "u1", 3), ("weight", "<f4"), ("delay", "<f4"), ("pre_area", "<f4"), ("post_area", "<f4")]
SYN_V7 = np.dtype(_SYN_HEAD + [("pre_autoreceptors", MOD_F, 8), ("pre_autoreceptor_count", "u1"),
                               ("res2", "u1", 3), ("post_receptors", MOD_F, 8), ("post_receptor_count", "u1"),
                               ("res3", "u1", 3), ("release_probability", "<f4"), ("release_amount", "<f4"),
                               ("cleft_neurotransmitter", "<f4", 14), ("facilitation", "<f4"),
                               ("depression", "<f4"), ("tau_clearance", "<f4"), ("refractory", "u1"),
                               ("res4", "u1", 3), ("RRP", "<f4"), ("RRP_max", "<f4"), ("RRP_recovery", "<f4")])
SYN_V13 = np.dtype(_SYN_HEAD + [("pre_autoreceptors", bion.MODIFIER_DT, 8), ("pre_autoreceptor_count", "u1"),
                                ("res2", "u1", 3), ("post_receptors", bion.MODIFIER_DT, 8),
                                ("post_receptor_count", "u1"), ("res3", "u1", 3), ("release_probability", "<f4"),
                                ("release_amount", "<f4"), ("cleft_neurotransmitter", "<f4", 15),
                                ("facilitation", "<f4"), ("depression", "<f4"), ("tau_clearance", "<f4"),
                                ("refractory", "u1"), ("res4", "u1", 3), ("RRP", "<f4"), ("RRP_max", "<f4"),
                                ("RRP_recovery", "<f4"), ("bouton_Ca", "<f4"), ("spine_Ca", "<f4"),
                                ("spine_camkii", "<f4"), ("spine_calcineurin", "<f4")])
_NUC = [("neuron_id", "<u4"), ("enabled", "u1"), ("consolidation_mode", "u1"), ("pad", "u1", 2),
        ("consolidation_interval_ticks", "<u4"), ("last_consolidation_tick", "<u4")]
_NUC_T = [("protein_synthesis_rate", "<f4"), ("pcreb_level", "<f4"), ("erk_level", "<f4"),
          ("nfkb_level", "<f4"), ("mtorc1_level", "<f4"), ("amino_acid_level", "<f4"), ("res", "u1", 4)]
NUC_V5 = np.dtype(_NUC + [("gene_activity", "<f4", 20), ("protein_level", "<f4", 20)] + _NUC_T)
NUC_V46 = np.dtype(_NUC + [("gene_activity", "<f4", 46), ("protein_level", "<f4", 46)] + _NUC_T)
NUC = np.dtype(_NUC + [("gene_activity", "<f4", 57), ("protein_level", "<f4", 57)] + _NUC_T)

SIZES = {"SEG_V3": (SEG_V3, 852), "SEG_V4": (SEG_V4, 1692), "SEG_V7": (SEG_V7, 1692),
         "SEG_V10ER": (SEG_V10ER, 1696), "SEG_V10": (SEG_V10, 1704), "SEG_V13": (SEG_V13, 3208),
         "CONN_V8": (CONN_V8, 12), "SYN_V7": (SYN_V7, 836), "SYN_V13": (SYN_V13, 1560),
         "NUC_V5": (NUC_V5, 204), "NUC_V46": (NUC_V46, 412), "NUC": (NUC, 500)}
for _n, (_d, _s) in SIZES.items():
    assert _d.itemsize == _s, (_n, _d.itemsize, _s)


def fill(dt: np.dtype, n: int, rng) -> np.ndarray:
    """n records of dt: finite floats in [-100, 100], random integers and bytes."""
    a = np.zeros(n, dt)
    _fill_fields(a, rng)
    return a


def _fill_fields(a, rng):
    for name in a.dtype.names:
        f = a[name]
        base = f.dtype.base if f.dtype.subdtype is None else f.dtype.subdtype[0]
        if base.names:
            _fill_fields(f.reshape(-1) if f.ndim else f, rng)
            continue
        if base.kind == "f":
            a[name] = rng.uniform(-100, 100, f.shape).astype(base)
        elif base.kind in "ui":
            info = np.iinfo(base)
            a[name] = rng.integers(0, min(info.max, 2**31 - 1), f.shape, dtype=np.int64).astype(base)
        else:
            a[name] = rng.integers(0, 256, f.shape, dtype=np.int64).astype(base)


def valid_modifier_counts(seg, limit, rng):
    seg["modifier_count"] = rng.integers(0, limit + 1, len(seg))


def network_parts(n_neurons, seg_per, rng):
    """Neurons (each owning seg_per segments) and chain connections inside each neuron."""
    neu = fill(bion.NEURON_DT, n_neurons, rng)
    neu["segment_start"] = np.arange(n_neurons) * seg_per
    neu["segment_count"] = seg_per
    a = [(k * seg_per + i, k * seg_per + i + 1) for k in range(n_neurons) for i in range(seg_per - 1)]
    return neu, np.array(a, np.uint32).reshape(-1, 2)


def header(version: int, n_neurons, n_seg, n_conn, n_syn, n_nuc=0, sizes=None) -> bytes:
    h = np.zeros(1, bion.HEADER_DT)
    h["magic"] = 0x42494F4E
    h["version"] = version
    h["mode"] = 2
    h["diffusion_type"] = 1
    h["num_neurons"], h["num_segments"] = n_neurons, n_seg
    h["num_intra_connections"], h["num_synapses"] = n_conn, n_syn
    h["grid_dim_x"], h["grid_dim_y"], h["grid_dim_z"] = 3, 4, 5
    h["voxel_size"] = 2.5
    h["timestamp"] = 123456789
    h["build_id"] = b"legacy-test"
    h["num_nuclei"] = n_nuc
    h["nucleus_version"] = 1
    for k, v in (sizes or {}).items():
        h["file_sz_" + k] = v
    raw = h.tobytes()
    return raw[:68] if version < 5 else raw[:84] if version < 10 else raw


# ---------------------------------------------------------------- seeded files of every version
def _synapses_v14(n, n_seg, rng):
    y = fill(bion.SYNAPSE_DT, n, rng)
    y["pre_segment"] = rng.integers(0, n_seg, n)
    y["post_segment"] = rng.integers(0, n_seg, n)
    y["type"] = rng.integers(0, 3, n)
    y["pre_autoreceptor_count"] = rng.integers(0, 9, n)
    y["post_receptor_count"] = rng.integers(0, 9, n)
    return y


def _legacy_syn(dt, n, n_seg, rng, counts=8):
    y = fill(dt, n, rng)
    y["pre_segment"] = rng.integers(0, n_seg, n)
    y["post_segment"] = rng.integers(0, n_seg, n)
    y["type"] = rng.integers(0, 3, n)
    y["pre_autoreceptor_count"] = rng.integers(0, counts + 1, n)
    y["post_receptor_count"] = rng.integers(0, counts + 1, n)
    return y


def build(tmp, name, version, seg="auto", syn="auto", nuc="auto", conn="auto", n_neurons=3, seg_per=7,
          n_syn=40, seed=0, mutate=None):
    """Writes one file; returns its path. seg/syn/nuc: record size on disk (v10+) or 'auto'."""
    rng = np.random.default_rng(seed)
    n_seg = max(n_neurons, 1) * seg_per
    neu, pairs = network_parts(max(n_neurons, 1), seg_per, rng)
    neu = neu[:n_neurons]
    n_conn = len(pairs)
    sizes = {}
    # segments
    if version == 3:
        s = fill(SEG_V3, n_seg, rng)
        valid_modifier_counts(s, 16, rng)
    elif version <= 6:
        s = fill(SEG_V4, n_seg, rng)
        valid_modifier_counts(s, 16, rng)
    elif version <= 9:
        s = fill(SEG_V7, n_seg, rng)
        valid_modifier_counts(s, 32, rng)
    else:
        dts = {"auto": bion.SEGMENT_DT, 0: bion.SEGMENT_DT, 3216: bion.SEGMENT_DT, 1692: SEG_V7,
               1696: SEG_V10ER, 1704: SEG_V10, 3208: SEG_V13}
        if seg in dts:
            s = fill(dts[seg], n_seg, rng)
            valid_modifier_counts(s, 32, rng)
            sizes["segment"] = 3216 if seg == "auto" else seg
        else:   # an unknown size: raw v14 prefix or v14 record plus a tail
            full = fill(bion.SEGMENT_DT, n_seg, rng)
            valid_modifier_counts(full, 32, rng)
            raw = full.view(np.uint8).reshape(n_seg, -1)
            if seg < 3216:
                s = raw[:, :seg].copy()
            else:
                s = np.concatenate([raw, rng.integers(0, 256, (n_seg, seg - 3216), dtype=np.uint8)], axis=1)
            sizes["segment"] = seg
    # connections
    if version <= 8:
        c = fill(CONN_V8, n_conn, rng)
        c["coupling_resistance"][::3] = 0.0
    else:
        c = fill(bion.CONN_DT, n_conn, rng)
    c["segment_a"], c["segment_b"] = pairs[:, 0], pairs[:, 1]
    # synapses
    if version <= 6:
        y = _synapses_v14(n_syn, n_seg, rng).view(np.uint8).reshape(n_syn, 1568)[:, :824].copy()
    elif version <= 9:
        y = _legacy_syn(SYN_V7, n_syn, n_seg, rng)
    else:
        if syn in ("auto", 0, 1568):
            y = _synapses_v14(n_syn, n_seg, rng)
        elif syn == 836:
            y = _legacy_syn(SYN_V7, n_syn, n_seg, rng)
        elif syn == 1560:
            y = _legacy_syn(SYN_V13, n_syn, n_seg, rng)
        else:
            y = _synapses_v14(n_syn, n_seg, rng).view(np.uint8).reshape(n_syn, 1568)[:, :syn].copy()
        sizes["synapse"] = 1568 if syn == "auto" else syn
    # nuclei
    n_nuc = 0
    nucs = b""
    if version == 5:
        n_nuc = n_neurons
        nucs = fill(NUC_V5, n_nuc, rng).tobytes()
    elif version == 6:
        n_nuc = n_neurons
        nucs = rng.integers(0, 255, (n_nuc, 332), dtype=np.uint8).tobytes()
    elif version >= 7 and nuc != "none":
        n_nuc = n_neurons
        if nuc == 412:
            nucs = fill(NUC_V46, n_nuc, rng).tobytes()
            sizes["nucleus"] = 412
        else:
            nucs = fill(NUC, n_nuc, rng).tobytes()
            if version >= 10:
                sizes["nucleus"] = 500
    if version >= 10:
        sizes.update(intraconnection=16, neuron=24, modifier=88)
    parts = [header(version, n_neurons, n_seg, n_conn, n_syn, n_nuc, sizes), neu.tobytes(), s.tobytes(),
             c.tobytes(), y.tobytes(), nucs]
    if mutate:
        parts = mutate(parts)
    p = os.path.join(tmp, name + ".bion")
    with open(p, "wb") as f:
        f.write(b"".join(parts))
    return p


def cases(tmp):
    c = {}
    for v in (3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 14):
        c[f"v{v}"] = (build(tmp, f"v{v}", v, seed=v), v)
    c["v4_nosyn"] = (build(tmp, "v4_nosyn", 4, n_syn=0, seed=40), 4)
    for sz in (0, 1692, 1696, 1704, 3208, 3000, 3400):
        c[f"v10_seg{sz}"] = (build(tmp, f"v10_seg{sz}", 10, seg=sz, seed=100 + sz), 10)
    for sz in (0, 836, 1560, 840):
        c[f"v10_syn{sz}"] = (build(tmp, f"v10_syn{sz}", 10, syn=sz, seed=200 + sz), 10)
    c["v10_nuc412"] = (build(tmp, "v10_nuc412", 10, nuc=412, seed=300), 10)
    c["v14_nonuc"] = (build(tmp, "v14_nonuc", 14, nuc="none", seed=301), 14)
    c["v14_noneuron"] = (build(tmp, "v14_noneuron", 14, n_neurons=0, seg_per=5, n_syn=0, seed=302), 14)
    return c


def defined(version, arrays):
    """Byte masks of the fields the reference loader writes (neurons, segments, connections,
    synapses, nuclei as uint8 arrays): reserved bytes of v3-v8 connections, v3-v6 synapses past
    824 bytes (except the vesicle pool) and v6 nuclei past 332 bytes stay unwritten."""
    neu, seg, con, syn, nuc = arrays
    mc = np.ones_like(con, bool)
    if version <= 8 and len(con):
        m = mc.reshape(-1, 16)
        m[:, 9:12] = False   # reserved[3]
    my = np.ones_like(syn, bool)
    if version <= 6 and len(syn):
        m = my.reshape(-1, 1568)
        m[:, 824:] = False
        for f in ("RRP", "RRP_max", "RRP_recovery"):
            o = bion.SYNAPSE_DT.fields[f][1]
            m[:, o:o + 4] = True
    mn = np.ones_like(nuc, bool)
    if version == 6 and len(nuc):
        m = mn.reshape(-1, bion.NUCLEUS_SIZE)
        m[:, 332:] = False
        o = 16 + 57 * 4 + 36 * 4   # protein_level[36..46), zeroed by the loader
        m[:, o:o + 40] = True
    return np.ones_like(neu, bool), np.ones_like(seg, bool), mc, my, mn


def error_cases(tmp):
    """Files each of which the loader must refuse, by the case's name."""
    def bad_magic(p):
        p[0] = b"XXXX" + p[0][4:]
        return p

    def version(v):
        def f(p):
            h = bytearray(p[0])
            h[4:8] = int(v).to_bytes(4, "little")
            p[0] = bytes(h)
            return p
        return f

    def zero_segs(p):
        h = bytearray(p[0])
        h[16:20] = (0).to_bytes(4, "little")
        p[0] = bytes(h)
        return p

    def truncate(k):
        def f(p):
            p[k] = p[k][:-5]
            return p[:k + 1]
        return f

    def poke(k, off, val):
        def f(p):
            b = bytearray(p[k])
            b[off:off + len(val)] = val
            p[k] = bytes(b)
            return p
        return f

    seg_sz, syn_sz = 3216, 1568
    return {
        "bad_magic": build(tmp, "e_magic", 14, mutate=bad_magic),
        "version2": build(tmp, "e_v2", 14, mutate=version(2)),
        "version15": build(tmp, "e_v15", 14, mutate=version(15)),
        "zero_segments": build(tmp, "e_zero", 14, mutate=zero_segs),
        "short_header": build(tmp, "e_hdr", 14, mutate=lambda p: [p[0][:50]]),
        "trunc_neurons": build(tmp, "e_tn", 14, mutate=truncate(1)),
        "trunc_segments": build(tmp, "e_ts", 14, mutate=truncate(2)),
        "trunc_segments_v7": build(tmp, "e_ts7", 7, mutate=truncate(2)),
        "trunc_connections": build(tmp, "e_tc", 14, mutate=truncate(3)),
        "trunc_synapses": build(tmp, "e_ty", 14, mutate=truncate(4)),
        "trunc_nuclei_v5": build(tmp, "e_tu", 5, mutate=truncate(5)),
        "conn_range": build(tmp, "e_cr", 14, mutate=poke(3, 0, (99999).to_bytes(4, "little"))),
        "syn_range": build(tmp, "e_yr", 14, mutate=poke(4, 8, (99999).to_bytes(4, "little"))),
        "syn_type": build(tmp, "e_yt", 14, mutate=poke(4, 12, bytes([3]))),
        "syn_receptors": build(tmp, "e_yc", 14, mutate=poke(4, bion.SYNAPSE_DT.fields["post_receptor_count"][1], bytes([9]))),
        "neuron_range": build(tmp, "e_nr", 14, mutate=poke(1, 8, (999).to_bytes(4, "little"))),
        "modcount": build(tmp, "e_mc", 14, mutate=poke(2, bion.SEGMENT_DT.fields["modifier_count"][1], bytes([33]))),
    }
