"""BION files of format versions 3-14 load into the same records as with the reference's own
loader (bion_load_host, src/bion_loader_host.cpp:26-425, compiled from the reference sources
into oracle/_ref), byte for byte, and v14 saves equal bion_save's (timestamp masked).

Synthetic files in each version's layout (tests/bion_legacy.py) cover: v3 (852-byte segments,
41-byte modifiers), v4-v6 (1692-byte segments), v5/v6 nuclei (204/332 bytes), v7-v9 (float
modifiers, 5 ions, 14 transmitters, 836-byte synapses), 12-byte connections up to v8, and from
v10 every segment/synapse/nucleus size the header's file_sz_* fields select (current, 1692, 1696,
1704, 3208, shorter and longer unknown sizes; 836, 1560, 840; 412, 500); plus every error code.
Bytes the reference leaves unwritten (reserved bytes of v3-v8 connections, v3-v6 synapses past
the first 824 bytes except the vesicle pool, v6 nuclei past 332 bytes) are not compared; the
reference loads run in a fresh interpreter, where those allocations come zeroed from the kernel.
"""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import bion_legacy as L
import harness as H

pytestmark = pytest.mark.skipif(not H.ref_available(), reason="oracle/_ref (the compiled reference) not built")

SYN_OFF = bion.SYNAPSE_DT.fields
NUC_SIZE = bion.NUCLEUS_SIZE

_REF_RUNNER = r"""
import ctypes as C, json, sys
import numpy as np
class BD(C.Structure):
    _fields_ = [("header", C.c_uint8 * 100), ("neurons", C.c_void_p), ("segments", C.c_void_p),
                ("connections", C.c_void_p), ("synapses", C.c_void_p), ("nuclei", C.c_void_p),
                ("num_neurons", C.c_uint32), ("num_segments", C.c_uint32), ("num_connections", C.c_uint32),
                ("num_synapses", C.c_uint32), ("num_nuclei", C.c_uint32)]
lib = C.CDLL(sys.argv[1])
load = lib._Z14bion_load_hostPKcP8BionData
load.argtypes = [C.c_char_p, C.c_void_p]
free = lib._Z16bion_unload_hostP8BionData
free.argtypes = [C.c_void_p]
out = {}
for path in json.loads(sys.argv[2]):
    b = BD()
    rc = load(path.encode(), C.byref(b))
    rec = {"rc": rc}
    if rc == 0:
        def grab(p, n, sz):
            return bytes((C.c_uint8 * (n * sz)).from_address(p)) if n and p else b""
        rec.update(header=bytes(b.header), n=[b.num_neurons, b.num_segments, b.num_connections,
                                              b.num_synapses, b.num_nuclei])
        np.save(path + ".ref.npy", np.frombuffer(
            grab(b.neurons, b.num_neurons, 24) + grab(b.segments, b.num_segments, 3216) +
            grab(b.connections, b.num_connections, 16) + grab(b.synapses, b.num_synapses, 1568) +
            grab(b.nuclei, b.num_nuclei, 500), np.uint8))
        rec["header"] = list(rec["header"])
        free(C.byref(b))
    out[path] = rec
print(json.dumps(out))
"""


def _reference(paths):
    r = subprocess.run([sys.executable, "-c", _REF_RUNNER, H.REF_SO, json.dumps(paths)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def _mine(path):
    lib = backend.load_library()
    lib.bchip_read_bion.argtypes = [C.c_char_p, C.c_void_p]
    lib.bchip_free_bion.argtypes = [C.c_void_p]
    b = backend.BionDataC()
    rc = lib.bchip_read_bion(path.encode(), C.byref(b))
    if rc != 0:
        return rc, None
    def grab(p, n, sz):
        return bytes((C.c_uint8 * (n * sz)).from_address(p)) if n and p else b""
    parts = dict(header=bytes(b.header), n=[b.num_neurons, b.num_segments, b.num_connections,
                                            b.num_synapses, b.num_nuclei],
                 neurons=grab(b.neurons, b.num_neurons, 24), segments=grab(b.segments, b.num_segments, 3216),
                 connections=grab(b.connections, b.num_connections, 16),
                 synapses=grab(b.synapses, b.num_synapses, 1568), nuclei=grab(b.nuclei, b.num_nuclei, 500))
    lib.bchip_free_bion(C.byref(b))
    return 0, parts


# ---------------------------------------------------------------- file builders
def _synapses_v14(n, n_seg, rng):
    y = L.fill(bion.SYNAPSE_DT, n, rng)
    y["pre_segment"] = rng.integers(0, n_seg, n)
    y["post_segment"] = rng.integers(0, n_seg, n)
    y["type"] = rng.integers(0, 3, n)
    y["pre_autoreceptor_count"] = rng.integers(0, 9, n)
    y["post_receptor_count"] = rng.integers(0, 9, n)
    return y


def _legacy_syn(dt, n, n_seg, rng, counts=8):
    y = L.fill(dt, n, rng)
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
    neu, pairs = L.network_parts(max(n_neurons, 1), seg_per, rng)
    neu = neu[:n_neurons]
    n_conn = len(pairs)
    sizes = {}
    # segments
    if version == 3:
        s = L.fill(L.SEG_V3, n_seg, rng)
        L.valid_modifier_counts(s, 16, rng)
    elif version <= 6:
        s = L.fill(L.SEG_V4, n_seg, rng)
        L.valid_modifier_counts(s, 16, rng)
    elif version <= 9:
        s = L.fill(L.SEG_V7, n_seg, rng)
        L.valid_modifier_counts(s, 32, rng)
    else:
        dts = {"auto": bion.SEGMENT_DT, 0: bion.SEGMENT_DT, 3216: bion.SEGMENT_DT, 1692: L.SEG_V7,
               1696: L.SEG_V10ER, 1704: L.SEG_V10, 3208: L.SEG_V13}
        if seg in dts:
            s = L.fill(dts[seg], n_seg, rng)
            L.valid_modifier_counts(s, 32, rng)
            sizes["segment"] = 3216 if seg == "auto" else seg
        else:   # an unknown size: raw v14 prefix or v14 record plus a tail
            full = L.fill(bion.SEGMENT_DT, n_seg, rng)
            L.valid_modifier_counts(full, 32, rng)
            raw = full.view(np.uint8).reshape(n_seg, -1)
            if seg < 3216:
                s = raw[:, :seg].copy()
            else:
                s = np.concatenate([raw, rng.integers(0, 256, (n_seg, seg - 3216), dtype=np.uint8)], axis=1)
            sizes["segment"] = seg
    # connections
    if version <= 8:
        c = L.fill(L.CONN_V8, n_conn, rng)
        c["coupling_resistance"][::3] = 0.0
    else:
        c = L.fill(bion.CONN_DT, n_conn, rng)
    c["segment_a"], c["segment_b"] = pairs[:, 0], pairs[:, 1]
    # synapses
    if version <= 6:
        y = _synapses_v14(n_syn, n_seg, rng).view(np.uint8).reshape(n_syn, 1568)[:, :824].copy()
    elif version <= 9:
        y = _legacy_syn(L.SYN_V7, n_syn, n_seg, rng)
    else:
        if syn in ("auto", 0, 1568):
            y = _synapses_v14(n_syn, n_seg, rng)
        elif syn == 836:
            y = _legacy_syn(L.SYN_V7, n_syn, n_seg, rng)
        elif syn == 1560:
            y = _legacy_syn(L.SYN_V13, n_syn, n_seg, rng)
        else:
            y = _synapses_v14(n_syn, n_seg, rng).view(np.uint8).reshape(n_syn, 1568)[:, :syn].copy()
        sizes["synapse"] = 1568 if syn == "auto" else syn
    # nuclei
    n_nuc = 0
    nucs = b""
    if version == 5:
        n_nuc = n_neurons
        nucs = L.fill(L.NUC_V5, n_nuc, rng).tobytes()
    elif version == 6:
        n_nuc = n_neurons
        nucs = rng.integers(0, 255, (n_nuc, 332), dtype=np.uint8).tobytes()
    elif version >= 7 and nuc != "none":
        n_nuc = n_neurons
        if nuc == 412:
            nucs = L.fill(L.NUC_V46, n_nuc, rng).tobytes()
            sizes["nucleus"] = 412
        else:
            nucs = L.fill(L.NUC, n_nuc, rng).tobytes()
            if version >= 10:
                sizes["nucleus"] = 500
    if version >= 10:
        sizes.update(intraconnection=16, neuron=24, modifier=88)
    parts = [L.header(version, n_neurons, n_seg, n_conn, n_syn, n_nuc, sizes), neu.tobytes(), s.tobytes(),
             c.tobytes(), y.tobytes(), nucs]
    if mutate:
        parts = mutate(parts)
    p = os.path.join(tmp, name + ".bion")
    with open(p, "wb") as f:
        f.write(b"".join(parts))
    return p


def _cases(tmp):
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


def _error_cases(tmp):
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
        "syn_receptors": build(tmp, "e_yc", 14, mutate=poke(4, SYN_OFF["post_receptor_count"][1], bytes([9]))),
        "neuron_range": build(tmp, "e_nr", 14, mutate=poke(1, 8, (999).to_bytes(4, "little"))),
        "modcount": build(tmp, "e_mc", 14, mutate=poke(2, bion.SEGMENT_DT.fields["modifier_count"][1], bytes([33]))),
    }


def _defined(name, version, arrays):
    """Byte masks of the fields the reference writes."""
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
            o = SYN_OFF[f][1]
            m[:, o:o + 4] = True
    mn = np.ones_like(nuc, bool)
    if version == 6 and len(nuc):
        m = mn.reshape(-1, NUC_SIZE)
        m[:, 332:] = False
        o = 16 + 57 * 4 + 36 * 4   # protein_level[36..46), zeroed by the loader
        m[:, o:o + 40] = True
    return np.ones_like(neu, bool), np.ones_like(seg, bool), mc, my, mn


def test_every_version_loads_as_the_reference_loads(tmp_path):
    cases = _cases(str(tmp_path))
    ref = _reference([p for p, _ in cases.values()])
    for name, (p, v) in cases.items():
        r = ref[p]
        rc, mine = _mine(p)
        assert rc == r["rc"] == 0, (name, rc, r["rc"])
        assert mine["n"] == r["n"], (name, mine["n"], r["n"])
        assert mine["header"] == bytes(r["header"]), name
        rb = np.load(p + ".ref.npy")
        n = r["n"]
        sizes = [n[0] * 24, n[1] * 3216, n[2] * 16, n[3] * 1568, n[4] * 500]
        cuts = np.cumsum([0] + sizes)
        ra = [rb[cuts[i]:cuts[i + 1]] for i in range(5)]
        ma = [np.frombuffer(mine[k], np.uint8) for k in ("neurons", "segments", "connections", "synapses", "nuclei")]
        masks = _defined(name, v, ra)
        for part, a, b, m in zip(("neurons", "segments", "connections", "synapses", "nuclei"), ra, ma, masks):
            assert len(a) == len(b), (name, part)
            bad = np.nonzero((a != b) & m)[0]
            assert len(bad) == 0, (name, part, int(bad[0]), len(bad))
        print(f"{name}: {n[1]} segments, {n[3]} synapses, {n[4]} nuclei equal", flush=True)


def test_error_codes_equal_the_reference(tmp_path):
    cases = _error_cases(str(tmp_path))
    ref = _reference(list(cases.values()))
    for name, p in cases.items():
        rc, _ = _mine(p)
        assert rc == ref[p]["rc"] != 0, (name, rc, ref[p]["rc"])
    rc, _ = _mine(str(tmp_path / "missing.bion"))
    assert rc == -1


def test_save_equals_bion_save(tmp_path):
    """A load + save through the library and through the reference give the same file."""
    p = build(str(tmp_path), "save_src", 9, seed=9)
    lib = backend.load_library()
    lib.bchip_read_bion.argtypes = [C.c_char_p, C.c_void_p]
    lib.bchip_write_bion.argtypes = [C.c_char_p, C.c_void_p]
    lib.bchip_free_bion.argtypes = [C.c_void_p]
    b = backend.BionDataC()
    assert lib.bchip_read_bion(p.encode(), C.byref(b)) == 0
    mine = str(tmp_path / "mine.bion")
    assert lib.bchip_write_bion(mine.encode(), C.byref(b)) == 0
    lib.bchip_free_bion(C.byref(b))
    theirs = str(tmp_path / "theirs.bion")
    code = ("import ctypes as C, sys\n" + _REF_RUNNER.split("lib = C.CDLL")[0] +
            "lib = C.CDLL(sys.argv[1])\nb = BD()\n"
            "assert lib._Z14bion_load_hostPKcP8BionData(sys.argv[2].encode(), C.byref(b)) == 0\n"
            "print(lib._Z9bion_savePKcPK8BionData(sys.argv[3].encode(), C.byref(b)))\n")
    r = subprocess.run([sys.executable, "-c", code, H.REF_SO, p, theirs], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stderr[-2000:]
    a = bytearray(open(mine, "rb").read())
    t = bytearray(open(theirs, "rb").read())
    assert len(a) == len(t)
    ts = bion.HEADER_DT.fields["timestamp"][1]   # offset 44 in the packed header
    a[ts:ts + 8] = t[ts:ts + 8] = b"\0" * 8
    # the reference's v9 connections carry unwritten reserved bytes: mask them
    hdr = np.frombuffer(bytes(a[:100]), bion.HEADER_DT)[0]
    off = 100 + int(hdr["num_neurons"]) * 24 + int(hdr["num_segments"]) * 3216
    bad = [i for i in range(len(a)) if a[i] != t[i] and not (off <= i < off + 16 * int(hdr["num_intra_connections"])
                                                             and (i - off) % 16 in (9, 10, 11))]
    assert not bad, ("differing byte offsets (header 100 B, neurons 24 B, segments 3216 B each from 100)",
                     off, bad[:32], [(a[i], t[i]) for i in bad[:8]])
