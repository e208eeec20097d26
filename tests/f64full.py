"""The oracle's FP64 build (oracle/liboracle_f64.so: the same restatement with every float
quantity in double, record storage as the ABI's; see oracle/f64.h) as the FP64 reference of the
per-field error reports (test infrastructure)."""
import ctypes as C
import os

import numpy as np

import harness as H

P = C.c_void_p
F64_SO = os.path.join(H.ROOT, "oracle", "liboracle_f64.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(F64_SO)
        L.oracle_create.restype = P
        L.oracle_destroy.argtypes = [P]
        # dt is a double in this build (every float parameter of the restatement is)
        L.oracle_run_tick.argtypes = [P, P, C.c_uint32, P, C.c_uint32, P, C.c_uint32, P, C.c_double] + [P] * 10
        L.oracle_spike_counts.argtypes = [P, C.POINTER(C.c_uint32), C.c_uint32]
        _lib = L
    return _lib


def run(net, ticks, dt, stim_fn=None, after=None):
    """Steps `net` in place for `ticks` ticks with the FP64 build; returns its spike counts.
    `after(t, net)` is called after each tick."""
    L = lib()
    ctx = L.oracle_create()
    tick = np.zeros(1, np.uint64)
    n = net
    for t in range(ticks):
        st = stim_fn(t).args() if stim_fn else [None] * 10
        L.oracle_run_tick(ctx, H._ptr(n.segments), len(n.segments), H._ptr(n.connections), len(n.connections),
                          H._ptr(n.synapses), len(n.synapses), tick.ctypes.data_as(P), dt, *st)
        if after:
            after(t, n)
    sp = np.zeros(len(n.segments), np.uint32)
    L.oracle_spike_counts(ctx, sp.ctypes.data_as(C.POINTER(C.c_uint32)), len(sp))
    L.oracle_destroy(ctx)
    return sp


ION = ("Na", "K", "Ca", "Cl", "Mg", "A", "H")


class ReleaseLog:
    """Release ticks of every chemical synapse: a release stores round(20 ms / dt) into the
    synapse's 8-bit refractory counter (so 800 wraps to 32 at dt = 0.025 ms) and the end of the
    same tick's step 3 decrements it once, so a tick that released leaves that value - 1; the
    counter only counts down from there."""

    def __init__(self, n_syn, dt, bits=8):
        self.full = (int(20.0 / dt + 0.5) % (1 << bits)) - 1
        self.ticks = [[] for _ in range(n_syn)]

    def __call__(self, t, synapses):
        for k in np.nonzero(synapses["refractory"] == self.full)[0]:
            self.ticks[k].append(t)

    def same_segments(self, other, synapses, n_seg):
        """Mask of segments every incoming synapse of which released on the same ticks in both logs."""
        ok = np.ones(n_seg, bool)
        for k in range(len(synapses)):
            if self.ticks[k] != other.ticks[k]:
                ok[synapses["post_segment"][k]] = False
        return ok


def field_errors(seg, ref, floor=1e-12, nt_rows=None):
    """Largest and median relative error of each field of `seg` against the FP64 records `ref`:
    V_m, the 7 intra- and extracellular ions, ATP and the transmitter levels (species with any
    nonzero FP64 value, over the segments `nt_rows` selects, all by default)."""
    def rel(x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        r = np.abs(x - y) / np.maximum(np.abs(y), floor)
        return float(r.max()), float(np.median(r))
    out = {"V_m": rel(seg["V_m"], ref["V_m"]), "ATP": rel(seg["ATP_level"], ref["ATP_level"])}
    for k, nm in enumerate(ION):
        out[f"intra {nm}"] = rel(seg["intra_ions"][:, k], ref["intra_ions"][:, k])
        out[f"extra {nm}"] = rel(seg["extra_ions"][:, k], ref["extra_ions"][:, k])
    nt, rnt = seg["neurotransmitter"], ref["neurotransmitter"]
    if nt_rows is not None:
        nt, rnt = nt[nt_rows], rnt[nt_rows]
    live = np.nonzero((rnt != 0).any(axis=0))[0]
    if len(live):
        out["transmitters"] = rel(nt[:, live], rnt[:, live])
    return out
