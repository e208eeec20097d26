"""Python host mirror of libbiocore_hip_backend.so's C ABI (ctypes; plain pointers and sizes).

Mirrors how BioNeuronCore's callers drive a GPU backend: create a sim on a device, load a
BION file (or an in-memory network), upload stimuli, run ticks, read the host AoS view
(reference src/gpu_backend_cuda.cu:316-390, tests/test_cpu_gpu_parity_cuda.cpp:45-170).
There is no fallback: if the library is missing or no device is present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import bion

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libbiocore_hip_backend.so")

COHERENCE_COMPAT = 0
COHERENCE_FAST = 1
ERR_ISLAND_TOO_LARGE = -20
ERR_UNSUPPORTED = -21
ERR_SCRATCH = -22

STIM_DT = np.dtype([
    ("current_stim_count", "<u4"), ("vclamp_count", "<u4"), ("ramp_count", "<u4"),
    ("ion_inj_count", "<u4"), ("lig_inj_count", "<u4"), ("current_tick", "<u8"),
    ("current_stim", bion.CURRENT_STIM_DT, 256), ("vclamps", bion.VCLAMP_DT, 256),
    ("ramps", bion.RAMP_DT, 256), ("ion_inj", bion.ION_INJ_DT, 256), ("lig_inj", bion.LIG_INJ_DT, 256)])
DEVICE_INFO_DT = np.dtype([("name", "S128"), ("vendor_id", "<u4"), ("is_cpu", "u1"),
                           ("global_mem_bytes", "<u8"), ("max_compute_units", "<u4"),
                           ("device_index", "<i4"), ("backend_type", "S16")])
assert DEVICE_INFO_DT.itemsize == 165

# Symbols a dlopen caller of the reference GPU backend resolves (tests/smoke_cuda.sh:24-28).
ABI_SYMBOLS = ("gpu_backend_create", "gpu_backend_create_on_device", "gpu_backend_destroy",
               "gpu_backend_load_bion", "gpu_backend_run_tick", "gpu_backend_bion_data",
               "gpu_backend_upload_stimuli", "gpu_backend_device_count", "gpu_backend_get_device_info",
               "gpu_backend_ibuf_offset", "gpu_backend_vm_offset", "gpu_backend_seg_size")
PLUGIN_SYMBOLS = ("register_modifier_current", "register_modulator_release",
                  "_Z22register_user_modifierhPFdP8ModifierP7SegmentfE",
                  "_Z23register_user_modulatorhPFfP8ModifierPffE")
EXT_SYMBOLS = ("bchip_load_bion_data", "bchip_save_bion", "bchip_set_coherence", "bchip_push_host_state",
               "bchip_run_ticks", "bchip_tick_count", "bchip_spike_counts", "bchip_delay_slots", "bchip_stats",
               "bchip_set_phase_timing", "bchip_stream", "bchip_plugin_load", "bchip_registered_modifier",
               "bchip_modifier_supported", "bchip_load_batch", "bchip_network_count", "bchip_set_stimuli",
               "bchip_batch_read", "bchip_get_schedule", "bchip_set_schedule", "bchip_autotune",
               "bchip_set_graphs", "bchip_set_jacobi", "bchip_set_cable_mode", "bchip_set_coupling_scale",
               "bchip_adjoint", "bchip_set_adjoint_memory", "bchip_read_bion", "bchip_free_bion",
               "bchip_write_bion", "bchip_set_stimulus_convention", "bchip_set_protein_mapping")


class BionDataC(C.Structure):
    _fields_ = [("header", C.c_uint8 * 100), ("neurons", C.c_void_p), ("segments", C.c_void_p),
                ("connections", C.c_void_p), ("synapses", C.c_void_p), ("nuclei", C.c_void_p),
                ("num_neurons", C.c_uint32), ("num_segments", C.c_uint32), ("num_connections", C.c_uint32),
                ("num_synapses", C.c_uint32), ("num_nuclei", C.c_uint32)]


class BchipStats(C.Structure):
    _fields_ = [("num_segments", C.c_uint32), ("num_connections", C.c_uint32), ("num_synapses", C.c_uint32),
                ("num_tiles", C.c_uint32), ("modifier_slots", C.c_uint32), ("receptor_slots", C.c_uint32),
                ("lds_bytes_per_tile", C.c_uint32), ("_pad", C.c_uint32),
                ("unsupported_modifiers", C.c_uint64), ("unsupported_synapses", C.c_uint64),
                ("order_hazard_synapses", C.c_uint64), ("last_phase_ms", C.c_float * 6),
                ("ordered_events", C.c_uint32), ("ordered_segments", C.c_uint32),
                ("big_islands", C.c_uint32), ("_pad2", C.c_uint32)]


class BchipSchedule(C.Structure):
    _fields_ = [("modifiers_block", C.c_uint32), ("diffusion_tiles", C.c_uint32), ("pre_grid", C.c_uint32),
                ("plast_block", C.c_uint32), ("rrp_block", C.c_uint32), ("step1_chunks", C.c_uint32)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


assert C.sizeof(BionDataC) == 168

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the backend library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libbiocore_hip_backend.so not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    P = C.c_void_p
    lib.gpu_backend_create.restype = P
    lib.gpu_backend_create_on_device.restype = P
    lib.gpu_backend_create_on_device.argtypes = [C.c_int]
    lib.gpu_backend_destroy.argtypes = [P]
    lib.gpu_backend_load_bion.argtypes = [P, C.c_char_p]
    lib.gpu_backend_run_tick.argtypes = [P, C.c_float]
    lib.gpu_backend_bion_data.restype = P
    lib.gpu_backend_bion_data.argtypes = [P]
    lib.gpu_backend_upload_stimuli.argtypes = [P, P]
    lib.gpu_backend_get_device_info.argtypes = [C.c_int, P]
    lib.bchip_load_bion_data.argtypes = [P, C.POINTER(BionDataC)]
    lib.bchip_save_bion.argtypes = [P, C.c_char_p]
    lib.bchip_set_coherence.argtypes = [P, C.c_int]
    lib.bchip_push_host_state.argtypes = [P]
    lib.bchip_run_ticks.argtypes = [P, C.c_float, C.c_uint32]
    lib.bchip_tick_count.restype = C.c_uint64
    lib.bchip_tick_count.argtypes = [P]
    lib.bchip_spike_counts.argtypes = [P, P, C.c_uint32]
    lib.bchip_delay_slots.argtypes = [P, P, C.c_uint64]
    lib.bchip_stats.argtypes = [P, C.POINTER(BchipStats)]
    lib.bchip_set_phase_timing.argtypes = [P, C.c_int]
    lib.bchip_stream.restype = P
    lib.bchip_stream.argtypes = [P]
    lib.bchip_plugin_load.argtypes = [C.c_char_p]
    lib.bchip_registered_modifier.argtypes = [C.c_uint8]
    lib.bchip_modifier_supported.argtypes = [C.c_uint8]
    lib.bchip_load_batch.argtypes = [P, C.POINTER(C.POINTER(BionDataC)), C.c_uint32]
    lib.bchip_network_count.restype = C.c_uint32
    lib.bchip_network_count.argtypes = [P]
    lib.bchip_set_stimuli.argtypes = [P, C.c_uint32] + [P, C.c_uint32] * 5
    lib.bchip_batch_read.argtypes = [P, C.c_uint32, P, P, P]
    lib.bchip_get_schedule.argtypes = [P, C.POINTER(BchipSchedule)]
    lib.bchip_set_schedule.argtypes = [P, C.POINTER(BchipSchedule)]
    lib.bchip_autotune.argtypes = [P, C.c_float, C.c_uint32]
    lib.bchip_set_graphs.argtypes = [P, C.c_int]
    lib.bchip_set_jacobi.argtypes = [P, C.c_double, C.c_int]
    lib.bchip_set_cable_mode.argtypes = [P, C.c_int]
    lib.bchip_set_coupling_scale.argtypes = [P, P, C.c_uint32]
    lib.bchip_adjoint.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32, P, C.c_uint32, P, P, P, P, P]
    lib.bchip_adjoint_noalloc.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32, P, C.c_uint32, P, P, P, P, P]
    lib.bchip_set_adjoint_memory.argtypes = [P, C.c_uint64]
    lib.bchip_set_stimulus_convention.argtypes = [P, C.c_int]
    lib.bchip_set_protein_mapping.argtypes = [P, C.c_int]
    lib.bchip_allow_unsupported.argtypes = [P, C.c_int]
    lib.bchip_eval_math.argtypes = [C.c_int, P, P, P, C.c_uint32]
    _lib = lib
    return lib


def device_count() -> int:
    return int(load_library().gpu_backend_device_count())


def device_info(index: int = 0) -> np.ndarray | None:
    out = np.zeros(1, DEVICE_INFO_DT)
    rc = load_library().gpu_backend_get_device_info(index, out.ctypes.data)
    return out[0] if rc == 0 else None


def make_stimuli(current=None, vclamp=None, ramp=None, ion=None, lig=None) -> np.ndarray:
    st = np.zeros(1, STIM_DT)
    for name, cnt, src in (("current_stim", "current_stim_count", current), ("vclamps", "vclamp_count", vclamp),
                           ("ramps", "ramp_count", ramp), ("ion_inj", "ion_inj_count", ion),
                           ("lig_inj", "lig_inj_count", lig)):
        if src is not None and len(src):
            if len(src) > 256:
                raise ValueError("at most 256 stimuli per type")
            st[0][name][:len(src)] = src
            st[0][cnt] = len(src)
    return st


class Simulation:
    """One network on one GPU (the backend's `sim` handle)."""

    def __init__(self, device: int = 0, coherence: int = COHERENCE_FAST):
        self.lib = load_library()
        self.h = self.lib.gpu_backend_create_on_device(device)
        if not self.h:
            raise RuntimeError("gpu_backend_create_on_device returned NULL (no HIP device)")
        self.lib.bchip_set_coherence(self.h, coherence)
        self._keep = None

    def close(self):
        if self.h:
            self.lib.gpu_backend_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_bion(self, path: str) -> int:
        return int(self.lib.gpu_backend_load_bion(self.h, path.encode()))

    def allow_unsupported(self, on: bool = True) -> None:
        """Load networks that need work the device does not do (off: such loads return
        ERR_UNSUPPORTED)."""
        self.lib.bchip_allow_unsupported(self.h, 1 if on else 0)

    def load_network(self, net: bion.Network, nuclei: np.ndarray | None = None) -> int:
        hdr = bion.make_header(net)
        b = BionDataC()
        C.memmove(b.header, hdr.ctypes.data, 100)
        arrs = [np.ascontiguousarray(a) for a in (net.neurons, net.segments, net.connections, net.synapses)]
        b.neurons, b.segments, b.connections, b.synapses = [a.ctypes.data if len(a) else None for a in arrs]
        b.num_neurons, b.num_segments, b.num_connections, b.num_synapses = [len(a) for a in arrs]
        if nuclei is not None and len(nuclei):
            nuc = np.ascontiguousarray(nuclei, dtype=bion.NUCLEUS_DT)
            b.nuclei, b.num_nuclei = nuc.ctypes.data, len(nuc)
        rc = int(self.lib.bchip_load_bion_data(self.h, C.byref(b)))
        return rc

    def set_protein_mapping(self, on: bool) -> None:
        """Nucleus proteins set the release parameters of chemical synapses (see the C header)."""
        if self.lib.bchip_set_protein_mapping(self.h, 1 if on else 0) != 0:
            raise RuntimeError("bchip_set_protein_mapping failed")

    def load_batch(self, nets: list[bion.Network]) -> int:
        """Load independent networks as one batch (network k's results equal a separate load)."""
        keep, ptrs = [], []
        for net in nets:
            hdr = bion.make_header(net)
            b = BionDataC()
            C.memmove(b.header, hdr.ctypes.data, 100)
            arrs = [np.ascontiguousarray(a) for a in (net.neurons, net.segments, net.connections, net.synapses)]
            b.neurons, b.segments, b.connections, b.synapses = [a.ctypes.data if len(a) else None for a in arrs]
            b.num_neurons, b.num_segments, b.num_connections, b.num_synapses = [len(a) for a in arrs]
            keep.append((b, arrs))
            ptrs.append(C.pointer(b))
        arr = (C.POINTER(BionDataC) * len(ptrs))(*ptrs)
        self._batch_sizes = [(len(n.segments), len(n.synapses)) for n in nets]
        return int(self.lib.bchip_load_batch(self.h, arr, len(ptrs)))

    def network_count(self) -> int:
        return int(self.lib.bchip_network_count(self.h))

    def set_stimuli(self, net: int, current=None, ramp=None, vclamp=None, ion=None, lig=None) -> None:
        """Unbounded stimuli of one network (structured arrays of the STIM_DT dtypes)."""
        args, keep = [], []
        for arr, dt in ((current, bion.CURRENT_STIM_DT), (ramp, bion.RAMP_DT), (vclamp, bion.VCLAMP_DT),
                        (ion, bion.ION_INJ_DT), (lig, bion.LIG_INJ_DT)):
            a = np.ascontiguousarray(arr if arr is not None else np.zeros(0, dt), dtype=dt)
            keep.append(a)   # alive until the call returns (the library copies)
            args += [a.ctypes.data if len(a) else None, len(a)]
        if self.lib.bchip_set_stimuli(self.h, net, *args) != 0:
            raise RuntimeError("bchip_set_stimuli failed")

    def batch_read(self, net: int):
        """(segments, synapses, spikes) of network `net`, in its own indices."""
        ns, ny = self._batch_sizes[net]
        segs = np.zeros(ns, bion.SEGMENT_DT)
        syns = np.zeros(ny, bion.SYNAPSE_DT)
        spk = np.zeros(ns, np.uint32)
        if self.lib.bchip_batch_read(self.h, net, segs.ctypes.data, syns.ctypes.data if ny else None,
                                     spk.ctypes.data) != 0:
            raise RuntimeError("bchip_batch_read failed")
        return segs, syns, spk

    def schedule(self) -> BchipSchedule:
        s = BchipSchedule()
        self.lib.bchip_get_schedule(self.h, C.byref(s))
        return s

    def set_schedule(self, **fields) -> None:
        s = self.schedule()
        for f, v in fields.items():
            setattr(s, f, v)
        if self.lib.bchip_set_schedule(self.h, C.byref(s)) != 0:
            raise ValueError(f"bad schedule {fields}")

    def autotune(self, dt: float, ticks_per_candidate: int = 3) -> int:
        r = int(self.lib.bchip_autotune(self.h, dt, ticks_per_candidate))
        if r < 0:
            raise RuntimeError("bchip_autotune failed")
        return r

    def set_jacobi(self, tol: float, max_sweeps: int = 200) -> None:
        """Residual-terminated Jacobi (tol <= 0: the reference's fixed 50 sweeps)."""
        if self.lib.bchip_set_jacobi(self.h, tol, max_sweeps) != 0:
            raise ValueError("bad Jacobi settings")

    def set_cable_mode(self, on: bool) -> None:
        if self.lib.bchip_set_cable_mode(self.h, 1 if on else 0) != 0:
            raise ValueError("cable mode needs a network without synapses")

    def set_coupling_scale(self, scale) -> None:
        """Per-connection axial scale (None: 1, the reference)."""
        if scale is None:
            rc = self.lib.bchip_set_coupling_scale(self.h, None, 0)
        else:
            s = np.ascontiguousarray(scale, dtype=np.float32)
            rc = self.lib.bchip_set_coupling_scale(self.h, s.ctypes.data, len(s))
        if rc != 0:
            raise ValueError("bad coupling scale")

    def adjoint(self, dt: float, ticks: int, checkpoint_every: int, soma, target, grads: bool = True,
                trace: bool = False, noalloc: bool = False):
        """Cable-mode loss and gradients: (loss, d loss/d density [n, 32], d loss/d coupling scale [nc]),
        plus the somatic V_m trace (ticks, cells) when trace=True.
        soma: host segment index per cell; target: (ticks, cells) V_m after each tick."""
        soma = np.ascontiguousarray(soma, dtype=np.uint32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        assert target.shape == (ticks, len(soma))
        loss = C.c_double(0.0)
        b = self.bion_data()
        n, nc = len(b.segments), len(b.connections)
        gd = np.zeros((n, 32)) if grads else None
        gs = np.zeros(max(nc, 1)) if grads else None
        tr = np.zeros((ticks, len(soma)), np.float32) if trace else None
        fn = self.lib.bchip_adjoint_noalloc if noalloc else self.lib.bchip_adjoint
        rc = fn(self.h, dt, ticks, checkpoint_every, soma.ctypes.data, len(soma),
                target.ctypes.data, C.byref(loss), gd.ctypes.data if grads else None,
                gs.ctypes.data if grads else None, tr.ctypes.data if trace else None)
        if rc != 0:
            raise RuntimeError(f"bchip_adjoint failed ({rc})", rc)
        out = (loss.value, gd, (gs[:nc] if grads else None))
        return out + (tr,) if trace else out

    def set_stimulus_convention(self, convention: int) -> None:
        """0: stimulus windows matched at the tick being run (CPU/CUDA); 1: one tick later (SYCL)."""
        if self.lib.bchip_set_stimulus_convention(self.h, convention) != 0:
            raise ValueError("convention must be 0 or 1")

    def set_adjoint_memory(self, max_bytes: int) -> None:
        """Cap on the adjoint's recorded-iterate memory (0: the reverse pass recomputes substeps)."""
        self.lib.bchip_set_adjoint_memory(self.h, max_bytes)

    def set_graphs(self, on: bool) -> None:
        self.lib.bchip_set_graphs(self.h, 1 if on else 0)

    def set_coherence(self, mode: int) -> None:
        if self.lib.bchip_set_coherence(self.h, mode) != 0:
            raise RuntimeError("bad coherence mode")

    def upload_stimuli(self, stim: np.ndarray) -> None:
        self.lib.gpu_backend_upload_stimuli(self.h, stim.ctypes.data)

    def run_tick(self, dt: float) -> None:
        self.lib.gpu_backend_run_tick(self.h, dt)

    def run_ticks(self, dt: float, count: int) -> None:
        if self.lib.bchip_run_ticks(self.h, dt, count) != 0:
            raise RuntimeError("bchip_run_ticks failed")

    def tick_count(self) -> int:
        return int(self.lib.bchip_tick_count(self.h))

    def bion_data(self) -> bion.Network:
        """Zero-copy numpy views of the sim-owned AoS (valid until the next load/destroy)."""
        p = self.lib.gpu_backend_bion_data(self.h)
        if not p:
            raise RuntimeError("no network loaded")
        b = BionDataC.from_address(p)

        def view(ptr, n, dt):
            if n == 0 or not ptr:
                return np.zeros(0, dt)
            return np.frombuffer((C.c_uint8 * (n * dt.itemsize)).from_address(ptr), dt)
        return bion.Network(view(b.neurons, b.num_neurons, bion.NEURON_DT),
                            view(b.segments, b.num_segments, bion.SEGMENT_DT),
                            view(b.connections, b.num_connections, bion.CONN_DT),
                            view(b.synapses, b.num_synapses, bion.SYNAPSE_DT))

    def push_host_state(self) -> None:
        if self.lib.bchip_push_host_state(self.h) != 0:
            raise RuntimeError("push failed")

    def spikes(self, n: int) -> np.ndarray:
        out = np.zeros(n, np.uint32)
        if self.lib.bchip_spike_counts(self.h, out.ctypes.data, n) != 0:
            raise RuntimeError("spike counts failed")
        return out

    def delay_slots(self, num_synapses: int) -> np.ndarray:
        out = np.zeros(num_synapses * 64, np.uint8)
        if num_synapses and self.lib.bchip_delay_slots(self.h, out.ctypes.data, out.size) != 0:
            raise RuntimeError("delay slots failed")
        return out

    def stats(self) -> BchipStats:
        s = BchipStats()
        self.lib.bchip_stats(self.h, C.byref(s))
        return s

    def set_phase_timing(self, on: bool) -> None:
        self.lib.bchip_set_phase_timing(self.h, 1 if on else 0)

    def save_bion(self, path: str) -> int:
        return int(self.lib.bchip_save_bion(self.h, path.encode()))
