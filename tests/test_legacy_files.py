"""BION files of format versions 3-14 (and the v10+ record-size variants, and the files the loader
must refuse) give the records the reference's own loader gives them, checked against the
committed fixture tests/golden/legacy_loads.json (written by tests/golden/gen_legacy.py from
the compiled reference): return code, counts, header, and the SHA-256 of each record array with
the bytes the reference leaves unwritten zeroed. The files are rebuilt from their seeds
(tests/bion_legacy.py). Once through the host-only reader bchip_read_bion, and once through the
drop-in's own gpu_backend_load_bion on the device, read back with gpu_backend_bion_data."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

from bioneuroncore_amd import backend
import bion_legacy as L
import harness as H

FIX = os.path.join(H.ROOT, "tests", "golden", "legacy_loads.json")
SIZES = (24, 3216, 16, 1568, 500)


def _grab(b):
    def grab(p, n, sz):
        return np.frombuffer(bytes((C.c_uint8 * (n * sz)).from_address(p)), np.uint8) if n and p else \
            np.zeros(0, np.uint8)
    n = [b.num_neurons, b.num_segments, b.num_connections, b.num_synapses, b.num_nuclei]
    ptrs = [b.neurons, b.segments, b.connections, b.synapses, b.nuclei]
    return bytes(b.header), n, [grab(p, k, s) for p, k, s in zip(ptrs, n, SIZES)]


def _check(name, want, header, n, parts):
    assert list(n) == want["n"], (name, n, want["n"])
    assert header.hex() == want["header"], name
    got = [hashlib.sha256(np.where(m, a, 0).astype(np.uint8).tobytes()).hexdigest()
           for a, m in zip(parts, L.defined(want["version"], parts))]
    for part, g, w in zip(("neurons", "segments", "connections", "synapses", "nuclei"), got, want["sha256"]):
        assert g == w, (name, part)


def test_host_reader_against_fixture(tmp_path):
    fx = json.load(open(FIX))
    lib = backend.load_library()
    lib.bchip_read_bion.argtypes = [C.c_char_p, C.c_void_p]
    lib.bchip_free_bion.argtypes = [C.c_void_p]
    for name, (p, v) in L.cases(str(tmp_path)).items():
        b = backend.BionDataC()
        rc = lib.bchip_read_bion(p.encode(), C.byref(b))
        assert rc == fx[name]["rc"] == 0, (name, rc)
        _check(name, fx[name], *_grab(b))
        lib.bchip_free_bion(C.byref(b))
    for name, p in L.error_cases(str(tmp_path)).items():
        b = backend.BionDataC()
        assert lib.bchip_read_bion(p.encode(), C.byref(b)) == fx["error_" + name]["rc"], name


@pytest.mark.gpu
def test_drop_in_load_against_fixture(tmp_path):
    """gpu_backend_load_bion on the device (types outside the device set allowed: the files hold
    random records) returns the reference loader's code, and gpu_backend_bion_data its records."""
    fx = json.load(open(FIX))
    lib = backend.load_library()
    cases = L.cases(str(tmp_path))
    errs = L.error_cases(str(tmp_path))
    for name, path in [(k, p) for k, (p, _) in cases.items()] + [("error_" + k, p) for k, p in errs.items()]:
        want = fx[name]
        h = lib.gpu_backend_create()
        assert h
        try:
            lib.bchip_allow_unsupported(h, 1)
            rc = lib.gpu_backend_load_bion(h, path.encode())
            expect = want["rc"] if want["rc"] != 0 else (-1 if want["n"][1] == 0 else 0)
            assert rc == expect, (name, rc, expect)
            if rc == 0:
                bd = C.cast(lib.gpu_backend_bion_data(h), C.POINTER(backend.BionDataC)).contents
                _check(name, want, *_grab(bd))
            print(f"{name}: rc {rc}", flush=True)
        finally:
            lib.gpu_backend_destroy(h)
