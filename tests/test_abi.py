"""The C-ABI library without a GPU: it loads, exports every function include/*.h declares
(the reference backend's 11 dlsym'd symbols first, tests/smoke_cuda.sh:24-28), reports the
reference's ABI offsets, and its NULL-handle paths are safe."""
import ctypes as C
import os
import re
import subprocess

import pytest

from bioneuroncore_amd import backend

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for h in ("biocore_hip_backend.h",):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"//.*", "", txt)
        txt = re.sub(r"^\s*typedef[^;]*;", "", txt, flags=re.M)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", txt, re.M):
            if m.group(1) not in ("typedef",):
                names.append(m.group(1))
    return names


def _exports():
    out = subprocess.run(["nm", "-D", "--defined-only", backend.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_symbol():
    exp = _exports()
    declared = _declared_functions()
    assert len(declared) >= 28
    plain = [d for d in declared if d not in ("register_user_modifier", "register_user_modulator")]
    missing = [d for d in plain if d not in exp]
    assert not missing, missing
    for s in backend.ABI_SYMBOLS + backend.PLUGIN_SYMBOLS + backend.EXT_SYMBOLS:
        assert s in exp, s


def test_abi_offsets_match_reference_layout():
    lib = backend.load_library()
    assert lib.gpu_backend_ibuf_offset() == 3128
    assert lib.gpu_backend_vm_offset() == 40
    assert lib.gpu_backend_seg_size() == 3216


def test_null_handles_are_safe():
    lib = backend.load_library()
    lib.gpu_backend_destroy(None)
    lib.gpu_backend_run_tick(None, C.c_float(0.025))
    assert lib.gpu_backend_bion_data(None) is None
    lib.gpu_backend_upload_stimuli(None, None)
    assert lib.gpu_backend_load_bion(None, b"x.bion") == -1
    assert lib.gpu_backend_get_device_info(0, None) == -1
    assert lib.bchip_set_coherence(None, 0) == -1


def test_no_device_means_null_sim(monkeypatch):
    lib = backend.load_library()
    if lib.gpu_backend_device_count() > 0:
        pytest.skip("a HIP device is present")
    assert lib.gpu_backend_create() is None
    assert lib.gpu_backend_get_device_info(0, C.create_string_buffer(165)) == -1
    with pytest.raises(RuntimeError):
        backend.Simulation()


def test_supported_modifier_set():
    lib = backend.load_library()
    from bioneuroncore_amd import bion
    for t in range(256):
        assert bool(lib.bchip_modifier_supported(t)) == (t in bion.SUPPORTED_MODIFIERS), t
