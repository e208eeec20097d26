"""bchip_adjoint_noalloc (the allocation-free variant of bchip_adjoint) through the C ABI: its
loss, gradients and somatic trace equal bchip_adjoint's bit for bit, on two calls in a row, the
second with a smaller input (fewer ticks, one cell); a call that needs more than its fixed
scratch array is refused with nothing run; pending stimuli and COMPAT coherence are refused."""
import ctypes as C

import numpy as np
import pytest

from bioneuroncore_amd import backend, bion
import harness as H

pytestmark = pytest.mark.gpu
DT = 0.025


def _cells(sizes=(200, 120)):
    return [bion.config5_cell(61 + k, n_seg=n) for k, n in enumerate(sizes)]


def _soma(cells):
    out, off = [], 0
    for c in cells:
        out.append(off + int(np.nonzero(c.segments["segment_type"] == 1)[0][0]))
        off += len(c.segments)
    return np.array(out, np.uint32)


def _sim(cells):
    sim = backend.Simulation(coherence=backend.COHERENCE_FAST)
    assert sim.load_batch([c.copy() for c in cells]) == 0
    sim.set_cable_mode(True)
    rng = np.random.default_rng(3)
    sim.set_coupling_scale(rng.uniform(0.5, 1.5, sum(len(c.connections) for c in cells)).astype(np.float32))
    for k, c in enumerate(cells):
        sim.set_stimuli(k, current=H.soma_clamp(c, 40.0 + 30.0 * k))
    sim.run_tick(DT)   # builds the stimulus lists (the no-allocation call refuses pending ones)
    return sim


def _target(ticks, ncells):
    t = np.arange(ticks, dtype=np.float32)[:, None]
    return (-58.0 + 0.15 * t - 4.0 * np.arange(ncells, dtype=np.float32)[None, :]).astype(np.float32)


def _run(cells, calls, noalloc):
    """The calls (ticks, checkpoint interval, cell indices) in a row on one simulation."""
    sim = _sim(cells)
    soma = _soma(cells)
    out = []
    for ticks, every, which in calls:
        sm = soma[list(which)]
        L, gd, gs, tr = sim.adjoint(DT, ticks, every, sm, _target(ticks, len(sm)), trace=True, noalloc=noalloc)
        out.append((L, gd, gs, tr, sim.bion_data().segments["V_m"].copy(), sim.tick_count()))
    sim.close()
    return out


def _same(a, b):
    assert a[0] == b[0]
    for x, y in zip(a[1:5], b[1:5]):
        np.testing.assert_array_equal(x, y)
    assert a[5] == b[5]


def test_noalloc_equals_adjoint_on_two_calls():
    """Two calls in a row, the second smaller (24 ticks, one cell): the second no-allocation call
    finds the first one's contents in the array and must clear what it uses."""
    cells = _cells()
    calls = [(48, 16, (0, 1)), (24, 8, (1,))]
    ref = _run(cells, calls, False)
    new = _run(cells, calls, True)
    for a, b in zip(ref, new):
        _same(a, b)
    assert np.abs(ref[0][1]).max() > 0 and np.abs(ref[0][2]).max() > 0
    assert np.abs(ref[1][1]).max() > 0


def test_noalloc_refusals():
    cells = _cells()
    soma = _soma(cells)
    sim = _sim(cells)
    t0 = sim.tick_count()
    # targets alone beyond the 64 MiB array: 3M ticks x 2 cells x 4 B = 24 MiB, checkpoints every
    # tick: 3M x (8 + 56 x 5) B per segment -> far beyond
    ticks = 3_000_000
    tgt = np.zeros((ticks, 2), np.float32)
    loss = C.c_double(0.0)
    lib = backend.load_library()
    rc = lib.bchip_adjoint_noalloc(sim.h, DT, ticks, 1, soma.ctypes.data, 2, tgt.ctypes.data, C.byref(loss),
                                   None, None, None)
    assert rc == backend.ERR_SCRATCH
    assert sim.tick_count() == t0   # nothing ran
    sim.set_stimuli(0, current=H.soma_clamp(cells[0], 10.0))   # pending upload
    tgt = _target(8, 2)
    rc = lib.bchip_adjoint_noalloc(sim.h, DT, 8, 4, soma.ctypes.data, 2, tgt.ctypes.data, C.byref(loss),
                                   None, None, None)
    assert rc == -1
    sim.run_tick(DT)
    sim.set_coherence(backend.COHERENCE_COMPAT)
    rc = lib.bchip_adjoint_noalloc(sim.h, DT, 8, 4, soma.ctypes.data, 2, tgt.ctypes.data, C.byref(loss),
                                   None, None, None)
    assert rc == -1
    sim.close()
